// zf_membrane.hip -- RBC membrane forces (coarse-grained spectrin network).
//
// Replaces, for every live cell, ZebraflowWT.c:
//   CGSM_AREA triangle geometry + node normals   :5757-5851, :6089-6093
//   anti-sliver terms                             :6042-6086
//   cell area / centroid                          :6717-6796
//   CGSM_LOCAL_FORCES (spring, bending, areas)    :6800-7384
//   CGSM_VOLUME_CALC                              :7512-7535
// and, on the host side of this file, the set-up of the shared mesh:
//   topology CGSM_ZFISHRBC_DATA_GEN  :1632-1901
//   equilibrium parameters           :1930-2223, :2338-2346, :2387-2460
//
// Kernels (all FP64 arithmetic; the membrane is a small fraction of a step):
//   k_tri_geom     one lane per (cell, triangle): area, outward normal, centroid
//   k_cell_reduce  one workgroup per cell: FP64 tree sums of area, volume,
//                  centroid, then the global area and volume force scalars
//   k_node_force   one lane per (cell, node): gathers the node's normal, its
//                  anti-sliver terms and every rhombus (edge) it belongs to in
//                  edge order, so F_Total is summed in the reference's order
//                  with no atomics and no per-edge scratch.
#include "zf_membrane_dev.h"

#include <algorithm>
#include <cmath>
#include <map>

namespace zf {
namespace {

using namespace memb;

// Triangle geometry, ZebraflowWT.c:5800-5831: area, outward normal, centroid.
__global__ void k_tri_geom(const double *__restrict__ X, size_t NT, int nn, int nt, int ncell,
                           const int32_t *__restrict__ tri, const int32_t *__restrict__ status,
                           double L, double *__restrict__ tg, size_t TT) {
  const int idx = xcd_range_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (idx >= ncell * nt) return;
  const int c = idx / nt, t = idx % nt;
  if (status[c] <= kStatusDead) return;
  const size_t base = (size_t)c * nn;
  const TriG<double> g = tri_geom(ld3(X, NT, base + tri[3 * t]), ld3(X, NT, base + tri[3 * t + 1]),
                                  ld3(X, NT, base + tri[3 * t + 2]), L);
  const size_t q = (size_t)c * nt + t;
  tg[q] = g.a;
  tg[TT + q] = g.m.x;
  tg[2 * TT + q] = g.m.y;
  tg[3 * TT + q] = g.m.z;
  tg[4 * TT + q] = g.C.x;
  tg[5 * TT + q] = g.C.y;
  tg[6 * TT + q] = g.C.z;
}

// Per-cell FP64 reductions (ZebraflowWT.c:6729-6791, :7517-7524).
__global__ __launch_bounds__(256) void k_cell_reduce(const double *__restrict__ X, size_t NT,
                                                     int nn, int nt, const double *__restrict__ tg,
                                                     size_t TT, const int32_t *__restrict__ status,
                                                     MembParams mp, double *__restrict__ cell) {
  __shared__ double sh[5][256];
  const int c = blockIdx.x;
  if (status[c] <= kStatusDead) return;
  double s[5] = {0, 0, 0, 0, 0};
  const size_t nb = (size_t)c * nn, tb = (size_t)c * nt;
  for (int n = threadIdx.x; n < nn; n += 256) {
    s[0] += X[nb + n];
    s[1] += X[NT + nb + n];
    s[2] += X[2 * NT + nb + n];
  }
  for (int t = threadIdx.x; t < nt; t += 256) {
    const size_t q = tb + t;
    const double a = tg[q];
    s[3] += a;
    s[4] += a * (tg[4 * TT + q] * tg[TT + q] + tg[5 * TT + q] * tg[2 * TT + q] +
                 tg[6 * TT + q] * tg[3 * TT + q]);
  }
  for (int k = 0; k < 5; k++) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 5; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double area = sh[3][0], vol = sh[4][0] * mp.L / 3.0;
    cell[8 * c + 0] = area;
    cell[8 * c + 1] = vol;
    cell[8 * c + 2] = sh[0][0] / nn;
    cell[8 * c + 3] = sh[1][0] / nn;
    cell[8 * c + 4] = sh[2][0] / nn;
    cell[8 * c + 5] = -mp.ka * mp.relax * (area - mp.area0) / mp.area0;  // F_AREA_GLOBAL
    cell[8 * c + 6] = -mp.kv * (vol - mp.vol0) / mp.vol0;               // F_VOLUME_SCALAR
  }
}

// One lane per (cell, node): the node's normal, its anti-sliver terms and
// every rhombus (edge) it belongs to in edge order, so F_Total is summed in
// the reference's order with no atomics and no per-edge scratch.  T: the
// arithmetic of the forces (float: relative to the node, ZF_FLAG_F32_CELLS);
// the triangle geometry and the cell scalars come in FP64 either way.
template <class T>
__global__ __launch_bounds__(128) void k_node_force(const double *__restrict__ X, size_t NT, int ncell, MeshTables mt,
                             const double *__restrict__ tg, size_t TT, const double *__restrict__ cell,
                             const int32_t *__restrict__ status, MembParams mp,
                             double *__restrict__ F, double *__restrict__ NN,
                             int32_t *__restrict__ fail) {
  const int nn = mt.nn;
  const int idx = xcd_range_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (idx >= ncell * nn) return;
  const int c = idx / nn, n = idx % nn;
  const size_t q = (size_t)c * nn + n;
  if (status[c] <= kStatusDead) {
    F[q] = F[NT + q] = F[2 * NT + q] = 0;
    return;
  }
  const V3<T> cc = liftp<T>(V3<double>{cell[8 * c + 2], cell[8 * c + 3], cell[8 * c + 4]},
                           ld3(X, NT, q));
  const NodeForce<T> r = node_force<T>(n, (size_t)c * nn, (size_t)c * mt.nt, X, NT, tg, TT, mt, cc,
                                       T(cell[8 * c + 5]), T(cell[8 * c + 6]), mp);
  NN[q] = r.nrm.x;
  NN[NT + q] = r.nrm.y;
  NN[2 * NT + q] = r.nrm.z;
  F[q] = (double)r.f.x;
  F[NT + q] = (double)r.f.y;
  F[2 * NT + q] = (double)r.f.z;
  if (r.fails) atomicAdd(&fail[c], r.fails);
}

__global__ void k_zero_i32(int32_t *a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// (cell, node, axis) host order <-> SoA device order
__global__ void k_cells_import(const double *__restrict__ h, int ncell, int nn, size_t NT,
                               double *__restrict__ X, double *__restrict__ V) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell * nn) return;
  for (int d = 0; d < 3; d++) {
    X[d * NT + idx] = h[3 * (size_t)idx + d];
    V[d * NT + idx] = 0.0;
  }
}

__global__ void k_soa_export(const double *__restrict__ A, int n, size_t NT, double *__restrict__ h) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  for (int d = 0; d < 3; d++) h[3 * (size_t)idx + d] = A[d * NT + idx];
}

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// ---- host: topology and equilibrium parameters ------------------------------

// Edge list by first appearance of the pairs (1,2),(2,3),(3,1) over the
// triangles; n3/t1 from the first triangle holding the edge, n4/t2 from the
// second (ZebraflowWT.c:1655-1805).
void build_topology(int nn, int nt, const int32_t *tri, std::vector<int32_t> &edge,
                    std::vector<int32_t> &nbr_node, std::vector<int32_t> &nbr_tri) {
  std::map<std::pair<int, int>, int> id;
  edge.clear();
  for (int t = 0; t < nt; t++)
    for (int s = 0; s < 3; s++) {
      const int p = tri[3 * t + s], q = tri[3 * t + (s + 1) % 3];
      const auto key = std::make_pair(std::min(p, q), std::max(p, q));
      if (id.count(key)) continue;
      id[key] = (int)edge.size() / 6;
      edge.insert(edge.end(), {p, q, 0, 0, 0, 0});
    }
  for (int t = 0; t < nt; t++)
    for (int s = 0; s < 3; s++) {
      const int p = tri[3 * t + s], q = tri[3 * t + (s + 1) % 3], w = tri[3 * t + (s + 2) % 3];
      int32_t *r = &edge[6 * id[std::make_pair(std::min(p, q), std::max(p, q))]];
      if (r[2] == 0) { r[2] = w; r[4] = t + 1; }
      else if (r[3] == 0) { r[3] = w; r[5] = t + 1; }
    }
  const int ne = (int)edge.size() / 6;
  nbr_node.assign((size_t)nn * 10, 0);
  nbr_tri.assign((size_t)nn * 10, 0);
  for (int e = 0; e < ne; e++)
    for (int s = 0; s < 2; s++) {
      const int n = edge[6 * e + s], other = edge[6 * e + 1 - s];
      int32_t *row = &nbr_node[(size_t)(n - 1) * 10];
      row[0]++;
      if (row[0] < 10) row[row[0]] = other;
    }
  for (int t = 0; t < nt; t++)
    for (int s = 0; s < 3; s++) {
      int32_t *row = &nbr_tri[(size_t)(tri[3 * t + s] - 1) * 10];
      row[0]++;
      if (row[0] < 10) row[row[0]] = t + 1;
    }
}

// WLC-POW fixed point for Plength, kp (ZebraflowWT.c:1942-1990) and the
// rest dihedral angle (:2081-2210).
void build_equilibrium(const zf_membrane_config &m, const double *x, const std::vector<int32_t> &edge,
                       std::vector<double> &eq) {
  const int ne = (int)edge.size() / 6;
  const double L = m.len_scale, x0 = m.x0;
  const double beta = 0.5 * x0 / std::pow(1 - x0, 3) - 0.25 / ((1 - x0) * (1 - x0)) + 0.25;
  eq.assign((size_t)6 * ne, 0.0);
  auto P = [&](int n, int d) { return x[3 * (n - 1) + d]; };
  for (int e = 0; e < ne; e++) {
    const int n1 = edge[6 * e], n2 = edge[6 * e + 1], n3 = edge[6 * e + 2], n4 = edge[6 * e + 3];
    double l2 = 0;
    for (int d = 0; d < 3; d++) l2 += (P(n1, d) - P(n2, d)) * (P(n1, d) - P(n2, d));
    const double l = std::sqrt(l2);
    const double alpha = std::sqrt(3.0) * 0.25 * m.kT / (l * L);
    const double eta = std::sqrt(3.0) * 0.25 * 3.0 / (l * l * l * L * L * L);
    double plen = 1e-9 * (L / 1e-6), kp = 0;
    for (int pass = 0; pass < m.passes; pass++) {
      kp = (m.es - alpha * beta / plen) / eta;
      plen = l * l * L * L * m.kT * (x0 - 0.25 + 1.0 / (4.0 * (1 - x0) * (1 - x0))) / kp;
    }
    double a1[3], a2[3], c1[3], c2[3], u1[3], v1[3], u2[3], v2[3];
    for (int d = 0; d < 3; d++) {
      u1[d] = P(n2, d) - P(n1, d); v1[d] = P(n3, d) - P(n1, d);
      u2[d] = P(n1, d) - P(n2, d); v2[d] = P(n4, d) - P(n2, d);
      c1[d] = (P(n1, d) + P(n2, d) + P(n3, d)) / 3.0;
      c2[d] = (P(n1, d) + P(n2, d) + P(n4, d)) / 3.0;
    }
    auto crs = [](const double *a, const double *b, double *o) {
      o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
    };
    crs(u1, v1, a1);
    crs(u2, v2, a2);
    const double m1 = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    const double m2 = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    double ct = 0, chk = 0;
    for (int d = 0; d < 3; d++) {
      ct += a1[d] / m1 * a2[d] / m2;
      chk += (a1[d] / m1 - a2[d] / m2) * (c1[d] - c2[d]);
    }
    ct = std::max(-1.0, std::min(1.0, ct));
    const double st = chk >= 0 ? std::sqrt(1 - ct * ct) : -std::sqrt(1 - ct * ct);
    eq[e] = L * l / x0;
    eq[ne + e] = plen;
    eq[2 * ne + e] = kp;
    eq[3 * ne + e] = std::asin(st);
    eq[4 * ne + e] = st;
    eq[5 * ne + e] = ct;
  }
}

void free_mesh(zf_ctx *ctx) {
  void **bufs[] = {(void **)&ctx->d_tri,       (void **)&ctx->d_edge,      (void **)&ctx->d_edge_eq,
                   (void **)&ctx->d_tri_area_eq, (void **)&ctx->d_node_tri, (void **)&ctx->d_inc_start,
                   (void **)&ctx->d_inc,       (void **)&ctx->d_x,         (void **)&ctx->d_v,
                   (void **)&ctx->d_mforce,    (void **)&ctx->d_nnormal,   (void **)&ctx->d_tgeo,
                   (void **)&ctx->d_cell,      (void **)&ctx->d_status,    (void **)&ctx->d_fail};
  for (auto b : bufs) dev_free(ctx, b);
  ctx->have_mesh = false;
  ctx->ncell = 0;
}

template <class T>
int upload(zf_ctx *ctx, T **d, const std::vector<T> &h) {
  int rc = dev_alloc(ctx, (void **)d, sizeof(T) * std::max<size_t>(h.size(), 1));
  if (rc) return rc;
  if (!h.empty()) ZF_HIP(ctx, hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return ZF_OK;
}

}  // namespace

void free_membrane(zf_ctx *ctx) {
  free_wrap(ctx);
  free_mesh(ctx);
}

}  // namespace zf

using namespace zf;

extern "C" {

void zf_membrane_config_reference(zf_membrane_config *m) {
  *m = zf_membrane_config{};
  m->len_scale = 1e-6;
  m->kT = 300 * 1.3806488e-23;  // TEMPERATURE*BOLTZ, ZebraflowWT.c:35-36
  m->x0 = 0.3154574132;         // e_ratio_eq, :133
  m->es = 3e-6;                 // Es, :128
  m->kbend = 1.3856406e-19;     // kBend, :127
  m->kd = 3.6e-5;
  m->ka = 7.7e-5;
  m->kv = 55.0;
  m->area0 = 144.916e-12;       // RBC_SURF, :126
  m->vol0 = 97.5671e-18;        // MCV, :125
  m->x_max = 0.95;
  m->x_floor = 0.02435;
  m->cos_cut = 0.99;
  m->cos_blend = 0.0;
  m->area_lo = 0.1;
  m->area_hi = 10.0;
  m->force_limit = 1e-10;
  m->vol_limit = 1e-10;
  m->sliver_len = 0.05;
  m->sliver_len_min = 0.001;
  m->sliver_f1 = 1e-14;
  m->sliver_f2 = 1e-11;
  m->relax = 1.0;
  m->relax2 = 1.0;
  m->passes = 51;
  m->max_cells = 1;
}

int zf_membrane_setup(zf_ctx *ctx, const zf_membrane_config *mc, int32_t nn, int32_t nt,
                      const int32_t *tri, const double *rest_xyz) {
  if (!ctx || !mc || !tri || !rest_xyz || nn <= 3 || nt <= 3 || mc->max_cells <= 0 || mc->passes < 1)
    return ZF_ERR_ARG;
  for (int i = 0; i < 3 * nt; i++)
    if (tri[i] < 1 || tri[i] > nn) return fail(ctx, ZF_ERR_ARG, "triangle node id out of range");
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  free_cells(ctx);  // sized by the old mesh
  free_ibm(ctx);
  free_nucleus(ctx);
  free_mesh(ctx);
  ctx->mcfg = *mc;
  ctx->nn = nn;
  ctx->nt = nt;
  build_topology(nn, nt, tri, ctx->h_edge, ctx->h_nbr_node, ctx->h_nbr_tri);
  const int ne = (int)ctx->h_edge.size() / 6;
  ctx->ne = ne;
  for (int e = 0; e < ne; e++)
    if (ctx->h_edge[6 * e + 3] == 0) return fail(ctx, ZF_ERR_GEOMETRY, "mesh is not closed (edge %d)", e + 1);
  for (int n = 0; n < nn; n++)
    if (ctx->h_nbr_tri[10 * n] > 9) return fail(ctx, ZF_ERR_GEOMETRY, "node %d has >9 triangles", n + 1);
  build_equilibrium(*mc, rest_xyz, ctx->h_edge, ctx->h_eq);
  // device tables
  std::vector<int32_t> dtri(3 * (size_t)nt), dedge(8 * (size_t)ne), ntri(10 * (size_t)nn);
  for (int i = 0; i < 3 * nt; i++) dtri[i] = tri[i] - 1;
  std::vector<int> owner(nt, -1);
  for (int e = 0; e < ne; e++)
    for (int s = 0; s < 2; s++) {
      const int t = ctx->h_edge[6 * e + 4 + s] - 1;
      if (owner[t] < 0) owner[t] = e;
    }
  for (int e = 0; e < ne; e++) {
    for (int k = 0; k < 6; k++) dedge[(size_t)k * ne + e] = ctx->h_edge[6 * e + k] - 1;
    dedge[6 * (size_t)ne + e] = owner[ctx->h_edge[6 * e + 4] - 1] == e;
    dedge[7 * (size_t)ne + e] = owner[ctx->h_edge[6 * e + 5] - 1] == e;
  }
  for (int n = 0; n < nn; n++) {
    ntri[10 * n] = ctx->h_nbr_tri[10 * n];
    for (int k = 1; k < 10; k++) ntri[10 * n + k] = ctx->h_nbr_tri[10 * n + k] - 1;
  }
  std::vector<double> eqd(5 * (size_t)ne), tae(nt);
  for (int e = 0; e < ne; e++) {
    eqd[e] = ctx->h_eq[e];
    eqd[ne + e] = ctx->h_eq[ne + e];
    eqd[2 * ne + e] = ctx->h_eq[2 * ne + e];
    eqd[3 * ne + e] = ctx->h_eq[4 * ne + e];
    eqd[4 * ne + e] = ctx->h_eq[5 * ne + e];
  }
  const double L = mc->len_scale;
  for (int t = 0; t < nt; t++) {
    const double *p1 = &rest_xyz[3 * (tri[3 * t] - 1)], *p2 = &rest_xyz[3 * (tri[3 * t + 1] - 1)],
                 *p3 = &rest_xyz[3 * (tri[3 * t + 2] - 1)];
    const double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double v[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
    const double ax = u[1] * v[2] - u[2] * v[1], ay = u[2] * v[0] - u[0] * v[2], az = u[0] * v[1] - u[1] * v[0];
    tae[t] = 0.5 * L * L * std::sqrt(ax * ax + ay * ay + az * az);
  }
  // node -> (edge, role) incidence in edge order
  std::vector<std::vector<int32_t>> per(nn);
  for (int e = 0; e < ne; e++)
    for (int r = 0; r < 4; r++) per[ctx->h_edge[6 * e + r] - 1].push_back((e << 2) | r);
  std::vector<int32_t> istart(nn + 1, 0), inc;
  for (int n = 0; n < nn; n++) {
    istart[n] = (int32_t)inc.size();
    inc.insert(inc.end(), per[n].begin(), per[n].end());
  }
  istart[nn] = (int32_t)inc.size();
  ctx->inc_max = 0;
  for (int n = 0; n < nn; n++) ctx->inc_max = std::max(ctx->inc_max, istart[n + 1] - istart[n]);
  int rc;
  if ((rc = upload(ctx, &ctx->d_tri, dtri)) || (rc = upload(ctx, &ctx->d_edge, dedge)) ||
      (rc = upload(ctx, &ctx->d_edge_eq, eqd)) || (rc = upload(ctx, &ctx->d_tri_area_eq, tae)) ||
      (rc = upload(ctx, &ctx->d_node_tri, ntri)) || (rc = upload(ctx, &ctx->d_inc_start, istart)) ||
      (rc = upload(ctx, &ctx->d_inc, inc)))
    return rc;
  const size_t ncap = (size_t)mc->max_cells, NT = ncap * nn, TT = ncap * nt;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_x, 24 * NT)) || (rc = dev_alloc(ctx, (void **)&ctx->d_v, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_mforce, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_nnormal, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_tgeo, 56 * TT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_cell, 64 * ncap)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_status, 4 * ncap)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_fail, 4 * ncap)))
    return rc;
  ctx->have_mesh = true;
  return ZF_OK;
}

int zf_download_mesh_topology(zf_ctx *ctx, int32_t *ne, int32_t *edge, int32_t *nbr_node,
                              int32_t *nbr_tri) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  if (ne) *ne = ctx->ne;
  if (edge) std::copy(ctx->h_edge.begin(), ctx->h_edge.end(), edge);
  if (nbr_node) std::copy(ctx->h_nbr_node.begin(), ctx->h_nbr_node.end(), nbr_node);
  if (nbr_tri) std::copy(ctx->h_nbr_tri.begin(), ctx->h_nbr_tri.end(), nbr_tri);
  return ZF_OK;
}

int zf_download_mesh_equilibrium(zf_ctx *ctx, double *lmax, double *plength, double *kp,
                                 double *theta0, double *sin0, double *cos0) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  double *outs[6] = {lmax, plength, kp, theta0, sin0, cos0};
  for (int k = 0; k < 6; k++)
    if (outs[k]) std::copy(ctx->h_eq.begin() + (size_t)k * ctx->ne, ctx->h_eq.begin() + (size_t)(k + 1) * ctx->ne, outs[k]);
  return ZF_OK;
}

int zf_upload_cells(zf_ctx *ctx, int32_t ncell, const double *xyz) {
  if (!ctx || (ncell > 0 && !xyz) || ncell < 0) return ZF_ERR_ARG;
  ctx->geom_gen++;  // positions change: repulsion bounds of the cell list expire
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  if (ncell > ctx->mcfg.max_cells) return fail(ctx, ZF_ERR_ARG, "%d cells > max_cells %d", ncell, ctx->mcfg.max_cells);
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn, n = (size_t)ncell * ctx->nn;
  ctx->ncell = ncell;
  ctx->nuc_lists_valid = false;
  std::vector<int32_t> st(ctx->mcfg.max_cells, -5000);
  // a parent lattice's cells start inside its window (status 2)
  for (int c = 0; c < ncell; c++) st[c] = ctx->have_wrap ? 2 : 1;
  ZF_HIP(ctx, hipMemcpy(ctx->d_status, st.data(), 4 * st.size(), hipMemcpyHostToDevice));
  if (ctx->have_wrap) ZF_HIP(ctx, hipMemset(ctx->d_regress, 0, 4 * (size_t)ctx->mcfg.max_cells));
  if (ctx->d_band) ZF_HIP(ctx, hipMemset(ctx->d_band, 0, 16 * (size_t)ctx->mcfg.max_cells));
  if (ctx->have_nuc) ZF_HIP(ctx, hipMemset(ctx->d_nuc_stale, 0, 4 * (size_t)ctx->mcfg.max_cells));
  if (n == 0) return ZF_OK;
  double *tmp = nullptr;
  int rc = dev_alloc(ctx, (void **)&tmp, 24 * n);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(tmp, xyz, 24 * n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_cells_import, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, tmp, ncell,
                       ctx->nn, NT, ctx->d_x, ctx->d_v);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= (int64_t)(24 * n);
  dev_free(ctx, (void **)&tmp);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "cell upload: %s", hipGetErrorString(e));
  return ZF_OK;
}

static int export_soa(zf_ctx *ctx, const double *src, double *dst) {
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn, n = (size_t)ctx->ncell * ctx->nn;
  if (!dst || n == 0) return ZF_OK;
  double *tmp = nullptr;
  int rc = dev_alloc(ctx, (void **)&tmp, 24 * n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_soa_export, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, src, (int)n, NT, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(dst, tmp, 24 * n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= (int64_t)(24 * n);
  dev_free(ctx, (void **)&tmp);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "export: %s", hipGetErrorString(e));
  return ZF_OK;
}

int zf_download_cells(zf_ctx *ctx, double *xyz, double *vel, int32_t *status) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  int rc;
  if ((rc = export_soa(ctx, ctx->d_x, xyz)) || (rc = export_soa(ctx, ctx->d_v, vel))) return rc;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (status && ctx->ncell)
    ZF_HIP(ctx, hipMemcpy(status, ctx->d_status, 4 * (size_t)ctx->ncell, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_membrane_forces(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  return membrane_forces_impl(ctx);
}

}  // extern "C"

memb::MeshTables zf::memb::mesh_tables(const zf_ctx *ctx) {
  return MeshTables{ctx->d_tri, ctx->d_node_tri, ctx->d_inc_start, ctx->d_inc, ctx->d_edge,
                    ctx->d_edge_eq, ctx->d_tri_area_eq, ctx->nn, ctx->nt, ctx->ne};
}

int zf::memb::membrane_geometry(zf_ctx *ctx, const double *X, double *tg, double *cell) {
  const MembParams mp = params_of(ctx->mcfg);
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn, TT = ncap * ctx->nt;
  const int nc = ctx->ncell;
  if (nc == 0) return ZF_OK;
  hipLaunchKernelGGL(k_tri_geom, dim3(nb((size_t)nc * ctx->nt, 256)), dim3(256), 0, ctx->stream, X,
                     NT, ctx->nn, ctx->nt, nc, ctx->d_tri, ctx->d_status, mp.L, tg, TT);
  hipLaunchKernelGGL(k_cell_reduce, dim3(nc), dim3(256), 0, ctx->stream, X, NT, ctx->nn, ctx->nt, tg,
                     TT, ctx->d_status, mp, cell);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int zf::membrane_forces_impl(zf_ctx *ctx) {
  ctx->geom_gen++;
  if (ctx->ncell == 0) return ZF_OK;
  const MembParams mp = params_of(ctx->mcfg);
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn, TT = ncap * ctx->nt;
  const int nc = ctx->ncell;
  hipLaunchKernelGGL(k_zero_i32, dim3(nb(nc, 256)), dim3(256), 0, ctx->stream, ctx->d_fail, nc);
  int rc = memb::membrane_geometry(ctx, cell_pos(ctx), ctx->d_tgeo, ctx->d_cell);
  if (rc) return rc;
  // FP32 node forces on request for the FP32 and FP16 storage forms
  if (ctx->cfg.storage != ZF_STORE_F64 && (ctx->cfg.flags & ZF_FLAG_F32_CELLS))
    hipLaunchKernelGGL(k_node_force<float>, dim3(nb((size_t)nc * ctx->nn, 128)), dim3(128), 0,
                       ctx->stream, cell_pos(ctx), NT, nc, memb::mesh_tables(ctx), ctx->d_tgeo, TT,
                       ctx->d_cell, ctx->d_status, mp, ctx->d_mforce, ctx->d_nnormal, ctx->d_fail);
  else
    hipLaunchKernelGGL(k_node_force<double>, dim3(nb((size_t)nc * ctx->nn, 128)), dim3(128), 0,
                       ctx->stream, cell_pos(ctx), NT, nc, memb::mesh_tables(ctx), ctx->d_tgeo, TT,
                       ctx->d_cell, ctx->d_status, mp, ctx->d_mforce, ctx->d_nnormal, ctx->d_fail);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

extern "C" {

int zf_download_membrane(zf_ctx *ctx, double *force, double *node_normal, double *cell_area,
                         double *cell_vol, double *cell_centroid, int32_t *guard_fail) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  int rc;
  if ((rc = export_soa(ctx, ctx->d_mforce, force)) || (rc = export_soa(ctx, ctx->d_nnormal, node_normal)))
    return rc;
  const int nc = ctx->ncell;
  if (nc == 0) return ZF_OK;
  std::vector<double> cell(8 * (size_t)nc);
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ZF_HIP(ctx, hipMemcpy(cell.data(), ctx->d_cell, 64 * (size_t)nc, hipMemcpyDeviceToHost));
  for (int c = 0; c < nc; c++) {
    if (cell_area) cell_area[c] = cell[8 * c];
    if (cell_vol) cell_vol[c] = cell[8 * c + 1];
    if (cell_centroid)
      for (int d = 0; d < 3; d++) cell_centroid[3 * c + d] = cell[8 * c + 2 + d];
  }
  if (guard_fail) ZF_HIP(ctx, hipMemcpy(guard_fail, ctx->d_fail, 4 * (size_t)nc, hipMemcpyDeviceToHost));
  return ZF_OK;
}

}  // extern "C"
