// zf_nucleus.hip -- the nucleus network of nucleated RBCs and its contact
// with the cell's membrane.
//
// Replaces, for the cells of this context, ZebraflowWT.c:
//   network data             NucleusEdges.dat / NucleusNodes.dat, :1903-1922
//   equilibrium parameters   :2003-2050 (WLC-POW fixed point, 0.1 Es),
//                            :2203-2221 (LMAX_N)
//   spring forces            :7388-7466 (F_Total_N)
//   nucleus-membrane contact :9519-9730 (triangle lists, 19-point probes)
// The nucleus nodes are immersed-boundary points of their cell: zf_ibm.hip
// interpolates, spreads and advances them with the membrane nodes.
//
// Spring forces: one lane per nucleus node gathers its edges in edge order
// (a CSR of (edge, role) per node), so each node's sum has the order of the
// reference's serial edge loop and no atomics are needed.  Contact: lists
// are rebuilt one lane per (cell, nucleus node) scanning the cell's
// triangles in order; the pushes go to the membrane nodes of the same
// cell, so one lane per cell walks its nucleus nodes, list entries and
// probe points in the reference's order and adds them one at a time
// (deterministic; cells touch disjoint membrane nodes).
#include "zf_internal.h"

#include <cmath>
#include <vector>

namespace zf {
namespace {

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct SpringParams {
  double L, kT, x_max, x_floor, relax2;
};

// WLC + POW force on node `role` (0: n1, 1: n2) of one filament
// (ZebraflowWT.c:7428-7463): the membrane's spring (:6888-6924).
__device__ __forceinline__ void spring_force(const double *p1, const double *p2, double lmax,
                                             double plen, double kp, int role,
                                             const SpringParams &sp, double f[3]) {
  double d[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double e[3] = {d[0] / len, d[1] / len, d[2] / len};
  double ratio = len * sp.L / lmax;
  if (ratio > sp.x_max) ratio = sp.x_max;
  if (ratio < sp.x_floor) len = sp.x_floor * lmax / sp.L;
  const double wlc = (-sp.kT / plen) * (ratio - 0.25 + 1.0 / (4.0 * (1 - ratio) * (1 - ratio)));
  const double comp = kp / (len * len * sp.L * sp.L);
  const double sg = role == 0 ? 1.0 : -1.0;
  for (int a = 0; a < 3; a++) f[a] = (sg * sp.relax2 * wlc) * e[a] + (sg * sp.relax2 * comp) * e[a];
}

__global__ void k_nuc_force(const double *__restrict__ X, size_t NT, int nn, int ncell,
                            const int32_t *__restrict__ status, const int32_t *__restrict__ inc_start,
                            const int32_t *__restrict__ inc, const int32_t *__restrict__ edge,
                            const double *__restrict__ eq, int ne, SpringParams sp,
                            double *__restrict__ F) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  const int c = g / nn, node = g % nn;
  double acc[3] = {0, 0, 0};
  if (status[c] > kStatusDead) {
    const size_t b0 = (size_t)c * nn;
    for (int t = inc_start[node]; t < inc_start[node + 1]; t++) {
      const int e = inc[t] >> 1, role = inc[t] & 1;
      const size_t n1 = b0 + edge[2 * e], n2 = b0 + edge[2 * e + 1];
      const double p1[3] = {X[n1], X[NT + n1], X[2 * NT + n1]};
      const double p2[3] = {X[n2], X[NT + n2], X[2 * NT + n2]};
      double f[3];
      spring_force(p1, p2, eq[e], eq[ne + e], eq[2 * ne + e], role, sp, f);
      for (int a = 0; a < 3; a++) acc[a] += f[a];
    }
  }
  for (int a = 0; a < 3; a++) F[a * NT + g] = acc[a];
}

// Lists (ZebraflowWT.c:9529-9590): the cell's triangles, in order, whose
// centroid is closer than r to the nucleus node.  Entries past `cap` are
// counted in *over.
__global__ void k_nuc_list(const double *__restrict__ Xn, size_t NTn, int nnuc, int ncell,
                           const int32_t *__restrict__ status, const double *__restrict__ tg,
                           size_t TT, int nt, double r, int cap, int32_t *__restrict__ list,
                           int32_t *__restrict__ cnt, int32_t *__restrict__ over,
                           const int32_t *__restrict__ stale, int all) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nnuc) return;
  const int c = g / nnuc;
  // all cells on the rebuild rule, else the cells marked since (new clones)
  if (!all && !stale[c]) return;
  int n = 0, lost = 0;
  if (status[c] > kStatusDead) {
    const double x = Xn[g], y = Xn[NTn + g], z = Xn[2 * NTn + g];
    const size_t t0 = (size_t)c * nt;
    for (int t = 0; t < nt; t++) {
      const double dx = x - tg[4 * TT + t0 + t], dy = y - tg[5 * TT + t0 + t],
                   dz = z - tg[6 * TT + t0 + t];
      if (sqrt(dx * dx + dy * dy + dz * dz) < r) {
        if (n < cap) list[(size_t)g * cap + n++] = t;
        else lost++;
      }
    }
  }
  cnt[g] = n;
  if (lost) atomicAdd(over, lost);
}

struct ContactParams {
  double near_r, contact, k;
};

// Contact (ZebraflowWT.c:9594-9728), one lane per cell.  The 19 probe
// points of a triangle (:9662-9699): its centroid, the three corner
// sub-triangle centroids, the three edge-midpoint pairs' midpoints and twelve
// midpoints towards them, in the reference's order.
__global__ void k_nuc_contact(const double *__restrict__ Xn, size_t NTn, int nnuc,
                              const double *__restrict__ X, size_t NT, int nn, int ncell,
                              const int32_t *__restrict__ status, const int32_t *__restrict__ tri,
                              const double *__restrict__ tg, size_t TT, int nt,
                              const int32_t *__restrict__ list, const int32_t *__restrict__ cnt,
                              int cap, ContactParams cp, double *__restrict__ vrep) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell || status[c] <= kStatusDead) return;
  const size_t b0 = (size_t)c * nn, t0 = (size_t)c * nt;
  for (int node = 0; node < nnuc; node++) {
    const size_t g = (size_t)c * nnuc + node;
    const double xc = Xn[g], yc = Xn[NTn + g], zc = Xn[2 * NTn + g];
    for (int l = 0; l < cnt[g]; l++) {
      const int t = list[g * cap + l];
      const size_t q = t0 + t;
      const double dcx = xc - tg[4 * TT + q], dcy = yc - tg[5 * TT + q], dcz = zc - tg[6 * TT + q];
      if (!(sqrt(dcx * dcx + dcy * dcy + dcz * dcz) < cp.near_r)) continue;
      const int n1 = tri[3 * t], n2 = tri[3 * t + 1], n3 = tri[3 * t + 2];
      const double nv[3] = {tg[TT + q], tg[2 * TT + q], tg[3 * TT + q]};
      double tp[20][3];
      for (int v = 0; v < 3; v++) {
        const double p1 = X[v * NT + b0 + n1], p2 = X[v * NT + b0 + n2], p3 = X[v * NT + b0 + n3];
        const double e1 = (p1 + p2) / 2., e2 = (p2 + p3) / 2., e3 = (p3 + p1) / 2.;
        tp[1][v] = (p1 + p2 + p3) / 3.;
        tp[2][v] = (e3 + p1 + e1) / 3.;
        tp[3][v] = (e1 + p2 + e2) / 3.;
        tp[4][v] = (e2 + p3 + e3) / 3.;
        tp[5][v] = (e1 + e3) / 2.;
        tp[6][v] = (e1 + e2) / 2.;
        tp[7][v] = (e2 + e3) / 2.;
        tp[8][v] = (p1 + tp[2][v]) / 2.;
        tp[9][v] = (e1 + tp[2][v]) / 2.;
        tp[10][v] = (e3 + tp[2][v]) / 2.;
        tp[11][v] = (e1 + tp[3][v]) / 2.;
        tp[12][v] = (p2 + tp[3][v]) / 2.;
        tp[13][v] = (e2 + tp[3][v]) / 2.;
        tp[14][v] = (e3 + tp[4][v]) / 2.;
        tp[15][v] = (e2 + tp[4][v]) / 2.;
        tp[16][v] = (p3 + tp[4][v]) / 2.;
        tp[17][v] = (e1 + tp[1][v]) / 2.;
        tp[18][v] = (e2 + tp[1][v]) / 2.;
        tp[19][v] = (e3 + tp[1][v]) / 2.;
      }
      for (int v = 1; v <= 19; v++) {
        const double dx = tp[v][0] - xc, dy = tp[v][1] - yc, dz = tp[v][2] - zc;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d <= cp.contact) {
          const double rf = (cp.contact - d) / cp.contact * cp.k;
          for (int a = 0; a < 3; a++) {
            const double add = rf * nv[a] / 3.;
            vrep[a * NT + b0 + n1] += add;
            vrep[a * NT + b0 + n2] += add;
            vrep[a * NT + b0 + n3] += add;
          }
        }
      }
    }
  }
}

__global__ void k_zero(double *__restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

__global__ void k_zero_i32(int32_t *__restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_soa_in(const double *__restrict__ h, int n, size_t NT, double *__restrict__ A) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  for (int a = 0; a < 3; a++) A[a * NT + g] = h[3 * (size_t)g + a];
}

__global__ void k_soa_out(const double *__restrict__ A, int n, size_t NT, double *__restrict__ h) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  for (int a = 0; a < 3; a++) h[3 * (size_t)g + a] = A[a * NT + g];
}

// Filament parameters of the rest network (ZebraflowWT.c:2003-2050,
// :2203-2221): the membrane's fixed point for Plength, kp with es_factor es.
void nucleus_equilibrium(const zf_membrane_config &m, double es, const double *x,
                         const std::vector<int32_t> &edge, std::vector<double> &eq) {
  const int ne = (int)edge.size() / 2;
  const double L = m.len_scale, x0 = m.x0;
  const double beta = 0.5 * x0 / std::pow(1 - x0, 3) - 0.25 / ((1 - x0) * (1 - x0)) + 0.25;
  eq.assign((size_t)3 * ne, 0.0);
  for (int e = 0; e < ne; e++) {
    const double *p1 = &x[3 * (size_t)(edge[2 * e] - 1)], *p2 = &x[3 * (size_t)(edge[2 * e + 1] - 1)];
    double l2 = 0;
    for (int d = 0; d < 3; d++) l2 += (p1[d] - p2[d]) * (p1[d] - p2[d]);
    const double l = std::sqrt(l2);
    const double alpha = std::sqrt(3.0) * 0.25 * m.kT / (l * L);
    const double eta = std::sqrt(3.0) * 0.25 * 3.0 / (l * l * l * L * L * L);
    double plen = 1e-9 * (L / 1e-6), kp = 0;
    for (int pass = 0; pass < m.passes; pass++) {
      kp = (es - alpha * beta / plen) / eta;
      plen = l * l * L * L * m.kT * (x0 - 0.25 + 1.0 / (4.0 * (1 - x0) * (1 - x0))) / kp;
    }
    eq[e] = L * l / x0;
    eq[ne + e] = plen;
    eq[2 * ne + e] = kp;
  }
}

template <class T>
int upload_vec(zf_ctx *ctx, T **d, const std::vector<T> &h) {
  int rc = dev_alloc(ctx, (void **)d, sizeof(T) * std::max<size_t>(h.size(), 1));
  if (rc) return rc;
  if (!h.empty()) ZF_HIP(ctx, hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return ZF_OK;
}

int require_nuc(zf_ctx *ctx) {
  if (!ctx->have_nuc) return fail(ctx, ZF_ERR_STATE, "zf_nucleus_setup not called");
  return ZF_OK;
}

}  // namespace

void free_nucleus(zf_ctx *ctx) {
  void **bufs[] = {(void **)&ctx->d_nuc_edge, (void **)&ctx->d_nuc_eq, (void **)&ctx->d_nuc_inc_start,
                   (void **)&ctx->d_nuc_inc, (void **)&ctx->d_xn, (void **)&ctx->d_vn,
                   (void **)&ctx->d_fn, (void **)&ctx->d_vrepn, (void **)&ctx->d_ffln,
                   (void **)&ctx->d_nlist, (void **)&ctx->d_nlist_cnt, (void **)&ctx->d_nuc_stale};
  for (auto b : bufs) dev_free(ctx, b);
  ctx->have_nuc = false;
  ctx->nnuc = ctx->ne_n = 0;
  ctx->nuc_lists_valid = false;
}

int nucleus_forces_impl(zf_ctx *ctx) {
  const int n = ctx->ncell * ctx->nnuc;
  if (n == 0) return ZF_OK;
  const size_t NTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc;
  const zf_membrane_config &m = ctx->mcfg;
  const SpringParams sp{m.len_scale, m.kT, m.x_max, m.x_floor, m.relax2};
  hipLaunchKernelGGL(k_nuc_force, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, nuc_pos(ctx), NTn,
                     ctx->nnuc, ctx->ncell, ctx->d_status, ctx->d_nuc_inc_start, ctx->d_nuc_inc,
                     ctx->d_nuc_edge, ctx->d_nuc_eq, ctx->ne_n, sp, ctx->d_fn);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int nucleus_contact_impl(zf_ctx *ctx, int iter) {
  const int n = ctx->ncell * ctx->nnuc;
  if (n == 0) return ZF_OK;
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const size_t NTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc;
  const size_t TT = (size_t)ctx->mcfg.max_cells * ctx->nt;
  const zf_nucleus_config &nc = ctx->ncfg;
  const int cap = nc.list_cap;
  const bool all = !ctx->nuc_lists_valid || (nc.list_interval > 0 && iter % nc.list_interval == 0);
  // a main lattice with parents' cells also rebuilds the lists of the
  // clones made since (the reference's NucMembUpdate of a new cell)
  if (all || !ctx->cell_parents.empty()) {
    hipLaunchKernelGGL(k_nuc_list, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, nuc_pos(ctx), NTn,
                       ctx->nnuc, ctx->ncell, ctx->d_status, ctx->d_tgeo, TT, ctx->nt,
                       nc.list_radius, cap, ctx->d_nlist, ctx->d_nlist_cnt,
                       ctx->d_nlist_cnt + NTn, ctx->d_nuc_stale, all ? 1 : 0);
    hipLaunchKernelGGL(k_zero_i32, dim3(nb(ctx->ncell, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_nuc_stale, ctx->ncell);
    ctx->nuc_lists_valid = true;
  }
  const ContactParams cp{nc.near_radius, nc.contact, nc.k_contact};
  hipLaunchKernelGGL(k_nuc_contact, dim3(nb(ctx->ncell, 64)), dim3(64), 0, ctx->stream, nuc_pos(ctx),
                     NTn, ctx->nnuc, cell_pos(ctx), NT, ctx->nn, ctx->ncell, ctx->d_status, ctx->d_tri,
                     ctx->d_tgeo, TT, ctx->nt, ctx->d_nlist, ctx->d_nlist_cnt, cap, cp, ctx->d_vrep);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

}  // namespace zf

using namespace zf;

extern "C" {

void zf_nucleus_config_reference(zf_nucleus_config *nc) {
  if (!nc) return;
  *nc = zf_nucleus_config{};
  nc->es_factor = 0.1;
  nc->list_radius = 0.5;
  nc->near_radius = 1.0;
  nc->contact = 0.2;
  nc->k_contact = 4.1666667e-07;
  nc->list_interval = 1000;
  nc->list_cap = 64;
}

int zf_nucleus_setup(zf_ctx *ctx, const zf_nucleus_config *nc, int32_t nnuc, int32_t ne,
                     const int32_t *edge, const double *rest_xyz) {
  if (!ctx || !nc || nnuc < 2 || ne < 1 || !edge || !rest_xyz || nc->list_cap < 1 ||
      nc->contact <= 0 || nc->list_radius <= 0)
    return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "zf_membrane_setup not called");
  for (size_t i = 0; i < 2 * (size_t)ne; i++)
    if (edge[i] < 1 || edge[i] > nnuc) return fail(ctx, ZF_ERR_ARG, "nucleus edge node out of range");
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  free_nucleus(ctx);
  ctx->ncfg = *nc;
  ctx->nnuc = nnuc;
  ctx->ne_n = ne;
  ctx->h_nuc_edge.assign(edge, edge + 2 * (size_t)ne);
  nucleus_equilibrium(ctx->mcfg, nc->es_factor * ctx->mcfg.es, rest_xyz, ctx->h_nuc_edge, ctx->h_nuc_eq);
  std::vector<int32_t> e0(2 * (size_t)ne), start(nnuc + 1, 0), inc(2 * (size_t)ne);
  for (size_t i = 0; i < 2 * (size_t)ne; i++) e0[i] = edge[i] - 1;
  for (size_t i = 0; i < 2 * (size_t)ne; i++) start[e0[i] + 1]++;
  for (int k = 0; k < nnuc; k++) start[k + 1] += start[k];
  std::vector<int32_t> fill(start.begin(), start.end() - 1);
  for (int e = 0; e < ne; e++)
    for (int r = 0; r < 2; r++) inc[fill[e0[2 * e + r]]++] = (e << 1) | r;
  const size_t NTn = (size_t)ctx->mcfg.max_cells * nnuc;
  int rc;
  if ((rc = upload_vec(ctx, &ctx->d_nuc_edge, e0)) || (rc = upload_vec(ctx, &ctx->d_nuc_eq, ctx->h_nuc_eq)) ||
      (rc = upload_vec(ctx, &ctx->d_nuc_inc_start, start)) || (rc = upload_vec(ctx, &ctx->d_nuc_inc, inc)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_xn, 24 * NTn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_vn, 24 * NTn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_fn, 24 * NTn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_vrepn, 24 * NTn)) ||
      (rc = dev_alloc(ctx, &ctx->d_ffln, 3 * 8 * NTn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_nlist, 4 * NTn * (size_t)nc->list_cap)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_nlist_cnt, 4 * (NTn + 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_nuc_stale, 4 * (size_t)ctx->mcfg.max_cells))) {
    free_nucleus(ctx);
    return rc;
  }
  double *zero[] = {ctx->d_xn, ctx->d_vn, ctx->d_fn, ctx->d_vrepn};
  for (double *z : zero)
    hipLaunchKernelGGL(k_zero, dim3(nb(3 * NTn, 256)), dim3(256), 0, ctx->stream, z, 3 * NTn);
  hipLaunchKernelGGL(k_zero_i32, dim3(nb(NTn + 1, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_nlist_cnt, NTn + 1);
  hipLaunchKernelGGL(k_zero_i32, dim3(nb(ctx->mcfg.max_cells, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_nuc_stale, (size_t)ctx->mcfg.max_cells);
  ZF_HIP(ctx, hipGetLastError());
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_nuc = true;
  ctx->nuc_lists_valid = false;
  // the spreading's sort keys hold the nucleus nodes after the membrane's
  if (ctx->have_ibm &&
      (rc = ensure_sort_keys(ctx, (size_t)ctx->mcfg.max_cells * (ctx->nn + nnuc)))) {
    free_nucleus(ctx);
    return rc;
  }
  return ZF_OK;
}

int zf_download_nucleus_equilibrium(zf_ctx *ctx, double *lmax, double *plength, double *kp) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  if (rc) return rc;
  double *outs[3] = {lmax, plength, kp};
  for (int a = 0; a < 3; a++)
    if (outs[a])
      for (int e = 0; e < ctx->ne_n; e++) outs[a][e] = ctx->h_nuc_eq[(size_t)a * ctx->ne_n + e];
  return ZF_OK;
}

int zf_upload_nuclei(zf_ctx *ctx, const double *xyz) {
  if (!ctx || !xyz) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  if (rc) return rc;
  const int n = ctx->ncell * ctx->nnuc;
  if (n == 0) return fail(ctx, ZF_ERR_STATE, "upload the cells first");
  const size_t NTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc;
  double *tmp = nullptr;
  if ((rc = dev_alloc(ctx, (void **)&tmp, 24 * (size_t)n))) return rc;
  hipError_t e = hipMemcpyAsync(tmp, xyz, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_soa_in, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, tmp, n, NTn, ctx->d_xn);
    hipLaunchKernelGGL(k_zero, dim3(nb(3 * NTn, 256)), dim3(256), 0, ctx->stream, ctx->d_vn, 3 * NTn);
    hipLaunchKernelGGL(k_zero, dim3(nb(3 * NTn, 256)), dim3(256), 0, ctx->stream, ctx->d_vrepn, 3 * NTn);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= 24 * (int64_t)n;
  dev_free(ctx, (void **)&tmp);
  ctx->nuc_lists_valid = false;
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "nucleus upload: %s", hipGetErrorString(e));
  return ZF_OK;
}

int zf_download_nuclei(zf_ctx *ctx, double *xyz, double *vel, double *force) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  if (rc) return rc;
  const int n = ctx->ncell * ctx->nnuc;
  if (n == 0) return ZF_OK;
  const size_t NTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc;
  double *tmp = nullptr;
  if ((rc = dev_alloc(ctx, (void **)&tmp, 24 * (size_t)n))) return rc;
  const double *src[3] = {ctx->d_xn, ctx->d_vn, ctx->d_fn};
  double *dst[3] = {xyz, vel, force};
  hipError_t e = hipSuccess;
  for (int a = 0; a < 3 && e == hipSuccess; a++) {
    if (!dst[a]) continue;
    hipLaunchKernelGGL(k_soa_out, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, src[a], n, NTn, tmp);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(dst[a], tmp, 24 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  ctx->device_bytes -= 24 * (int64_t)n;
  dev_free(ctx, (void **)&tmp);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "nucleus download: %s", hipGetErrorString(e));
  return ZF_OK;
}

int zf_get_nucleus_info(zf_ctx *ctx, int32_t *nnuc, int32_t *list_cap) {
  if (!ctx || !nnuc || !list_cap) return ZF_ERR_ARG;
  *nnuc = ctx->have_nuc ? ctx->nnuc : 0;
  *list_cap = ctx->have_nuc ? ctx->ncfg.list_cap : 0;
  return ZF_OK;
}

int zf_download_nucleus_lists(zf_ctx *ctx, int32_t *list, int32_t *count, int32_t *valid) {
  if (!ctx || !list || !count || !valid) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  if (rc) return rc;
  const size_t n = (size_t)ctx->ncell * ctx->nnuc, cap = ctx->ncfg.list_cap;
  *valid = ctx->nuc_lists_valid ? 1 : 0;
  if (n == 0) return ZF_OK;
  ZF_HIP(ctx, hipMemcpyAsync(list, ctx->d_nlist, 4 * n * cap, hipMemcpyDeviceToHost, ctx->stream));
  ZF_HIP(ctx, hipMemcpyAsync(count, ctx->d_nlist_cnt, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ZF_OK;
}

int zf_upload_nucleus_lists(zf_ctx *ctx, const int32_t *list, const int32_t *count, int32_t valid) {
  if (!ctx || !list || !count) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  if (rc) return rc;
  const size_t n = (size_t)ctx->ncell * ctx->nnuc, cap = ctx->ncfg.list_cap;
  if (n == 0) return fail(ctx, ZF_ERR_STATE, "upload the cells first");
  const int32_t ntri = ctx->nt;
  for (size_t q = 0; q < n; q++) {
    if (count[q] < 0 || (size_t)count[q] > cap)
      return fail(ctx, ZF_ERR_ARG, "nucleus node %zu: list count %d outside [0, %zu]", q, count[q], cap);
    for (int32_t k = 0; k < count[q]; k++)
      if (list[q * cap + k] < 0 || list[q * cap + k] >= ntri)
        return fail(ctx, ZF_ERR_ARG, "nucleus node %zu: triangle %d outside the mesh", q,
                    list[q * cap + k]);
  }
  ZF_HIP(ctx, hipMemcpyAsync(ctx->d_nlist, list, 4 * n * cap, hipMemcpyHostToDevice, ctx->stream));
  ZF_HIP(ctx, hipMemcpyAsync(ctx->d_nlist_cnt, count, 4 * n, hipMemcpyHostToDevice, ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->nuc_lists_valid = valid != 0;
  return ZF_OK;
}

int zf_nucleus_forces(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  return rc ? rc : nucleus_forces_impl(ctx);
}

int zf_nucleus_contact(zf_ctx *ctx, int32_t iter, int32_t *list_overflow) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_nuc(ctx);
  if (rc) return rc;
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "zf_ibm_setup not called (the contact adds to v_rep)");
  if ((rc = nucleus_contact_impl(ctx, iter))) return rc;
  if (list_overflow) {
    const size_t NTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc;
    ZF_HIP(ctx, hipMemcpyAsync(list_overflow, ctx->d_nlist_cnt + NTn, 4, hipMemcpyDeviceToHost,
                               ctx->stream));
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return ZF_OK;
}

}  // extern "C"
