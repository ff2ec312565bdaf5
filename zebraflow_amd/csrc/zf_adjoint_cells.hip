// zf_adjoint_cells.hip -- transpose of the membrane forces (vector-Jacobian
// product with respect to the node positions and the two modulus scales).
//
// F_n(X) of zf_membrane.hip depends on X through the rhombi (edges) that
// contain node n, the anti-sliver terms of its triangles, and three cell
// scalars: the centroid cc = mean(X), the global area force
// f_ag = -ka relax (A - A0)/A0 and the volume force f_vs = -kv (V - V0)/V0.
// For an adjoint lam (one 3-vector per node):
//   lam_X(m) = sum_{edges e containing m} sum_{roles r} lam(node_r(e)) . dF_e,r/dX_m
//            + sum_{triangles t of m} sum_{vertices n of t} lam(n) . dsliver(n, t)/dX_m
//            + Lag df_ag/dX_m + Lvs df_vs/dX_m + Lcc / nn,
// with Lag = sum_n lam(n) . dF_n/df_ag (and Lvs, Lcc alike) per cell.  The
// local derivatives come from the force code itself evaluated in dual
// numbers (zf_membrane_dev.h): one lane per (cell, node m) seeds X_m and
// gathers every term that X_m enters, in a fixed order, so no atomics.
//
// Modulus scales.  With the lattice-unit constants (kT, kd, ka, kv, kp and
// the anti-sliver forces proportional to es; kbend to kb), F = F_nb + F_b
// with F_nb proportional to es and F_b (bending) to kbend, so
//   dJ/d ln es = sum lam . F_nb,   dJ/d ln kbend = sum lam . F_b.
#include "zf_membrane_dev.h"

#include <vector>

namespace zf {
namespace {

using namespace memb;
using D3 = Dual<3>;
using D5 = Dual<5>;

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// The transposes run one lane per (cell, node m, slot j), j < J = inc_max + 1:
// slot j < count(m) takes the j-th rhombus containing m, slot count(m) the
// node's own terms (anti-sliver, and in pass 3 its triangles' area and
// volume); the rest write zeros.  Partial results go to slot-major scratch
// and are summed in slot order, so the sums do not depend on scheduling.

// Pass 1: lam(m) . d(contribution)/d(f_ag, cc_x, cc_y, cc_z, f_vs), and the
// contribution's non-bending and bending parts against lam(m).  part: 7 x
// NT*J.  (Dual<5> terms: 256 VGPRs without spills at 128 lanes per group.)
__global__ __launch_bounds__(128) void k_memb_cell_terms(
    const double *__restrict__ X, size_t NT, int ncell, MeshTables mt, int J,
    const double *__restrict__ tg, size_t TT, const double *__restrict__ cell,
    const int32_t *__restrict__ status, MembParams mp, const double *__restrict__ lam,
    double *__restrict__ part) {
  const int nn = mt.nn, ne = mt.ne;
  const size_t NJ = NT * J;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ncell * nn * J) return;
  const int j = (int)(idx % J);
  const size_t q = idx / J;
  const int c = (int)(q / nn), m = (int)(q % nn);
  const int k0 = mt.inc_start[m], cnt = mt.inc_start[m + 1] - k0;
  double out[7] = {0, 0, 0, 0, 0, 0, 0};
  if (status[c] >= 0 && j <= cnt) {
    const size_t nbase = (size_t)c * nn, tbase = (size_t)c * mt.nt;
    const double l0 = lam[q], l1 = lam[NT + q], l2 = lam[2 * NT + q];
    if (j < cnt) {
      D5 f_ag(cell[8 * c + 5]), f_vs(cell[8 * c + 6]);
      V3<D5> cc = {D5(cell[8 * c + 2]), D5(cell[8 * c + 3]), D5(cell[8 * c + 4])};
      f_ag.d[0] = 1.0;
      cc.x.d[1] = 1.0;
      cc.y.d[2] = 1.0;
      cc.z.d[3] = 1.0;
      f_vs.d[4] = 1.0;
      const int code = mt.inc[k0 + j], e = code >> 2;
      const V3<D5> p1 = lift<D5>(ld3(X, NT, nbase + mt.E[e])),
                   p2 = lift<D5>(ld3(X, NT, nbase + mt.E[ne + e])),
                   p3 = lift<D5>(ld3(X, NT, nbase + mt.E[2 * ne + e])),
                   p4 = lift<D5>(ld3(X, NT, nbase + mt.E[3 * ne + e]));
      const TriG<D5> g1 = tri_load<D5>(tg, TT, tbase + mt.E[4 * ne + e]);
      const TriG<D5> g2 = tri_load<D5>(tg, TT, tbase + mt.E[5 * ne + e]);
      const EdgeConst ec = edge_const(mt, e);
      EdgeOut<D5> o;
      switch (code & 3) {  // one inlined copy per role
        case 0: o = edge_role(0, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp); break;
        case 1: o = edge_role(1, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp); break;
        case 2: o = edge_role(2, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp); break;
        default: o = edge_role(3, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp);
      }
      // this rhombus's share of F(m): f + relax f_vs fv
      const V3<D5> t = o.f + (mp.relax * f_vs) * o.fv;
#pragma unroll
      for (int k = 0; k < 5; k++) out[k] = l0 * t.x.d[k] + l1 * t.y.d[k] + l2 * t.z.d[k];
      const double lb = l0 * o.fb.x.v + l1 * o.fb.y.v + l2 * o.fb.z.v;
      out[5] = l0 * t.x.v + l1 * t.y.v + l2 * t.z.v - lb;
      out[6] = lb;
    } else {  // the anti-sliver terms of m (shear-modulus part)
      const V3<double> xn = ld3(X, NT, q);
      V3<double> f = {0, 0, 0};
      const int ntri = mt.node_tri[10 * m];
      for (int k = 1; k <= ntri && k < 10; k++) {
        const size_t qt = tbase + mt.node_tri[10 * m + k];
        f = f + sliver(xn, V3<double>{tg[4 * TT + qt], tg[5 * TT + qt], tg[6 * TT + qt]}, mp);
      }
      out[5] = l0 * f.x + l1 * f.y + l2 * f.z;
    }
  }
#pragma unroll
  for (int k = 0; k < 7; k++) part[k * NJ + idx] = out[k];
}

// Pass 2, one workgroup per cell: fixed-order tree sums of the 7 terms over
// the cell's slots.  out: 8 per cell.
__global__ __launch_bounds__(256) void k_memb_cell_sum(const double *__restrict__ part, size_t NJ,
                                                       int nslot, const int32_t *__restrict__ status,
                                                       double *__restrict__ out) {
  __shared__ double sh[7][256];
  const int c = blockIdx.x;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  if (status[c] >= 0)
    for (int i = threadIdx.x; i < nslot; i += 256)
      for (int k = 0; k < 7; k++) s[k] += part[k * NJ + (size_t)c * nslot + i];
  for (int k = 0; k < 7; k++) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 7; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < 7) out[8 * c + threadIdx.x] = sh[threadIdx.x][0];
}

__device__ __forceinline__ V3<D3> seeded(const double *__restrict__ X, size_t NT, size_t nbase, int node,
                                         int m) {
  V3<D3> p = lift<D3>(ld3(X, NT, nbase + node));
  if (node == m) {
    p.x.d[0] = 1.0;
    p.y.d[1] = 1.0;
    p.z.d[2] = 1.0;
  }
  return p;
}

__device__ __forceinline__ void acc(double g[3], const double *__restrict__ lam, size_t NT, size_t q,
                                    const V3<D3> &f) {
  const double l0 = lam[q], l1 = lam[NT + q], l2 = lam[2 * NT + q];
#pragma unroll
  for (int a = 0; a < 3; a++) g[a] += l0 * f.x.d[a] + l1 * f.y.d[a] + l2 * f.z.d[a];
}

// Pass 3: d/dX_m of the slot's terms against lam.  Slot j < count(m): the
// j-th rhombus containing m, all four roles; slot count(m): the anti-sliver
// terms of m's triangles (every vertex of each) and the cell-scalar chain
// through those triangles' area and volume and the centroid.  part: 3 x NT*J.
__global__ __launch_bounds__(128) void k_memb_adj_pos(
    const double *__restrict__ X, size_t NT, int ncell, MeshTables mt, int J,
    const double *__restrict__ cell, const double *__restrict__ cterm,
    const int32_t *__restrict__ status, MembParams mp, const double *__restrict__ lam,
    double *__restrict__ part) {
  const int nn = mt.nn, ne = mt.ne;
  const size_t NJ = NT * J;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)ncell * nn * J) return;
  const int j = (int)(idx % J);
  const size_t q = idx / J;
  const int c = (int)(q / nn), m = (int)(q % nn);
  const int k0 = mt.inc_start[m], cnt = mt.inc_start[m + 1] - k0;
  double g[3] = {0, 0, 0};
  if (status[c] >= 0 && j <= cnt) {
    const size_t nbase = (size_t)c * nn;
    if (j < cnt) {
      const V3<D3> cc = {D3(cell[8 * c + 2]), D3(cell[8 * c + 3]), D3(cell[8 * c + 4])};
      const D3 f_ag(cell[8 * c + 5]);
      const double fvs = mp.relax * cell[8 * c + 6];
      const int e = mt.inc[k0 + j] >> 2;
      int nd[4];
#pragma unroll
      for (int r = 0; r < 4; r++) nd[r] = mt.E[r * ne + e];
      const V3<D3> p1 = seeded(X, NT, nbase, nd[0], m), p2 = seeded(X, NT, nbase, nd[1], m),
                   p3 = seeded(X, NT, nbase, nd[2], m), p4 = seeded(X, NT, nbase, nd[3], m);
      const int t1 = mt.E[4 * ne + e], t2 = mt.E[5 * ne + e];
      const TriG<D3> g1 = tri_geom(seeded(X, NT, nbase, mt.tri[3 * t1], m),
                                   seeded(X, NT, nbase, mt.tri[3 * t1 + 1], m),
                                   seeded(X, NT, nbase, mt.tri[3 * t1 + 2], m), mp.L);
      const TriG<D3> g2 = tri_geom(seeded(X, NT, nbase, mt.tri[3 * t2], m),
                                   seeded(X, NT, nbase, mt.tri[3 * t2 + 1], m),
                                   seeded(X, NT, nbase, mt.tri[3 * t2 + 2], m), mp.L);
      const EdgeConst ec = edge_const(mt, e);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const EdgeOut<D3> o = edge_role(r, p1, p2, p3, p4, g1, g2, ec, cc, f_ag, mp);
        acc(g, lam, NT, nbase + nd[r], o.f + fvs * o.fv);
      }
    } else {
      double dA[3] = {0, 0, 0}, dV[3] = {0, 0, 0};
      const int ntri = mt.node_tri[10 * m];
      for (int k = 1; k <= ntri && k < 10; k++) {
        const int t = mt.node_tri[10 * m + k];
        const int v[3] = {mt.tri[3 * t], mt.tri[3 * t + 1], mt.tri[3 * t + 2]};
        const V3<D3> pv[3] = {seeded(X, NT, nbase, v[0], m), seeded(X, NT, nbase, v[1], m),
                              seeded(X, NT, nbase, v[2], m)};
        const TriG<D3> tgm = tri_geom(pv[0], pv[1], pv[2], mp.L);
#pragma unroll
        for (int i = 0; i < 3; i++) acc(g, lam, NT, nbase + v[i], sliver(pv[i], tgm.C, mp));
        const D3 vol = tgm.a * dot(tgm.C, tgm.m);
#pragma unroll
        for (int a = 0; a < 3; a++) {
          dA[a] += tgm.a.d[a];
          dV[a] += vol.d[a];
        }
      }
      const double *L = cterm + 8 * c;  // Lag, Lcc(3), Lvs
      const double kA = -mp.ka * mp.relax / mp.area0, kV = -mp.kv / mp.vol0 * (mp.L / 3.0);
      const double inv_nn = 1.0 / nn;
#pragma unroll
      for (int a = 0; a < 3; a++) g[a] += L[0] * kA * dA[a] + L[4] * kV * dV[a] + L[1 + a] * inv_nn;
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++) part[a * NJ + idx] = g[a];
}

// lamX(m) += the node's slots, in slot order.
__global__ void k_memb_adj_sum(const double *__restrict__ part, size_t NT, int n, int J,
                               const int32_t *__restrict__ inc_start, int nn,
                               const int32_t *__restrict__ status, double *__restrict__ lamX) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n || status[q / nn] < 0) return;
  const int m = q % nn, cnt = inc_start[m + 1] - inc_start[m];
  const size_t NJ = NT * J;
  for (int a = 0; a < 3; a++) {
    double s = 0;
    for (int j = 0; j <= cnt; j++) s += part[a * NJ + (size_t)q * J + j];
    lamX[a * NT + q] += s;
  }
}

// Modulus-scale gradients: sum over live cells in cell order.
__global__ void k_memb_param_sum(const double *__restrict__ cterm, int ncell,
                                 const int32_t *__restrict__ status, double *__restrict__ g) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0, b = 0;
  for (int c = 0; c < ncell; c++)
    if (status[c] >= 0) {
      a += cterm[8 * c + 5];
      b += cterm[8 * c + 6];
    }
  g[0] += a;
  g[1] += b;
}

// G = gamma_p v_rep + F_Total, the node forces spreading received.
__global__ void k_make_g(const double *__restrict__ vrep, const double *__restrict__ F, size_t n3,
                         double gamma_p, double *__restrict__ G) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) G[i] = gamma_p * vrep[i] + F[i];
}

__global__ void k_scale(const double *__restrict__ a, size_t n, double s, double *__restrict__ b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = s * a[i];
}

__global__ void k_zero(double *__restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

__global__ void k_aos_soa(const double *__restrict__ h, int n, size_t NT, double *__restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; a++) A[a * NT + i] = h[3 * (size_t)i + a];
}
__global__ void k_soa_aos(const double *__restrict__ A, int n, size_t NT, double *__restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; a++) h[3 * (size_t)i + a] = A[a * NT + i];
}

}  // namespace

// Device-level membrane transpose at positions X (SoA over the cell
// capacity).  lamX += (dF/dX)^T lam; gparam[0..1] += dJ/d ln(es), dJ/d ln(kbend).
// scratch: 7*ncap*nt (geometry) + 8*ncap (cell scalars) + 7*ncap*nn
// (node terms) + 8*ncap (cell terms) doubles.
size_t membrane_vjp_scratch(const zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells;
  return 8 * (7 * ncap * ctx->nt + 16 * ncap + 7 * ncap * ctx->nn * (ctx->inc_max + 1));
}

int membrane_vjp_device(zf_ctx *ctx, const double *X, const double *lam, double *lamX, double *gparam,
                        double *scratch) {
  const int nc = ctx->ncell;
  if (nc == 0) return ZF_OK;
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn, TT = ncap * ctx->nt;
  double *tg = scratch, *cell = tg + 7 * TT, *cterm = cell + 8 * ncap, *part = cterm + 8 * ncap;
  int rc = membrane_geometry(ctx, X, tg, cell);
  if (rc) return rc;
  const MembParams mp = params_of(ctx->mcfg);
  const MeshTables mt = mesh_tables(ctx);
  const int n = nc * ctx->nn, J = ctx->inc_max + 1;
  const size_t lanes = (size_t)n * J;
  hipLaunchKernelGGL(k_memb_cell_terms, dim3(nb(lanes, 128)), dim3(128), 0, ctx->stream, X, NT, nc, mt,
                     J, tg, TT, cell, ctx->d_status, mp, lam, part);
  hipLaunchKernelGGL(k_memb_cell_sum, dim3(nc), dim3(256), 0, ctx->stream, part, NT * J,
                     ctx->nn * J, ctx->d_status, cterm);
  hipLaunchKernelGGL(k_memb_adj_pos, dim3(nb(lanes, 128)), dim3(128), 0, ctx->stream, X, NT, nc, mt, J,
                     cell, cterm, ctx->d_status, mp, lam, part);
  hipLaunchKernelGGL(k_memb_adj_sum, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, part, NT, n, J,
                     ctx->d_inc_start, ctx->nn, ctx->d_status, lamX);
  if (gparam)
    hipLaunchKernelGGL(k_memb_param_sum, dim3(1), dim3(64), 0, ctx->stream, cterm, nc, ctx->d_status,
                       gparam);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

void free_adjoint_cells(zf_ctx *ctx) {
  void **bufs[] = {(void **)&ctx->d_rec_x, (void **)&ctx->d_rec_g, (void **)&ctx->d_rec_field,
                   (void **)&ctx->d_lamX,  (void **)&ctx->d_lamV,  (void **)&ctx->d_lamG,
                   (void **)&ctx->d_mscratch, (void **)&ctx->d_cgrad};
  for (auto b : bufs) dev_free(ctx, b);
  ctx->adj_cells = false;
  ctx->adj_rec_cap = ctx->adj_nrec = 0;
}

int adjoint_cells_alloc(zf_ctx *ctx) {
  if (!ctx->have_ibm || !ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "cells need zf_ibm_setup");
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const int cap = ctx->acfg.max_steps / ctx->icfg.rbc_interval + 1;
  int rc;
  if ((rc = ensure_force_field(ctx)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_rec_x, 24 * NT * cap)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_rec_g, 24 * NT * cap)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_rec_field, 3 * ctx->stride * field_elem(ctx) * (cap + 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_lamX, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_lamV, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_lamG, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_mscratch, membrane_vjp_scratch(ctx))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_cgrad, 16)))
    return rc;
  ctx->adj_rec_cap = cap;
  ctx->adj_cells = true;
  return ZF_OK;
}

int adjoint_cells_record(zf_ctx *ctx, int j) {
  if (j < 0 || j >= ctx->adj_rec_cap) return fail(ctx, ZF_ERR_STATE, "RBC record %d out of range", j);
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  ZF_HIP(ctx, hipMemcpyAsync(ctx->d_rec_x + 3 * NT * j, ctx->d_x, 24 * NT, hipMemcpyDeviceToDevice,
                             ctx->stream));
  hipLaunchKernelGGL(k_make_g, dim3(nb(3 * NT, 256)), dim3(256), 0, ctx->stream, ctx->d_vrep,
                     ctx->d_mforce, 3 * NT, ctx->icfg.gamma_p, ctx->d_rec_g + 3 * NT * j);
  const size_t fb = 3 * ctx->stride * field_elem(ctx);
  ZF_HIP(ctx, hipMemcpyAsync(static_cast<char *>(ctx->d_rec_field) + fb * (j + 1), ctx->d_force, fb,
                             hipMemcpyDeviceToDevice, ctx->stream));
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int adjoint_cells_backward(zf_ctx *ctx, int j, const void *f, void *lam) {
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const double *X = ctx->d_rec_x + 3 * NT * j, *G = ctx->d_rec_g + 3 * NT * j;
  // x' = x + dt V, V = interpolation + v_rep (v_rep held fixed)
  hipLaunchKernelGGL(k_scale, dim3(nb(3 * NT, 256)), dim3(256), 0, ctx->stream, ctx->d_lamX, 3 * NT,
                     ctx->icfg.dt_advance, ctx->d_lamV);
  int rc = ibm_vjp_device(ctx, X, f, G, ctx->d_lamV, ctx->d_gF, ctx->d_lamG, ctx->d_lamX, lam,
                          ctx->cfg.storage == ZF_STORE_F64);
  if (rc) return rc;
  // G = gamma_p v_rep + F_Total(X): the membrane part
  if ((rc = membrane_vjp_device(ctx, X, ctx->d_lamG, ctx->d_lamX, ctx->d_cgrad, ctx->d_mscratch)))
    return rc;
  hipLaunchKernelGGL(k_zero, dim3(nb(3 * ctx->stride, 256)), dim3(256), 0, ctx->stream, ctx->d_gF,
                     3 * ctx->stride);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

}  // namespace zf

using namespace zf;

extern "C" {

int zf_adjoint_cell_gradients(zf_ctx *ctx, double dparam[2], double *dJ_dx0) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_adj || !ctx->adj_cells) return fail(ctx, ZF_ERR_STATE, "no adjoint with cells set up");
  if (!ctx->adj_ready) return fail(ctx, ZF_ERR_STATE, "zf_adjoint_forward not run");
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  double g[2] = {0, 0};
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ZF_HIP(ctx, hipMemcpy(g, ctx->d_cgrad, 16, hipMemcpyDeviceToHost));
  if (dparam) {
    dparam[0] = g[0];
    dparam[1] = g[1];
  }
  if (dJ_dx0 && n > 0) {
    // lamV is free between backward passes
    hipLaunchKernelGGL(k_soa_aos, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_lamX, n, NT,
                       ctx->d_lamV);
    ZF_HIP(ctx, hipGetLastError());
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ZF_HIP(ctx, hipMemcpy(dJ_dx0, ctx->d_lamV, 24 * (size_t)n, hipMemcpyDeviceToHost));
  }
  return ZF_OK;
}

int zf_membrane_vjp(zf_ctx *ctx, const double *lam_force, double *lam_pos, double dparam[2]) {
  if (!ctx || !lam_force || !lam_pos) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const size_t sbytes = membrane_vjp_scratch(ctx);
  const size_t bytes = sbytes + 8 * (6 * NT + 3 * (size_t)n + 2);
  char *buf = nullptr;
  int rc = dev_alloc(ctx, (void **)&buf, bytes);
  if (rc) return rc;
  double *scratch = (double *)buf, *lam = (double *)(buf + sbytes), *lamX = lam + 3 * NT,
         *stage = lamX + 3 * NT, *gp = stage + 3 * (size_t)n;
  hipError_t e = hipMemsetAsync(lam, 0, 8 * (6 * NT + 3 * (size_t)n + 2), ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(stage, lam_force, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_aos_soa, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, stage, n, NT, lam);
    e = hipGetLastError();
  }
  if (e == hipSuccess && (rc = membrane_vjp_device(ctx, ctx->d_x, lam, lamX, gp, scratch))) {
    ctx->device_bytes -= (int64_t)bytes;
    dev_free(ctx, (void **)&buf);
    return rc;
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_soa_aos, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, lamX, n, NT, stage);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(lam_pos, stage, 24 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  double g[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(g, gp, 16, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= (int64_t)bytes;
  dev_free(ctx, (void **)&buf);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "membrane vjp: %s", hipGetErrorString(e));
  if (dparam) {
    dparam[0] = g[0];
    dparam[1] = g[1];
  }
  return ZF_OK;
}

}  // extern "C"
