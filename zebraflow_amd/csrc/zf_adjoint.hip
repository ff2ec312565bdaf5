// zf_adjoint.hip -- hand-written adjoint of the LBM time step with a
// wall-shear-stress objective, with checkpointing.
//
// Forward iteration k maps the populations S_{k-1} to S_k
// (ZebraflowWT.c:6101-6145 stream + collide, :6292-6433 boundary values,
// the kernel of zf_lbm.hip).  With lambda_k = dJ/dS_k:
//   lambda_{k-1} = A_k^T lambda_k + dJ/dS_{k-1}
// and every parameter theta of iteration k collects lambda_k . dS_k/dtheta.
//
// A_k^T, one lane per interior node s (k_adjoint): the forward pulls
// g_c(s) = S(src(s,c), c) and writes S_k(s, :) = C(g(s)) (BGK + Guo); the
// pull map is one-to-one, so the transpose is a push: the lane forms
// mu = dC/dg^T lambda_k(s, :) and writes mu_c to lambda_{k-1}(src(s,c), c),
// each target written by exactly one lane.  Boundary nodes whose BCNLIST
// neighbour is s take their values from s's moments; their lambda_k folds
// into s's moment adjoint before mu is formed.  Boundary nodes fed by a
// non-interior neighbour ("late" entries) add to that neighbour's lambda in
// a second, sequential pass (k_adjoint_late).  Entries that the adjoint
// reads but no lane pulls are cleared (k_zero_idx).
//
// Collision transpose.  out_c = (1-w) g_c + w geq_c(drho, u) + p S_c(u, F),
// w = 1/tau, p = DT (1 - w/2), geq_c = w_c (drho + (1+drho) X_c),
// X_c = 3 c.u + 4.5 (c.u)^2 - 1.5 u^2, S_c = w_c (3 (c-u).F + 9 (c.u)(c.F)),
// u = v/(1+drho), v = j + DT F/2.  With l = lambda_k(s, :):
//   a_drho = w sum_c l_c w_c (1 + X_c)
//   a_u    = sum_c l_c w_c [w (1+drho) ((3 + 9 c.u) c - 3 u) + p (9 (c.F) c - 3 F)]
//   a_F    = p sum_c l_c w_c (3 (c - u) + 9 (c.u) c)  +  DT/2 a_u/(1+drho)
//   mu_c   = (1-w) l_c + a_drho - a_u.u/(1+drho) + a_u.c/(1+drho)
//
// Objective: mean over the window and the wall points of |tau|, tau the
// zf_wss shear stress; its gradient reaches the populations of each
// stencil node through u = j/(1+drho): dJ/dg_c = dJ/du . (c - u)/(1+drho).
#include "zf_wss_dev.h"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

namespace zf {
namespace {

using namespace d3q19;
using namespace wssdev;

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// arithmetic and storage type of the adjoint populations
template <int S>
struct Adj {
  using A = double;
};
template <>
struct Adj<ZF_STORE_F32DEV> {
  using A = float;
};

struct AdjParams {
  double u0, amp, sin_phase;  // inlet velocity u0 (1 + amp sin_phase) dir
  double dir[3];
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Adds the workgroup's sums of five per-lane values (256 lanes) to its own
// row part[5 * blockIdx.x ..], over all adjoint iterations; k_reduce_grads
// adds the rows in row order at the end, so the gradients do not depend on
// scheduling.
__device__ __forceinline__ void block_sum(double r[5], double *__restrict__ part) {
  __shared__ double wp[4][5];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = 0; i < 5; i++) {
    const double v = wave_sum(r[i]);
    if (lane == 0) wp[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < 5)
    part[5 * (size_t)blockIdx.x + threadIdx.x] +=
        wp[0][threadIdx.x] + wp[1][threadIdx.x] + wp[2][threadIdx.x] + wp[3][threadIdx.x];
}

// grads[i] = sum over the n partial rows, in row order (one lane per i).
__global__ void k_reduce_grads(const double *__restrict__ part, int n, double *__restrict__ grads) {
  const int i = threadIdx.x;
  if (i >= 5) return;
  double acc = 0;
  for (int r = 0; r < n; r++) acc += part[5 * (size_t)r + i];
  grads[i] = acc;
}

// Storage index the forward pulls g_c(s) from (the adjoint's push target),
// 32-bit: the set-up checks that 19 * stride fits, so the loads and stores
// take a scalar base plus one vector offset register.
__device__ __forceinline__ uint32_t pull_index(int c, uint32_t w, const int d[20], int s,
                                               uint32_t stride) {
  if (c == 18) return 18u * stride + (uint32_t)s;
  return ((w >> c) & 1u) ? (uint32_t)(c ^ 1) * stride + (uint32_t)s
                         : (uint32_t)c * stride + (uint32_t)(s + d[c]);
}

template <int S, bool FIELD>
__global__ __launch_bounds__(256, S == ZF_STORE_F64 ? 1 : 3) void k_adjoint(
    const typename Store<S>::T *__restrict__ src, const typename Adj<S>::A *__restrict__ lin,
    typename Adj<S>::A *__restrict__ lout, uint32_t stride, int n_int, int nblk, int remap,
    const uint32_t *__restrict__ word, const int32_t *__restrict__ block0,
    const int32_t *__restrict__ chords, const typename Store<S>::F *__restrict__ force, StepParams prm,
    const BcEntry *__restrict__ bc, const int32_t *__restrict__ fused_s,
    const int32_t *__restrict__ att_start, int n_fused, BcParams bp, AdjParams ap,
    const int64_t *__restrict__ unpulled, int64_t n_unpulled, double *__restrict__ gF,
    double *__restrict__ part) {
  using R = typename Store<S>::R;
  using A = typename Adj<S>::A;
  // entries read by the adjoint that no lane pushes to: cleared here, each
  // workgroup one slice of the list (no lane of this launch writes them)
  {
    const int64_t chunk = (n_unpulled + nblk - 1) / nblk;
    const int64_t e1 = min(n_unpulled, ((int64_t)blockIdx.x + 1) * chunk);
    for (int64_t e = (int64_t)blockIdx.x * chunk + threadIdx.x; e < e1; e += 256) lout[unpulled[e]] = (A)0;
  }
  const int s = logical_block(blockIdx.x, nblk, remap != 0) * 256 + threadIdx.x;
  double red[5] = {0, 0, 0, 0, 0};
  if (s < n_int) {
    const uint32_t w = word[s];
    const int ch = block0[s >> 6] + (int)(w >> kWordChordShift);
    const int4 *cr = reinterpret_cast<const int4 *>(chords + (size_t)ch * kChordInts);
    int d[20];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const int4 x = cr[v];
      d[4 * v] = x.x; d[4 * v + 1] = x.y; d[4 * v + 2] = x.z; d[4 * v + 3] = x.w;
    }
    // moments of the pulled populations (the forward's, not kept)
    R drho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int c = 0; c < Q; c++) {
      const R gc = Store<S>::ld(src, pull_index(c, w, d, s, stride), WQ[c]);
      drho += gc;
      jx += (R)EX[c] * gc;
      jy += (R)EY[c] * gc;
      jz += (R)EZ[c] * gc;
    }
    R fx = (R)prm.g[0], fy = (R)prm.g[1], fz = (R)prm.g[2];
    if (FIELD) {
      fx += (R)force[s];
      fy += (R)force[stride + s];
      fz += (R)force[2 * stride + s];
    }
    const R dt = (R)prm.dt;
    const R rho = (R)1 + drho;
    const R q = drho / rho;
    const R vx = jx + (R)0.5 * dt * fx, vy = jy + (R)0.5 * dt * fy, vz = jz + (R)0.5 * dt * fz;
    const R ux = vx - vx * q, uy = vy - vy * q, uz = vz - vz * q;
    const R usq = ux * ux + uy * uy + uz * uz;
    const R om = (R)1 / ((R)0.5 + (R)prm.visc_scale * (R)d[18]);
    const R fp = dt * ((R)1 - (R)0.5 * om);
    // lambda_k(s, :) is streamed twice (sums here, mu below; the second
    // read hits the cache) rather than held in 19 registers
    R add = 0, aux = 0, auy = 0, auz = 0, afx = 0, afy = 0, afz = 0;
#pragma unroll
    for (int c = 0; c < Q; c++) {
      const R ex = (R)EX[c], ey = (R)EY[c], ez = (R)EZ[c];
      const R cu = ex * ux + ey * uy + ez * uz;
      const R cf = ex * fx + ey * fy + ez * fz;
      const R X = (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usq;
      const R wl = wmul<R>(c, (R)lin[(uint32_t)c * stride + (uint32_t)s]);
      add += om * wl * ((R)1 + X);
      const R k1 = om * wl * rho, t1 = (R)3 + (R)9 * cu;
      const R k2 = fp * wl, t2 = (R)9 * cf;
      aux += k1 * (t1 * ex - (R)3 * ux) + k2 * (t2 * ex - (R)3 * fx);
      auy += k1 * (t1 * ey - (R)3 * uy) + k2 * (t2 * ey - (R)3 * fy);
      auz += k1 * (t1 * ez - (R)3 * uz) + k2 * (t2 * ez - (R)3 * fz);
      const R t3 = (R)9 * cu;
      afx += k2 * ((R)3 * (ex - ux) + t3 * ex);
      afy += k2 * ((R)3 * (ey - uy) + t3 * ey);
      afz += k2 * ((R)3 * (ez - uz) + t3 * ez);
    }
    R gix = 0, giy = 0, giz = 0;  // d(out)/d(inlet velocity) . lambda
    if (w & kWordMacro) {
      int lo = 0, hi = n_fused - 1, k = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1, v = fused_s[mid];
        if (v == s) { k = mid; break; }
        if (v < s) lo = mid + 1; else hi = mid - 1;
      }
      if (k >= 0)
        for (int t = att_start[k]; t < att_start[k + 1]; t++) {
          const BcEntry e = bc[t];
          const bool wall = e.code == 0 || e.code == 4998;
          const bool inlet = e.code == ZF_B_VELOCITY_INLET;
          R drb = 0;
          if (e.code >= 1001 && e.code <= 1009) drb = (R)bp.pdrho[e.code - 1001];
          else if (e.code == ZF_B_PRESSURE_OUTLET) drb = (R)bp.outlet_drho;
          const R ubx = inlet ? (R)bp.inlet[0] : ux, uby = inlet ? (R)bp.inlet[1] : uy,
                  ubz = inlet ? (R)bp.inlet[2] : uz;
          const R ubsq = ubx * ubx + uby * uby + ubz * ubz;
#pragma nounroll
          for (int c = 0; c < Q; c++) {
            const R wl = wmul<R>(c, (R)lin[(size_t)c * stride + e.s]);
            const R ex = (R)EX[c], ey = (R)EY[c], ez = (R)EZ[c];
            if (wall) {
              add += wl;
            } else if (inlet) {
              const R cu = ex * ubx + ey * uby + ez * ubz;
              add += wl * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * ubsq);
              const R k1 = wl * rho, t1 = (R)3 + (R)9 * cu;
              gix += k1 * (t1 * ex - (R)3 * ubx);
              giy += k1 * (t1 * ey - (R)3 * uby);
              giz += k1 * (t1 * ez - (R)3 * ubz);
            } else {
              const R cu = ex * ux + ey * uy + ez * uz;
              const R k1 = wl * ((R)1 + drb), t1 = (R)3 + (R)9 * cu;
              aux += k1 * (t1 * ex - (R)3 * ux);
              auy += k1 * (t1 * ey - (R)3 * uy);
              auz += k1 * (t1 * ez - (R)3 * uz);
            }
          }
        }
    }
    // through u = v / (1 + drho)
    const R avx = aux / rho, avy = auy / rho, avz = auz / rho;
    const R adt = add - (avx * ux + avy * uy + avz * uz);
    afx += (R)0.5 * dt * avx;
    afy += (R)0.5 * dt * avy;
    afz += (R)0.5 * dt * avz;
#pragma unroll
    for (int c = 0; c < Q; c++) {
      const R lc = (R)lin[(uint32_t)c * stride + (uint32_t)s];
      const R mu = ((R)1 - om) * lc + adt + (R)EX[c] * avx + (R)EY[c] * avy + (R)EZ[c] * avz;
      lout[pull_index(c, w, d, s, stride)] = (A)mu;
    }
    if (FIELD && gF) {
      gF[s] += (double)afx;
      gF[stride + s] += (double)afy;
      gF[2 * stride + s] += (double)afz;
    }
    const double gdir = (double)gix * ap.dir[0] + (double)giy * ap.dir[1] + (double)giz * ap.dir[2];
    red[0] = gdir * ap.u0 * ap.sin_phase;
    red[1] = gdir * (1.0 + ap.amp * ap.sin_phase);
    red[2] = afx;
    red[3] = afy;
    red[4] = afz;
  }
  block_sum(red, part);
}

// Boundary entries fed by a non-interior neighbour, one entry after the
// other (entries may share a neighbour), one lane per direction.  The
// forward took the neighbour's unshifted moments from S_{k-1}.
template <int S>
__global__ __launch_bounds__(64) void k_adjoint_late(
    const typename Store<S>::T *__restrict__ src, const typename Adj<S>::A *__restrict__ lin,
    typename Adj<S>::A *__restrict__ lout, size_t stride, const BcEntry *__restrict__ bc, int t0,
    int t1, BcParams bp, AdjParams ap, double *__restrict__ part_row) {
  using A = typename Adj<S>::A;
  const int c = threadIdx.x;
  const bool on = c < Q;
  const double ex = on ? EX[c] : 0, ey = on ? EY[c] : 0, ez = on ? EZ[c] : 0, wc = on ? WQ[c] : 0;
  double gacc = 0, gacc0 = 0;
  for (int t = t0; t < t1; t++) {
    const BcEntry e = bc[t];
    const double gv = on ? (double)Store<S>::ld(src, (size_t)c * stride + e.sm, wc) : 0.0;
    const double drho = wave_sum(gv);
    const double rm = 1.0 + drho;
    const double ux = wave_sum(ex * gv) / rm, uy = wave_sum(ey * gv) / rm, uz = wave_sum(ez * gv) / rm;
    const double l = on ? (double)lin[(size_t)c * stride + e.s] : 0.0;
    const double wl = wc * l;
    double cdd = 0, cux = 0, cuy = 0, cuz = 0, cgx = 0, cgy = 0, cgz = 0;
    if (e.code == 0 || e.code == 4998) {
      cdd = wl;
    } else if (e.code == ZF_B_VELOCITY_INLET) {
      const double ubx = bp.inlet[0], uby = bp.inlet[1], ubz = bp.inlet[2];
      const double cu = ex * ubx + ey * uby + ez * ubz;
      cdd = wl * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (ubx * ubx + uby * uby + ubz * ubz));
      const double k1 = wl * rm, t1v = 3.0 + 9.0 * cu;
      cgx = k1 * (t1v * ex - 3.0 * ubx);
      cgy = k1 * (t1v * ey - 3.0 * uby);
      cgz = k1 * (t1v * ez - 3.0 * ubz);
    } else {
      const double drb = (e.code >= 1001 && e.code <= 1009) ? bp.pdrho[e.code - 1001] : bp.outlet_drho;
      const double cu = ex * ux + ey * uy + ez * uz;
      const double k1 = wl * (1.0 + drb), t1v = 3.0 + 9.0 * cu;
      cux = k1 * (t1v * ex - 3.0 * ux);
      cuy = k1 * (t1v * ey - 3.0 * uy);
      cuz = k1 * (t1v * ez - 3.0 * uz);
    }
    const double add = wave_sum(cdd);
    const double aux = wave_sum(cux), auy = wave_sum(cuy), auz = wave_sum(cuz);
    const double gx = wave_sum(cgx), gy = wave_sum(cgy), gz = wave_sum(cgz);
    if (on) {
      const double v = add + (aux * (ex - ux) + auy * (ey - uy) + auz * (ez - uz)) / rm;
      const size_t o = (size_t)c * stride + e.sm;
      lout[o] = (A)((double)lout[o] + v);
    }
    const double gdir = gx * ap.dir[0] + gy * ap.dir[1] + gz * ap.dir[2];
    gacc += gdir * ap.u0 * ap.sin_phase;
    gacc0 += gdir * (1.0 + ap.amp * ap.sin_phase);
  }
  if (c == 0) {
    part_row[0] += gacc;
    part_row[1] += gacc0;
  }
}

template <class A>
__global__ void k_zero_idx(A *__restrict__ lam, const int64_t *__restrict__ idx, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lam[idx[i]] = (A)0;
}

template <class A>
__global__ void k_fill_a(A *__restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (A)0;
}

// Marks every (c, s) an interior node pulls.
__global__ void k_mark_pulled(size_t stride, int n_int, const uint32_t *__restrict__ word,
                              const int32_t *__restrict__ block0, const int32_t *__restrict__ chords,
                              uint8_t *__restrict__ mark) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_int) return;
  const uint32_t w = word[s];
  const int ch = block0[s >> 6] + (int)(w >> kWordChordShift);
  const int32_t *d = chords + (size_t)ch * kChordInts;
  for (int c = 0; c < 18; c++) {
    const bool bb = (w >> c) & 1u;
    mark[bb ? (size_t)(c ^ 1) * stride + s : (size_t)c * stride + (s + d[c])] = 1;
  }
  mark[(size_t)18 * stride + s] = 1;
}

// Stencil of every probe: up to kStencilMax (storage index, weight) pairs
// per probe, count in cnt[m].
constexpr int kStencilMax = 64;
__global__ void k_stencil_table(const int32_t *__restrict__ b, const int32_t *__restrict__ sbox,
                                WGrid g, int n, const double *__restrict__ pts,
                                int32_t *__restrict__ cnt, int32_t *__restrict__ sidx,
                                double *__restrict__ wgt) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double *x = pts + 3 * (size_t)m, *nv = pts + 3 * (size_t)n + 3 * (size_t)m;
  int k = 0;
  wss_stencil(g, b, sbox, x[0] + g.dx * nv[0], x[1] + g.dx * nv[1], x[2] + g.dx * nv[2],
              [&](int sq, double w) {
                if (k < kStencilMax) {
                  sidx[(size_t)m * kStencilMax + k] = sq;
                  wgt[(size_t)m * kStencilMax + k] = w;
                }
                k++;
              });
  cnt[m] = k;
}

// Unshifted velocity and density of every support node (stencil nodes of
// the probes): su[4 i ..] = (u_x, u_y, u_z, 1 + drho).  Batched over states
// (blockIdx.y = b): state r0 + b is f0 for r0 + b == 0, else the
// (r0 + b - 1)-th state of the array `fr` (state_elems each); its moments go
// to su + 4 nlist b.
template <int S>
__global__ void k_support_moments(const typename Store<S>::T *__restrict__ f0,
                                  const typename Store<S>::T *__restrict__ fr, size_t state_elems,
                                  int r0, size_t stride, const int32_t *__restrict__ list,
                                  int nlist, double *__restrict__ su) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlist) return;
  const int rs = r0 + (int)blockIdx.y;
  const typename Store<S>::T *f = rs == 0 ? f0 : fr + (size_t)(rs - 1) * state_elems;
  su += 4 * (size_t)nlist * blockIdx.y;
  double drho, jv[3];
  node_moments<S>(f, stride, list[i], drho, jv);
  const double r = 1.0 + drho;
  su[4 * (size_t)i] = jv[0] / r;
  su[4 * (size_t)i + 1] = jv[1] / r;
  su[4 * (size_t)i + 2] = jv[2] / r;
  su[4 * (size_t)i + 3] = r;
}

// Wall shear stress of each probe from its stencil table (support index,
// weight); one 64-lane wave per probe, lane e taking entries e, e+64, ...
// and a fixed-order wave reduction.  grad == 0: out[m] = |tau|.
// grad != 0: out[3m..] = a_m = -(mu/DX) dJ/dtau with J gaining scale |tau|,
// so dJ/du at stencil node q is w_q a_m (dJ/dtau is tangential, so the
// projection I - n n leaves it unchanged).
__global__ __launch_bounds__(256) void k_probe_tau(int n, const int32_t *__restrict__ off,
                                                   const int32_t *__restrict__ sup,
                                                   const double *__restrict__ wgt,
                                                   const double *__restrict__ su,
                                                   const double *__restrict__ pts, double mu_over_dx,
                                                   double scale, int grad, int nsup,
                                                   double *__restrict__ out) {
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (m >= n) return;  // whole waves leave together
  // batch b = blockIdx.y: support moments of state b, outputs of state b
  su += 4 * (size_t)nsup * blockIdx.y;
  out += (grad ? 3 : 1) * (size_t)n * blockIdx.y;
  double u0 = 0, u1 = 0, u2 = 0;
  for (int e = off[m] + lane; e < off[m + 1]; e += 64) {
    const double w = wgt[e];
    const double *v = su + 4 * (size_t)sup[e];
    u0 += w * v[0];
    u1 += w * v[1];
    u2 += w * v[2];
  }
  u0 = wave_sum(u0);
  u1 = wave_sum(u1);
  u2 = wave_sum(u2);
  if (lane != 0) return;
  const double *nv = pts + 3 * (size_t)n + 3 * (size_t)m;
  const double un = u0 * nv[0] + u1 * nv[1] + u2 * nv[2];
  const double t0 = -mu_over_dx * (u0 - un * nv[0]), t1 = -mu_over_dx * (u1 - un * nv[1]),
               t2 = -mu_over_dx * (u2 - un * nv[2]);
  const double mag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
  if (!grad) {
    out[m] = mag;
    return;
  }
  const double k = mag > 0 ? -mu_over_dx * scale / mag : 0.0;
  out[3 * (size_t)m] = k * t0;
  out[3 * (size_t)m + 1] = k * t1;
  out[3 * (size_t)m + 2] = k * t2;
}

// dJ/du at each support node i: the sum over the probes whose stencil holds
// the node of w a_m; 8 lanes per node take entries e, e+8, ... and a
// fixed-order shuffle reduction finishes it; sg[3 i ..].
__global__ __launch_bounds__(256) void k_obj_gather(int nlist, const int32_t *__restrict__ off,
                                                    const int32_t *__restrict__ probe,
                                                    const double *__restrict__ weight,
                                                    const double *__restrict__ a, int npts,
                                                    double *__restrict__ sg) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t >> 3, sub = t & 7;
  a += 3 * (size_t)npts * blockIdx.y;  // batch b = blockIdx.y
  sg += 3 * (size_t)nlist * blockIdx.y;
  double d0 = 0, d1 = 0, d2 = 0;
  if (i < nlist)
    for (int e = off[i] + sub; e < off[i + 1]; e += 8) {
      const double w = weight[e];
      const double *am = a + 3 * (size_t)probe[e];
      d0 += w * am[0];
      d1 += w * am[1];
      d2 += w * am[2];
    }
  for (int o = 4; o > 0; o >>= 1) {
    d0 += __shfl_xor(d0, o);
    d1 += __shfl_xor(d1, o);
    d2 += __shfl_xor(d2, o);
  }
  if (i < nlist && sub == 0) {
    sg[3 * (size_t)i] = d0;
    sg[3 * (size_t)i + 1] = d1;
    sg[3 * (size_t)i + 2] = d2;
  }
}

// lambda(s_i, c) += dJ/du_i . (c - u_i)/(1 + drho_i), one lane per
// (direction c, support node i), consecutive lanes on consecutive nodes.
template <int S>
__global__ void k_obj_apply(const int32_t *__restrict__ list, int nlist, size_t stride,
                            const double *__restrict__ sg, const double *__restrict__ su,
                            typename Adj<S>::A *__restrict__ lam) {
  using A = typename Adj<S>::A;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Q * nlist) return;
  const int c = t / nlist, i = t - c * nlist;
  const double *d = sg + 3 * (size_t)i, *v = su + 4 * (size_t)i;
  const double g = (d[0] * (EX[c] - v[0]) + d[1] * (EY[c] - v[1]) + d[2] * (EZ[c] - v[2])) / v[3];
  const size_t o = (size_t)c * stride + list[i];
  lam[o] = (A)((double)lam[o] + g);
}

// lambda (storage order) -> host layout, in double.
template <class A>
__global__ void k_export_lam(const A *__restrict__ lam, size_t stride, int n,
                             const int32_t *__restrict__ s_of_p, double *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = s_of_p[p];
  for (int c = 0; c < Q; c++) out[(size_t)c * n + p] = (double)lam[(size_t)c * stride + s];
}

template <int S>
const typename Store<S>::T *st(const void *p) {
  return static_cast<const typename Store<S>::T *>(p);
}

WGrid wgrid(const zf_ctx *ctx) {
  const GridGeom gg = grid_geom(ctx);
  return WGrid{gg.x0, gg.y0, gg.z0, gg.dx, ctx->cfg.im, ctx->cfg.jm, ctx->cfg.km};
}

template <class F>
int with_store(int storage, F &&f) {
  if (storage == ZF_STORE_F64) return f(std::integral_constant<int, ZF_STORE_F64>());
  return f(std::integral_constant<int, ZF_STORE_F32DEV>());
}

void *state_at(zf_ctx *ctx, void *base, size_t i) {
  return static_cast<char *>(base) + i * ctx->adj_state_bytes;
}

int require_adj(zf_ctx *ctx) {
  if (!ctx->have_adj) return fail(ctx, ZF_ERR_STATE, "zf_adjoint_setup not called");
  return ZF_OK;
}

// Support moments of nb states: r0 + b == 0 is f0, else the (r0 + b - 1)-th
// state of fr; batch slot b of d_sup_u.
int support_moments(zf_ctx *ctx, const void *f0, const void *fr, int r0, int nbatch) {
  if (ctx->n_support == 0) return ZF_OK;
  with_store(ctx->cfg.storage, [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    hipLaunchKernelGGL(k_support_moments<S>, dim3(nb(ctx->n_support, 256), nbatch), dim3(256), 0,
                       ctx->stream, st<S>(f0), st<S>(fr), (size_t)Q * ctx->stride, r0, ctx->stride,
                       ctx->d_support, ctx->n_support, ctx->d_sup_u);
    return 0;
  });
  return ZF_OK;
}

// Objective terms |tau| of the current state, stored for iteration `it`.
int obj_forward(zf_ctx *ctx, const void *f, int it) {
  const zf_adjoint_config &a = ctx->acfg;
  double *out = ctx->d_objbuf + (size_t)(it - a.t_begin) * ctx->adj_npts;
  support_moments(ctx, f, f, 0, 1);
  hipLaunchKernelGGL(k_probe_tau, dim3(nb(ctx->adj_npts, 4)), dim3(256), 0, ctx->stream,
                     ctx->adj_npts, ctx->d_prb_off, ctx->d_prb_sup, ctx->d_prb_w, ctx->d_sup_u,
                     ctx->d_adj_pts, a.mu_over_dx, 1.0, 0, ctx->n_support, out);
  return ZF_OK;
}

// dJ/du at the support nodes for nb states (see support_moments), in batch
// slots 0..nb-1 of d_sup_g (with their moments in d_sup_u).  The three
// passes read only the states, so the states of a checkpoint segment are
// processed in a few batched launches instead of three launches each.
int obj_prepare(zf_ctx *ctx, const void *f0, const void *fr, int r0, int nbatch) {
  const zf_adjoint_config &a = ctx->acfg;
  if (nbatch < 1 || nbatch > ctx->obj_batch) return fail(ctx, ZF_ERR_ARG, "objective batch");
  const double scale = 1.0 / ((double)(a.t_end - a.t_begin + 1) * ctx->adj_npts);
  support_moments(ctx, f0, fr, r0, nbatch);
  hipLaunchKernelGGL(k_probe_tau, dim3(nb(ctx->adj_npts, 4), nbatch), dim3(256), 0, ctx->stream,
                     ctx->adj_npts, ctx->d_prb_off, ctx->d_prb_sup, ctx->d_prb_w, ctx->d_sup_u,
                     ctx->d_adj_pts, a.mu_over_dx, scale, 1, ctx->n_support, ctx->d_adj_a);
  if (ctx->n_support > 0)
    hipLaunchKernelGGL(k_obj_gather, dim3(nb(8 * (size_t)ctx->n_support, 256), nbatch), dim3(256), 0,
                       ctx->stream, ctx->n_support, ctx->d_sup_off, ctx->d_sup_probe, ctx->d_sup_w,
                       ctx->d_adj_a, ctx->adj_npts, ctx->d_sup_g);
  return ZF_OK;
}

// lambda += dJ/dS for the state prepared in batch slot `slot`.
int obj_apply(zf_ctx *ctx, int slot, void *lam) {
  if (ctx->n_support == 0) return ZF_OK;
  const size_t ns = ctx->n_support;
  with_store(ctx->cfg.storage, [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    using A = typename Adj<S>::A;
    hipLaunchKernelGGL(k_obj_apply<S>, dim3(nb((size_t)Q * ns, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_support, ctx->n_support, ctx->stride, ctx->d_sup_g + 3 * ns * slot,
                       ctx->d_sup_u + 4 * ns * slot, static_cast<A *>(lam));
    return 0;
  });
  return ZF_OK;
}

// One adjoint iteration k: lout = A_k^T lin, given S_{k-1} = src.
int adjoint_step(zf_ctx *ctx, const void *src, const void *lin, void *lout, int k) {
  StepParams prm;
  BcParams bp;
  step_params(ctx, k, prm, bp);
  const zf_config &c = ctx->cfg;
  const double ph = 2.0 * 3.14159265358979323846 * k / (c.inlet_period > 0 ? c.inlet_period : 1.0);
  const AdjParams ap{c.inlet_u0, c.inlet_amp, std::sin(ph), {c.inlet_dir[0], c.inlet_dir[1], c.inlet_dir[2]}};
  const int nblk = (int)nb(ctx->n_int, 256);
  const int remap = (c.flags & ZF_FLAG_LINEAR_BLOCKS) ? 0 : 1;
  with_store(c.storage, [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    using A = typename Adj<S>::A;
    if (nblk > 0) {
      if (ctx->d_force)
        hipLaunchKernelGGL((k_adjoint<S, true>), dim3(nblk), dim3(256), 0, ctx->stream, st<S>(src),
                           static_cast<const A *>(lin), static_cast<A *>(lout), ctx->stride,
                           ctx->n_int, nblk, remap, ctx->d_word, ctx->d_block_chord0, ctx->d_chords,
                           static_cast<const typename Store<S>::F *>(ctx->d_force), prm, ctx->d_bc,
                           ctx->d_fused_s, ctx->d_att_start,
                           ctx->n_fused_nodes, bp, ap, ctx->d_unpulled, ctx->n_unpulled,
                           ctx->d_gF, ctx->d_gpart);
      else
        hipLaunchKernelGGL((k_adjoint<S, false>), dim3(nblk), dim3(256), 0, ctx->stream, st<S>(src),
                           static_cast<const A *>(lin), static_cast<A *>(lout), ctx->stride,
                           ctx->n_int, nblk, remap, ctx->d_word, ctx->d_block_chord0, ctx->d_chords,
                           nullptr, prm, ctx->d_bc, ctx->d_fused_s, ctx->d_att_start,
                           ctx->n_fused_nodes, bp, ap, ctx->d_unpulled, ctx->n_unpulled,
                           nullptr, ctx->d_gpart);
    }
    if (nblk == 0 && ctx->n_unpulled > 0)
      hipLaunchKernelGGL(k_zero_idx<A>, dim3(nb((size_t)ctx->n_unpulled, 256)), dim3(256), 0,
                         ctx->stream, static_cast<A *>(lout), ctx->d_unpulled, ctx->n_unpulled);
    if (ctx->n_bc > ctx->n_bc_fused)
      hipLaunchKernelGGL(k_adjoint_late<S>, dim3(1), dim3(64), 0, ctx->stream, st<S>(src),
                         static_cast<const A *>(lin), static_cast<A *>(lout), ctx->stride, ctx->d_bc,
                         ctx->n_bc_fused, ctx->n_bc, bp, ap, ctx->d_gpart + 5 * (size_t)nblk);
    return 0;
  });
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

}  // namespace

void free_adjoint(zf_ctx *ctx) {
  void **bufs[] = {(void **)&ctx->d_adj_pts, &ctx->d_ckpt, &ctx->d_seg, &ctx->d_lam[0],
                   &ctx->d_lam[1], (void **)&ctx->d_adj_a, (void **)&ctx->d_support,
                   (void **)&ctx->d_sup_off, (void **)&ctx->d_sup_probe, (void **)&ctx->d_sup_w,
                   (void **)&ctx->d_sup_u, (void **)&ctx->d_sup_g, (void **)&ctx->d_prb_off,
                   (void **)&ctx->d_prb_sup,
                   (void **)&ctx->d_prb_w,
                   (void **)&ctx->d_gpart,
                   (void **)&ctx->d_unpulled, (void **)&ctx->d_grads, (void **)&ctx->d_gF,
                   (void **)&ctx->d_objbuf};
  for (auto b : bufs) dev_free(ctx, b);
  free_adjoint_cells(ctx);
  ctx->have_adj = ctx->adj_ready = false;
  ctx->n_support = 0;
  ctx->n_unpulled = 0;
}

}  // namespace zf

using namespace zf;

extern "C" {

int zf_set_forcing(zf_ctx *ctx, double inlet_u0, double inlet_amp, const double body_force[3]) {
  if (!ctx || !body_force) return ZF_ERR_ARG;
  ctx->cfg.inlet_u0 = inlet_u0;
  ctx->cfg.inlet_amp = inlet_amp;
  for (int a = 0; a < 3; a++) ctx->cfg.body_force[a] = body_force[a];
  return ZF_OK;
}

int zf_adjoint_setup(zf_ctx *ctx, const zf_adjoint_config *ac, int64_t npts, const double *xyz,
                     const double *normal) {
  if (!ctx || !ac || npts <= 0 || !xyz || !normal) return ZF_ERR_ARG;
  if (ac->checkpoint_interval < 1 || ac->max_steps < 1 || ac->t_begin > ac->t_end || npts > (1 << 24))
    return fail(ctx, ZF_ERR_ARG, "bad adjoint configuration");
  if (!ctx->have_lattice || !ctx->d_f[0]) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  if (ctx->cfg.storage == ZF_STORE_F16DEV)
    return fail(ctx, ZF_ERR_ARG, "the adjoint supports FP64 and FP32 storage only");
  if (ctx->link_main || !ctx->linked.empty())
    return fail(ctx, ZF_ERR_STATE, "the adjoint does not cover parent-lattice coupling");
  if (ac->cells && ctx->have_nuc)
    return fail(ctx, ZF_ERR_STATE, "the adjoint does not cover the nucleus network");
  if ((uint64_t)Q * ctx->stride >= (1ull << 32))
    return fail(ctx, ZF_ERR_ARG, "lattice too large for the adjoint's 32-bit population offsets");
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  free_adjoint(ctx);
  ctx->acfg = *ac;
  ctx->adj_npts = (int32_t)npts;
  // objective batch: the states of up to 64 iterations of a segment
  ctx->obj_batch = std::min(64, ac->checkpoint_interval);
  const size_t stride = ctx->stride;
  const size_t eb = store_bytes(ctx->cfg.storage);
  const size_t ab = ctx->cfg.storage == ZF_STORE_F64 ? 8 : 4;
  ctx->adj_state_bytes = eb * Q * stride;
  const int K = ac->checkpoint_interval;
  const size_t nck = (size_t)(ac->max_steps + K - 1) / K;
  const size_t nobj = (size_t)(ac->t_end - ac->t_begin + 1) * npts;
  int rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_adj_pts, 48 * (size_t)npts)) ||
      (rc = dev_alloc(ctx, &ctx->d_ckpt, nck * ctx->adj_state_bytes)) ||
      (rc = dev_alloc(ctx, &ctx->d_seg, (size_t)std::max(K - 1, 1) * ctx->adj_state_bytes)) ||
      (rc = dev_alloc(ctx, &ctx->d_lam[0], ab * Q * stride)) ||
      (rc = dev_alloc(ctx, &ctx->d_lam[1], ab * Q * stride)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_adj_a, 24 * (size_t)npts * ctx->obj_batch)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_gpart, 40 * ((size_t)nb(ctx->n_int, 256) + 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_grads, 8 * 8)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_objbuf, 8 * nobj))) {
    free_adjoint(ctx);
    return rc;
  }
  if (ac->cells && (rc = adjoint_cells_alloc(ctx))) {
    free_adjoint(ctx);
    return rc;
  }
  if (ctx->d_force && (rc = dev_alloc(ctx, (void **)&ctx->d_gF, 24 * stride))) {
    free_adjoint(ctx);
    return rc;
  }
  ZF_HIP(ctx, hipMemcpy(ctx->d_adj_pts, xyz, 24 * (size_t)npts, hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_adj_pts + 3 * npts, normal, 24 * (size_t)npts, hipMemcpyHostToDevice));
  // pulled entries and probe support, marked on the device, listed here
  uint8_t *mark = nullptr;
  if ((rc = dev_alloc(ctx, (void **)&mark, Q * stride))) {
    free_adjoint(ctx);
    return rc;
  }
  ZF_HIP(ctx, hipMemsetAsync(mark, 0, Q * stride, ctx->stream));
  if (ctx->n_int > 0)
    hipLaunchKernelGGL(k_mark_pulled, dim3(nb(ctx->n_int, 256)), dim3(256), 0, ctx->stream, stride,
                       ctx->n_int, ctx->d_word, ctx->d_block_chord0, ctx->d_chords, mark);
  std::vector<uint8_t> hm(Q * stride);
  ZF_HIP(ctx, hipGetLastError());
  // the context stream does not synchronise with hipMemcpy's stream
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ZF_HIP(ctx, hipMemcpy(hm.data(), mark, Q * stride, hipMemcpyDeviceToHost));
  std::vector<BcEntry> hbc(ctx->n_bc);
  if (ctx->n_bc)
    ZF_HIP(ctx, hipMemcpy(hbc.data(), ctx->d_bc, sizeof(BcEntry) * ctx->n_bc, hipMemcpyDeviceToHost));
  // entries the adjoint reads: interior nodes, boundary nodes, late neighbours
  std::vector<uint8_t> read(stride, 0);
  for (int s = 0; s < ctx->n_int; s++) read[s] = 1;
  for (int t = 0; t < ctx->n_bc; t++) {
    read[hbc[t].s] = 1;
    if (t >= ctx->n_bc_fused) read[hbc[t].sm] = 1;
  }
  std::vector<int64_t> unpulled;
  for (int c = 0; c < Q; c++)
    for (size_t s = 0; s < stride; s++)
      if (read[s] && !hm[(size_t)c * stride + s]) unpulled.push_back((int64_t)(c * stride + s));
  ctx->device_bytes -= (int64_t)(Q * stride);
  dev_free(ctx, (void **)&mark);
  // probe stencils -> per support node (probe, weight) lists, probe order
  int32_t *tcnt = nullptr, *tidx = nullptr;
  double *tw = nullptr;
  const size_t tslots = (size_t)npts * kStencilMax;
  if ((rc = dev_alloc(ctx, (void **)&tcnt, 4 * (size_t)npts)) ||
      (rc = dev_alloc(ctx, (void **)&tidx, 4 * tslots)) || (rc = dev_alloc(ctx, (void **)&tw, 8 * tslots))) {
    dev_free(ctx, (void **)&tcnt);
    dev_free(ctx, (void **)&tidx);
    free_adjoint(ctx);
    return rc;
  }
  hipLaunchKernelGGL(k_stencil_table, dim3(nb((size_t)npts, 64)), dim3(64), 0, ctx->stream, ctx->d_b,
                     ctx->d_sbox, wgrid(ctx), (int)npts, ctx->d_adj_pts, tcnt, tidx, tw);
  std::vector<int32_t> hcnt(npts), hidx(tslots);
  std::vector<double> hw(tslots);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(hcnt.data(), tcnt, 4 * (size_t)npts, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hidx.data(), tidx, 4 * tslots, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hw.data(), tw, 8 * tslots, hipMemcpyDeviceToHost);
  ctx->device_bytes -= (int64_t)(4 * (size_t)npts + 12 * tslots);
  dev_free(ctx, (void **)&tcnt);
  dev_free(ctx, (void **)&tidx);
  dev_free(ctx, (void **)&tw);
  if (e != hipSuccess) {
    free_adjoint(ctx);
    return fail(ctx, ZF_ERR_HIP, "adjoint stencil table: %s", hipGetErrorString(e));
  }
  std::vector<std::vector<std::pair<int32_t, double>>> per(stride);
  for (int64_t m = 0; m < npts; m++) {
    if (hcnt[m] > kStencilMax) {
      free_adjoint(ctx);
      return fail(ctx, ZF_ERR_ARG, "probe %lld has %d stencil nodes (> %d)", (long long)m, hcnt[m],
                  kStencilMax);
    }
    for (int k = 0; k < hcnt[m]; k++)
      per[hidx[(size_t)m * kStencilMax + k]].push_back({(int32_t)m, hw[(size_t)m * kStencilMax + k]});
  }
  std::vector<int32_t> support, off(1, 0), sprobe, sup_of(stride, -1);
  std::vector<double> sw;
  for (size_t q = 0; q < stride; q++)
    if (!per[q].empty()) {
      sup_of[q] = (int32_t)support.size();
      support.push_back((int32_t)q);
      for (auto &pw : per[q]) {
        sprobe.push_back(pw.first);
        sw.push_back(pw.second);
      }
      off.push_back((int32_t)sprobe.size());
    }
  // probe-major table in stencil order: (support index, weight)
  std::vector<int32_t> poff(1, 0), psup;
  std::vector<double> pw;
  for (int64_t m = 0; m < npts; m++) {
    for (int k = 0; k < hcnt[m]; k++) {
      psup.push_back(sup_of[hidx[(size_t)m * kStencilMax + k]]);
      pw.push_back(hw[(size_t)m * kStencilMax + k]);
    }
    poff.push_back((int32_t)psup.size());
  }
  ctx->n_unpulled = (int64_t)unpulled.size();
  ctx->n_support = (int32_t)support.size();
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_unpulled, 8 * std::max<size_t>(unpulled.size(), 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_support, 4 * std::max<size_t>(support.size(), 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_sup_off, 4 * off.size())) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_sup_probe, 4 * std::max<size_t>(sprobe.size(), 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_sup_w, 8 * std::max<size_t>(sw.size(), 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_sup_u, 32 * std::max<size_t>(support.size(), 1) * ctx->obj_batch)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_sup_g, 24 * std::max<size_t>(support.size(), 1) * ctx->obj_batch)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_prb_off, 4 * poff.size())) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_prb_sup, 4 * std::max<size_t>(psup.size(), 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_prb_w, 8 * std::max<size_t>(pw.size(), 1)))) {
    free_adjoint(ctx);
    return rc;
  }
  ZF_HIP(ctx, hipMemcpy(ctx->d_prb_off, poff.data(), 4 * poff.size(), hipMemcpyHostToDevice));
  if (!psup.empty()) {
    ZF_HIP(ctx, hipMemcpy(ctx->d_prb_sup, psup.data(), 4 * psup.size(), hipMemcpyHostToDevice));
    ZF_HIP(ctx, hipMemcpy(ctx->d_prb_w, pw.data(), 8 * pw.size(), hipMemcpyHostToDevice));
  }
  ZF_HIP(ctx, hipMemcpy(ctx->d_sup_off, off.data(), 4 * off.size(), hipMemcpyHostToDevice));
  if (!sprobe.empty()) {
    ZF_HIP(ctx, hipMemcpy(ctx->d_sup_probe, sprobe.data(), 4 * sprobe.size(), hipMemcpyHostToDevice));
    ZF_HIP(ctx, hipMemcpy(ctx->d_sup_w, sw.data(), 8 * sw.size(), hipMemcpyHostToDevice));
  }
  if (!unpulled.empty())
    ZF_HIP(ctx, hipMemcpy(ctx->d_unpulled, unpulled.data(), 8 * unpulled.size(), hipMemcpyHostToDevice));
  if (!support.empty())
    ZF_HIP(ctx, hipMemcpy(ctx->d_support, support.data(), 4 * support.size(), hipMemcpyHostToDevice));
  ctx->have_adj = true;
  return ZF_OK;
}

int zf_adjoint_forward(zf_ctx *ctx, int32_t iter0, int32_t nsteps, double *objective) {
  if (!ctx || nsteps < 1) return ZF_ERR_ARG;
  int rc = require_adj(ctx);
  if (rc) return rc;
  const zf_adjoint_config &a = ctx->acfg;
  if (nsteps > a.max_steps) return fail(ctx, ZF_ERR_ARG, "nsteps %d > max_steps %d", nsteps, a.max_steps);
  if (a.t_begin <= iter0 || a.t_end > iter0 + nsteps)
    return fail(ctx, ZF_ERR_ARG, "objective window [%d, %d] outside iterations %d..%d", a.t_begin,
                a.t_end, iter0 + 1, iter0 + nsteps);
  const int K = a.checkpoint_interval;
  const bool cells = ctx->adj_cells && ctx->ncell > 0;
  const int I = cells ? ctx->icfg.rbc_interval : 1;
  if (ctx->adj_cells) {
    // the field in effect at the start; the cell states to compare at the end
    if ((rc = ensure_force_field(ctx))) return rc;
    ctx->h_status0.assign(ctx->ncell, 0);
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->ncell)
      ZF_HIP(ctx, hipMemcpy(ctx->h_status0.data(), ctx->d_status, 4 * (size_t)ctx->ncell,
                            hipMemcpyDeviceToHost));
    ZF_HIP(ctx, hipMemcpyAsync(ctx->d_rec_field, ctx->d_force, 3 * ctx->stride * field_elem(ctx),
                               hipMemcpyDeviceToDevice, ctx->stream));
  }
  ctx->adj_nrec = 0;
  ZF_HIP(ctx, hipEventRecord(ctx->events[12], ctx->stream));
  for (int k = iter0 + 1; k <= iter0 + nsteps; k++) {
    const int r = k - 1 - iter0;  // state S_{k-1}
    if (r % K == 0)
      ZF_HIP(ctx, hipMemcpyAsync(state_at(ctx, ctx->d_ckpt, r / K), ctx->d_f[ctx->cur],
                                 ctx->adj_state_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    const bool rbc = cells && k % I == 0;
    if ((rc = zf_lbm_step(ctx, k, rbc ? 1 : 0))) return rc;
    if (rbc) {
      // the cell stages, recorded between spreading and the position update
      if ((rc = rbc_update(ctx, false)) || (rc = adjoint_cells_record(ctx, ctx->adj_nrec)) ||
          (rc = advance_impl(ctx, nullptr)))
        return rc;
      ctx->adj_nrec++;
    }
    if (k >= a.t_begin && k <= a.t_end) obj_forward(ctx, ctx->d_f[ctx->cur], k);
  }
  ZF_HIP(ctx, hipEventRecord(ctx->events[13], ctx->stream));
  ZF_HIP(ctx, hipGetLastError());
  if (ctx->adj_cells && ctx->ncell) {
    std::vector<int32_t> st(ctx->ncell);
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ZF_HIP(ctx, hipMemcpy(st.data(), ctx->d_status, 4 * (size_t)ctx->ncell, hipMemcpyDeviceToHost));
    if (st != ctx->h_status0)
      return fail(ctx, ZF_ERR_STATE, "a cell was removed during the recorded forward pass");
  }
  const size_t nobj = (size_t)(a.t_end - a.t_begin + 1) * ctx->adj_npts;
  std::vector<double> h(nobj);
  ZF_HIP(ctx, hipMemcpyAsync(h.data(), ctx->d_objbuf, 8 * nobj, hipMemcpyDeviceToHost, ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ZF_HIP(ctx, hipEventElapsedTime(&ctx->adj_ms[0], ctx->events[12], ctx->events[13]));
  double sum = 0;
  for (double v : h) sum += v;
  if (objective) *objective = sum / (double)nobj;
  ctx->adj_iter0 = iter0;
  ctx->adj_nsteps = nsteps;
  ctx->adj_ready = true;
  return ZF_OK;
}

int zf_adjoint_backward(zf_ctx *ctx, double grad[5]) {
  if (!ctx || !grad) return ZF_ERR_ARG;
  int rc = require_adj(ctx);
  if (rc) return rc;
  if (!ctx->adj_ready) return fail(ctx, ZF_ERR_STATE, "zf_adjoint_forward not run");
  const zf_adjoint_config &a = ctx->acfg;
  const int K = a.checkpoint_interval, iter0 = ctx->adj_iter0, N = ctx->adj_nsteps;
  const size_t stride = ctx->stride, ab = ctx->cfg.storage == ZF_STORE_F64 ? 8 : 4;
  // a force field uploaded after the set-up
  if (ctx->d_force && !ctx->d_gF && (rc = dev_alloc(ctx, (void **)&ctx->d_gF, 24 * ctx->stride)))
    return rc;
  const bool cells = ctx->adj_cells && ctx->ncell > 0;
  const int I = cells ? ctx->icfg.rbc_interval : 1;
  void *const field0 = ctx->d_force;
  // the force field iteration k read: the initial one, or the one spread at
  // the last RBC iteration before k
  auto field_of = [&](int k) -> void * {
    if (!cells) return field0;
    const int m = (k - 1) / I - iter0 / I;
    return static_cast<char *>(ctx->d_rec_field) + 3 * ctx->stride * field_elem(ctx) * (size_t)m;
  };
  ZF_HIP(ctx, hipEventRecord(ctx->events[12], ctx->stream));
  if (ctx->adj_cells) {
    const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
    hipLaunchKernelGGL(k_fill_a<double>, dim3(nb(3 * NT, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_lamX, 3 * NT);
    hipLaunchKernelGGL(k_fill_a<double>, dim3(1), dim3(64), 0, ctx->stream, ctx->d_cgrad, (size_t)2);
  }
  const int rows = (int)nb(ctx->n_int, 256) + 1;
  hipLaunchKernelGGL(k_fill_a<double>, dim3(nb(5 * (size_t)rows, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_gpart, 5 * (size_t)rows);
  if (ctx->d_gF)
    hipLaunchKernelGGL(k_fill_a<double>, dim3(nb(3 * stride, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_gF, 3 * stride);
  for (int b = 0; b < 2; b++)
    with_store(ctx->cfg.storage, [&](auto tag) {
      constexpr int S = decltype(tag)::value;
      using A = typename Adj<S>::A;
      hipLaunchKernelGGL(k_fill_a<A>, dim3(nb(Q * stride, 256)), dim3(256), 0, ctx->stream,
                         static_cast<A *>(ctx->d_lam[b]), Q * stride);
      return 0;
    });
  (void)ab;
  int cur = 0;
  // lambda_N = dJ/dS_N
  if (iter0 + N >= a.t_begin && iter0 + N <= a.t_end) {
    if ((rc = obj_prepare(ctx, ctx->d_f[ctx->cur], ctx->d_f[ctx->cur], 0, 1))) return rc;
    obj_apply(ctx, 0, ctx->d_lam[cur]);
  }
  const int nseg = (N + K - 1) / K;
  for (int j = nseg - 1; j >= 0; j--) {
    // segment j: iterations iter0 + jK + 1 .. iter0 + min((j+1)K, N), which
    // read S_{iter0+jK} (checkpoint j) .. S_{iter0+min((j+1)K,N)-1}
    const int len = std::min(K, N - j * K);
    const void *ck = state_at(ctx, ctx->d_ckpt, j);
    auto state = [&](int r) -> const void * {  // S_{iter0 + jK + r}
      return r == 0 ? ck : state_at(ctx, ctx->d_seg, r - 1);
    };
    for (int r = 1; r < len; r++) {
      ctx->d_force = field_of(iter0 + j * K + r);
      if ((rc = step_into(ctx, state(r - 1), state_at(ctx, ctx->d_seg, r - 1), iter0 + j * K + r))) {
        ctx->d_force = field0;
        return rc;
      }
    }
    // objective data of the segment's states in the window, prepared in
    // batches of up to obj_batch states as the sweep reaches them
    int blo = 0, bhi = -1;
    const int rwin = std::max(0, a.t_begin - (iter0 + j * K));  // first state index in the window
    for (int r = len; r >= 1; r--) {
      const int k = iter0 + j * K + r;  // iteration k reads S_{k-1} = state(r - 1)
      ctx->d_force = field_of(k);
      rc = adjoint_step(ctx, state(r - 1), ctx->d_lam[cur], ctx->d_lam[cur ^ 1], k);
      ctx->d_force = field0;
      if (rc) return rc;
      cur ^= 1;
      if (k - 1 >= a.t_begin && k - 1 <= a.t_end) {
        const int rs = r - 1;
        if (rs < blo || rs > bhi) {
          bhi = rs;
          blo = std::max(rwin, rs - ctx->obj_batch + 1);
          if ((rc = obj_prepare(ctx, ck, ctx->d_seg, blo, bhi - blo + 1))) return rc;
        }
        obj_apply(ctx, rs - blo, ctx->d_lam[cur]);
      }
      // RBC iteration k - 1: its cell stages read S_{k-1}; the field they
      // spread was read by iterations k .. (its gradient is complete now)
      if (cells && k - 1 > iter0 && (k - 1) % I == 0 &&
          (rc = adjoint_cells_backward(ctx, (k - 1) / I - iter0 / I - 1, state(r - 1), ctx->d_lam[cur])))
        return rc;
    }
  }
  ctx->d_force = field0;
  ctx->adj_cur = cur;
  hipLaunchKernelGGL(k_reduce_grads, dim3(1), dim3(64), 0, ctx->stream, ctx->d_gpart, rows,
                     ctx->d_grads);
  ZF_HIP(ctx, hipEventRecord(ctx->events[13], ctx->stream));
  ZF_HIP(ctx, hipGetLastError());
  double g[8];
  ZF_HIP(ctx, hipMemcpyAsync(g, ctx->d_grads, sizeof g, hipMemcpyDeviceToHost, ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ZF_HIP(ctx, hipEventElapsedTime(&ctx->adj_ms[1], ctx->events[12], ctx->events[13]));
  for (int i = 0; i < 5; i++) grad[i] = g[i];
  return ZF_OK;
}

int zf_adjoint_download(zf_ctx *ctx, double *dJ_df0, double *dJ_dforce) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_adj(ctx);
  if (rc) return rc;
  const int n = ctx->dommax;
  if (dJ_df0 && n > 0) {
    double *tmp = nullptr;
    if ((rc = dev_alloc(ctx, (void **)&tmp, 8 * (size_t)Q * n))) return rc;
    with_store(ctx->cfg.storage, [&](auto tag) {
      constexpr int S = decltype(tag)::value;
      using A = typename Adj<S>::A;
      hipLaunchKernelGGL(k_export_lam<A>, dim3(nb(n, 256)), dim3(256), 0, ctx->stream,
                         static_cast<const A *>(ctx->d_lam[ctx->adj_cur]), ctx->stride, n,
                         ctx->d_s_of_p, tmp);
      return 0;
    });
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(dJ_df0, tmp, 8 * (size_t)Q * n, hipMemcpyDeviceToHost);
    ctx->device_bytes -= 8 * (int64_t)Q * n;
    dev_free(ctx, (void **)&tmp);
    if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "adjoint download: %s", hipGetErrorString(e));
  }
  if (dJ_dforce && n > 0) {
    std::vector<double> g(3 * ctx->stride, 0.0);
    std::vector<int32_t> sp(n);
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->d_gF)
      ZF_HIP(ctx, hipMemcpy(g.data(), ctx->d_gF, 24 * ctx->stride, hipMemcpyDeviceToHost));
    ZF_HIP(ctx, hipMemcpy(sp.data(), ctx->d_s_of_p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    for (int p = 0; p < n; p++)
      for (int a = 0; a < 3; a++) dJ_dforce[(size_t)a * n + p] = g[(size_t)a * ctx->stride + sp[p]];
  }
  return ZF_OK;
}

int zf_adjoint_timing(zf_ctx *ctx, double ms[2]) {
  if (!ctx || !ms) return ZF_ERR_ARG;
  ms[0] = ctx->adj_ms[0];
  ms[1] = ctx->adj_ms[1];
  return ZF_OK;
}

}  // extern "C"
