// zf_adjoint_ibm.hip -- hand-written vector-Jacobian products of the
// immersed-boundary coupling (the transposes of zf_ibm.hip's operators).
//
// Forward (ZebraflowWT.c:8702-8996 interpolation, :10746-10867 spreading),
// with w(d) = prod_a (1 + cos(pi |d_a| / 2)) / 64, d = (x_lattice - X)/DX
// in lattice units and |d| <= 2:
//   V_n    = vs  sum_q w(x_q - X_n) u_q        q: 0 < B < 4998
//   F(s)   = fs  sum_n w(x_s - X_n) G_n        s: B >= 0
// Transposes, for adjoints lam_V (per membrane node) and lam_F (per node):
//   lam_G_n  = fs  sum_s w lam_F(s)
//   lam_u(q) = vs  sum_n w lam_V(n)
//   lam_X_n  = fs  sum_s grad_X w (lam_F(s) . G_n) + vs sum_q grad_X w (lam_V(n) . u_q)
// with grad_X w_a = (pi / (2 DX)) sin(pi |d_a| / 2) sign(d_a) prod_{b != a} phi_b / 64.
// lam_u reaches the populations through u = j / (1 + drho):
//   lam_f(q, c) += lam_u(q) . (c - u_q) / (1 + drho_q).
// The membrane-node sums run one lane per node over its 7^3 window; lam_u
// is gathered per lattice node from the membrane nodes of the 27
// surrounding 4^3 tiles (the spreading kernel's scheme), so every sum has a
// fixed order and no atomics.
#include "zf_d3q19.h"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

namespace zf {
namespace {

using namespace d3q19;

constexpr int kTile = 4;
constexpr double kPiA = 3.14159265358979;

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct AP {
  double x0, y0, z0, dx, vs, fs;
  int im, jm, km, ntx, nty, ntz;
  int cube;  // ZF_IBM_CUBE_SUPPORT
};

// Outside the kernel support (d in lattice units), as the forward kernels.
__device__ __forceinline__ bool off_support(double dx, double dy, double dz, int cube) {
  if (cube) return fabs(dx) >= 2.0 || fabs(dy) >= 2.0 || fabs(dz) >= 2.0;
  return dx * dx + dy * dy + dz * dz > 4.0;
}

AP aparams(const zf_ctx *ctx) {
  return AP{ctx->icfg.x0, ctx->icfg.y0, ctx->icfg.z0, ctx->icfg.dx, ctx->icfg.vel_scale,
            ctx->icfg.force_scale, ctx->cfg.im, ctx->cfg.jm, ctx->cfg.km, ctx->ntx, ctx->nty,
            ctx->ntz, (ctx->icfg.flags & ZF_IBM_CUBE_SUPPORT) ? 1 : 0};
}

// Weights in the arithmetic of the forward's fields: FP64 for the FP64
// form, FP32 otherwise (as the forward's spreading).
template <int S>
using W = typename Store<S>::F;
__device__ __forceinline__ float cos_pi(float x) { return cospif(x); }
__device__ __forceinline__ double cos_pi(double x) { return cospi(x); }
__device__ __forceinline__ float sin_pi(float x) { return sinpif(x); }
__device__ __forceinline__ double sin_pi(double x) { return sinpi(x); }

// Kernel weight w(d) = prod_a (1 + cos(pi |d_a| / 2)) / 64 and its gradient
// with respect to the membrane node position, in units of 1/DX.
template <class R>
__device__ __forceinline__ R weight_grad(R dxl, R dyl, R dzl, R g[3]) {
  const R h = (R)0.5;
  const R ax = fabs(dxl), ay = fabs(dyl), az = fabs(dzl);
  const R px = (R)1 + cos_pi(h * ax), py = (R)1 + cos_pi(h * ay), pz = (R)1 + cos_pi(h * az);
  const R k = (R)(kPiA * 0.5 / 64.0);
  const R sx = sin_pi(h * ax) * (dxl > 0 ? (R)1 : dxl < 0 ? (R)-1 : (R)0);
  const R sy = sin_pi(h * ay) * (dyl > 0 ? (R)1 : dyl < 0 ? (R)-1 : (R)0);
  const R sz = sin_pi(h * az) * (dzl > 0 ? (R)1 : dzl < 0 ? (R)-1 : (R)0);
  g[0] = k * sx * py * pz;
  g[1] = k * px * sy * pz;
  g[2] = k * px * py * sz;
  return px * py * pz * (R)(1.0 / 64.0);
}

// Unshifted velocity of every listed node (the forward's interpolation
// input), in the field type.
template <int S>
__global__ void k_adj_velocity(const typename Store<S>::T *__restrict__ f, size_t stride, int n,
                               W<S> *__restrict__ u) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double drho = 0, j0 = 0, j1 = 0, j2 = 0;
#pragma unroll
  for (int c = 0; c < Q; c++) {
    const double g = (double)Store<S>::ld(f, (size_t)c * stride + s, WQ[c]);
    drho += g;
    j0 += EX[c] * g;
    j1 += EY[c] * g;
    j2 += EZ[c] * g;
  }
  const double rho = 1.0 + drho;
  u[s] = (W<S>)(j0 / rho);
  u[stride + s] = (W<S>)(j1 / rho);
  u[2 * stride + s] = (W<S>)(j2 / rho);
}

template <int S>
__device__ __forceinline__ void node_u(const typename Store<S>::T *__restrict__ f, size_t stride,
                                       int s, double u[3], double &rho) {
  double drho = 0, j0 = 0, j1 = 0, j2 = 0;
#pragma unroll
  for (int c = 0; c < Q; c++) {
    const double g = (double)Store<S>::ld(f, (size_t)c * stride + s, WQ[c]);
    drho += g;
    j0 += EX[c] * g;
    j1 += EY[c] * g;
    j2 += EZ[c] * g;
  }
  rho = 1.0 + drho;
  u[0] = j0 / rho;
  u[1] = j1 / rho;
  u[2] = j2 / rho;
}

// One lane per membrane node: lam_G and the position adjoint (velocities
// of the support nodes from uq).
template <int S>
__global__ __launch_bounds__(128) void k_ibm_adj_node(
    const double *__restrict__ X, size_t NT, int nn, int ncell, const int32_t *__restrict__ status,
    const int32_t *__restrict__ b, const int32_t *__restrict__ sbox, const W<S> *__restrict__ uq,
    size_t stride, AP p, const int32_t *__restrict__ vslot, const double *__restrict__ uvoid,
    const double *__restrict__ G, const double *__restrict__ lamV, const double *__restrict__ lamF,
    double *__restrict__ lamG, double *__restrict__ lamX) {
  // 8 lanes per membrane node: lane `sub` takes the window plane i - 3 + sub
  // (sub < 7), and a fixed-order shuffle reduction over the 8 lanes sums the
  // planes (deterministic); 8x the lanes of one lane per node, which left
  // most of the device idle for a few tens of cells
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = t >> 3, sub = t & 7;
  const bool live = g < ncell * nn && status[min(g, ncell * nn - 1) / nn] >= 0;
  double lg0 = 0, lg1 = 0, lg2 = 0, lx0 = 0, lx1 = 0, lx2 = 0;
  if (live) {
    const double xp = X[g], yp = X[NT + g], zp = X[2 * NT + g];
    const double G0 = G ? G[g] : 0, G1 = G ? G[NT + g] : 0, G2 = G ? G[2 * NT + g] : 0;
    const double V0 = lamV ? lamV[g] : 0, V1 = lamV ? lamV[NT + g] : 0, V2 = lamV ? lamV[2 * NT + g] : 0;
    const int i = (int)((xp - p.x0) / p.dx + 1.01), j = (int)((yp - p.y0) / p.dx + 1.01),
              k = (int)((zp - p.z0) / p.dx + 1.01);
    const int j1 = max(1, j - 3), j2 = min(p.jm, j + 3), k1 = max(1, k - 3), k2 = min(p.km, k + 3);
    const double inv = 1.0 / p.dx;
    const int ii = i - 3 + sub;
    if (sub < 7 && ii >= 1 && ii <= p.im) {
      const double dxl = (p.x0 + (ii - 1) * p.dx - xp) * inv;
      for (int jj = j1; jj <= j2; jj++) {
        const double dyl = (p.y0 + (jj - 1) * p.dx - yp) * inv;
        for (int kk = k1; kk <= k2; kk++) {
          const double dzl = (p.z0 + (kk - 1) * p.dx - zp) * inv;
          if (off_support(dxl, dyl, dzl, p.cube)) continue;
          const size_t q = ((size_t)(ii - 1) * p.jm + (jj - 1)) * p.km + (kk - 1);
          const int bv = b[q];
          const int sl = (vslot && !(bv > 0 && bv < 4998)) ? vslot[q] : -1;
          if (bv < 0 && sl < 0) continue;
          const int s = bv >= 0 ? sbox[q] : -1;
          W<S> gr[3];
          const double w = (double)weight_grad((W<S>)dxl, (W<S>)dyl, (W<S>)dzl, gr);
          const double gw[3] = {(double)gr[0] * inv, (double)gr[1] * inv, (double)gr[2] * inv};
          if (lamF && bv >= 0) {  // spreading onto every B >= 0 node
            const double l0 = lamF[s], l1 = lamF[stride + s], l2 = lamF[2 * stride + s];
            lg0 += p.fs * w * l0;
            lg1 += p.fs * w * l1;
            lg2 += p.fs * w * l2;
            const double lg = p.fs * (l0 * G0 + l1 * G1 + l2 * G2);
            lx0 += lg * gw[0];
            lx1 += lg * gw[1];
            lx2 += lg * gw[2];
          }
          if (lamV && bv > 0 && bv < 4998) {  // interpolation from fluid nodes
            const double lv = p.vs * (V0 * (double)uq[s] + V1 * (double)uq[stride + s] +
                                      V2 * (double)uq[2 * stride + s]);
            lx0 += lv * gw[0];
            lx1 += lv * gw[1];
            lx2 += lv * gw[2];
          } else if (lamV && sl >= 0) {  // the wall layer's fixed Uvoid
            const double lv = p.vs * (V0 * uvoid[3 * (size_t)sl] + V1 * uvoid[3 * (size_t)sl + 1] +
                                      V2 * uvoid[3 * (size_t)sl + 2]);
            lx0 += lv * gw[0];
            lx1 += lv * gw[1];
            lx2 += lv * gw[2];
          }
        }
      }
    }
  }
  // every lane of the 8 takes part (the group never straddles a wave)
  for (int o = 4; o > 0; o >>= 1) {
    lg0 += __shfl_xor(lg0, o, 8);
    lg1 += __shfl_xor(lg1, o, 8);
    lg2 += __shfl_xor(lg2, o, 8);
    lx0 += __shfl_xor(lx0, o, 8);
    lx1 += __shfl_xor(lx1, o, 8);
    lx2 += __shfl_xor(lx2, o, 8);
  }
  if (!live || sub != 0) return;
  if (lamG) {
    lamG[g] = lg0;
    lamG[NT + g] = lg1;
    lamG[2 * NT + g] = lg2;
  }
  lamX[g] += lx0;
  lamX[NT + g] += lx1;
  lamX[2 * NT + g] += lx2;
}

__device__ __forceinline__ int tile_of(double pos, double origin, double dx, int nt) {
  int t = (int)floor((pos - origin) / dx / kTile);
  return t < 0 ? 0 : t >= nt ? nt - 1 : t;
}

// tile << 32 | node; dead cells sort past the last tile.
__global__ void k_tile_keys(const double *__restrict__ X, size_t NT, int nn, int ncell,
                            const int32_t *__restrict__ status, AP p, uint64_t *__restrict__ keys) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  const uint64_t ntiles = (uint64_t)p.ntx * p.nty * p.ntz;
  if (status[g / nn] < 0) {
    keys[g] = (ntiles << 32) | (uint64_t)g;
    return;
  }
  const int tx = tile_of(X[g], p.x0, p.dx, p.ntx), ty = tile_of(X[NT + g], p.y0, p.dx, p.nty),
            tz = tile_of(X[2 * NT + g], p.z0, p.dx, p.ntz);
  keys[g] = ((((uint64_t)tx * p.nty + ty) * p.ntz + tz) << 32) | (uint64_t)g;
}

__global__ void k_ranges(const uint64_t *__restrict__ keys, int n, uint64_t ntiles,
                         int32_t *__restrict__ range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t t = keys[i] >> 32;
  if (t >= ntiles) return;
  if (i == 0 || (keys[i - 1] >> 32) != t) range[2 * t] = i;
  if (i == n - 1 || (keys[i + 1] >> 32) != t) range[2 * t + 1] = i + 1;
}

__global__ void k_fill32(int32_t *a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// lam_u per fluid lattice node (one 64-lane wave per 4^3 tile, membrane
// nodes of the 27 neighbouring tiles in sorted order), then
// lam(q, c) += lam_u . (c - u_q)/(1 + drho_q).
template <int S, class A>
__global__ __launch_bounds__(64) void k_interp_adj_gather(
    const uint64_t *__restrict__ keys, const int32_t *__restrict__ range,
    const double *__restrict__ X, const double *__restrict__ lamV, size_t NT,
    const int32_t *__restrict__ b, const int32_t *__restrict__ sbox,
    const typename Store<S>::T *__restrict__ f, size_t stride, AP p, A *__restrict__ lam) {
  using R = W<S>;
  __shared__ double s0[64], s1[64], s2[64];
  __shared__ R kx[4][64], ky[4][64], kz[4][64];  // (1 + cos(pi |d| / 2)) for tile row r
  __shared__ double sx[64], sy[64], sz[64];
  const int tile = blockIdx.x;
  const int tz = tile % p.ntz, ty = (tile / p.ntz) % p.nty, tx = tile / (p.ntz * p.nty);
  const int lane = threadIdx.x;
  const int ix = lane / 16, iy = (lane / 4) % 4, iz = lane % 4;
  const int li = tx * kTile + ix, lj = ty * kTile + iy, lk = tz * kTile + iz;
  const double qx = p.x0 + li * p.dx, qy = p.y0 + lj * p.dx, qz = p.z0 + lk * p.dx;
  const double inv = 1.0 / p.dx;
  const R h = (R)0.5;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int dxT = -1; dxT <= 1; dxT++)
    for (int dyT = -1; dyT <= 1; dyT++)
      for (int dzT = -1; dzT <= 1; dzT++) {
        const int nx = tx + dxT, ny = ty + dyT, nz = tz + dzT;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= p.ntx || ny >= p.nty || nz >= p.ntz) continue;
        const int nt = (nx * p.nty + ny) * p.ntz + nz;
        const int beg = range[2 * nt], end = range[2 * nt + 1];
        for (int c0 = beg; c0 < end; c0 += 64) {
          // The weight is a product of one factor per axis, and a lane's
          // offset from a node along an axis is one of 4 (its tile row): the
          // staging lane evaluates the node's 12 offsets and factors with the
          // lanes' own expressions, and nodes at more than 2 DX from every
          // row of some axis (off the support of all 64 lanes) are dropped;
          // the others are compacted in order, so the sums are unchanged.
          bool rel = false;
          double px = 0, py = 0, pz = 0, dv[3][4];
          int g = 0;
          if (lane < end - c0) {
            g = (int)(keys[c0 + lane] & 0xffffffffu);
            px = X[g];
            py = X[NT + g];
            pz = X[2 * NT + g];
            bool ax[3] = {false, false, false};
#pragma unroll
            for (int r = 0; r < 4; r++) {
              dv[0][r] = (p.x0 + (tx * kTile + r) * p.dx - px) * inv;
              dv[1][r] = (p.y0 + (ty * kTile + r) * p.dx - py) * inv;
              dv[2][r] = (p.z0 + (tz * kTile + r) * p.dx - pz) * inv;
#pragma unroll
              for (int a = 0; a < 3; a++) ax[a] = ax[a] || fabs(dv[a][r]) <= 2.0;
            }
            rel = ax[0] && ax[1] && ax[2];
          }
          const unsigned long long relm = __ballot(rel);
          const int cnt = __popcll(relm);
          if (rel) {
            const int sl = __popcll(relm & ((1ull << lane) - 1ull));
            sx[sl] = px;
            sy[sl] = py;
            sz[sl] = pz;
            s0[sl] = lamV[g];
            s1[sl] = lamV[NT + g];
            s2[sl] = lamV[2 * NT + g];
#pragma unroll
            for (int r = 0; r < 4; r++) {
              kx[r][sl] = (R)1 + cos_pi(h * fabs((R)dv[0][r]));
              ky[r][sl] = (R)1 + cos_pi(h * fabs((R)dv[1][r]));
              kz[r][sl] = (R)1 + cos_pi(h * fabs((R)dv[2][r]));
            }
          }
          __syncthreads();
          for (int m = 0; m < cnt; m++) {
            const double dxl = (qx - sx[m]) * inv, dyl = (qy - sy[m]) * inv, dzl = (qz - sz[m]) * inv;
            if (off_support(dxl, dyl, dzl, p.cube)) continue;
            const double w = (double)(kx[ix][m] * ky[iy][m] * kz[iz][m] * (R)(1.0 / 64.0));
            a0 += w * s0[m];
            a1 += w * s1[m];
            a2 += w * s2[m];
          }
          __syncthreads();
        }
      }
  if (li >= p.im || lj >= p.jm || lk >= p.km) return;
  if (a0 == 0.0 && a1 == 0.0 && a2 == 0.0) return;
  const size_t q = ((size_t)li * p.jm + lj) * p.km + lk;
  const int bv = b[q];
  if (!(bv > 0 && bv < 4998)) return;
  const int s = sbox[q];
  double u[3], rho;
  node_u<S>(f, stride, s, u, rho);
  a0 *= p.vs;
  a1 *= p.vs;
  a2 *= p.vs;
#pragma unroll
  for (int c = 0; c < Q; c++) {
    const double v = (a0 * (EX[c] - u[0]) + a1 * (EY[c] - u[1]) + a2 * (EZ[c] - u[2])) / rho;
    const size_t o = (size_t)c * stride + s;
    lam[o] = (A)((double)lam[o] + v);
  }
}

template <class A>
__global__ void k_export_lam_ibm(const A *__restrict__ lam, size_t stride, int n,
                                 const int32_t *__restrict__ s_of_p, double *__restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = s_of_p[p];
  for (int c = 0; c < Q; c++) out[(size_t)c * n + p] = (double)lam[(size_t)c * stride + s];
}

// host (c, n, axis) <-> device SoA [axis][c*nn + n]
__global__ void k_aos_to_soa(const double *__restrict__ h, int n, size_t NT, double *__restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; a++) A[a * NT + i] = h[3 * (size_t)i + a];
}
__global__ void k_soa_to_aos(const double *__restrict__ A, int n, size_t NT, double *__restrict__ h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; a++) h[3 * (size_t)i + a] = A[a * NT + i];
}
// host per-list-node [3][dommax] -> storage [3][stride]
__global__ void k_p_to_s3(const double *__restrict__ h, int n, const int32_t *__restrict__ s_of_p,
                          size_t stride, double *__restrict__ A) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int s = s_of_p[p];
  for (int a = 0; a < 3; a++) A[a * stride + s] = h[(size_t)a * n + p];
}

template <class F>
int with_form(int storage, F &&f) {
  if (storage == ZF_STORE_F64) return f(std::integral_constant<int, ZF_STORE_F64>());
  if (storage == ZF_STORE_F32DEV) return f(std::integral_constant<int, ZF_STORE_F32DEV>());
  return f(std::integral_constant<int, ZF_STORE_F16DEV>());
}

}  // namespace

int ibm_vjp_device(zf_ctx *ctx, const double *X, const void *f, const double *G,
                   const double *lamV, const double *lamF, double *lamG, double *lamX, void *lam,
                   bool lam_double) {
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  const AP p = aparams(ctx);
  // the interpolation's velocity field of f, in the forward's velocity
  // buffer (which the next forward interpolation overwrites)
  with_form(ctx->cfg.storage, [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    using T = typename Store<S>::T;
    W<S> *uq = static_cast<W<S> *>(ctx->d_umac);
    if (lamV)
      hipLaunchKernelGGL(k_adj_velocity<S>, dim3(nb(ctx->dommax, 256)), dim3(256), 0, ctx->stream,
                         static_cast<const T *>(f), ctx->stride, ctx->dommax, uq);
    hipLaunchKernelGGL(k_ibm_adj_node<S>, dim3(nb(8 * (size_t)n, 128)), dim3(128), 0, ctx->stream, X,
                       NT, ctx->nn, ctx->ncell, ctx->d_status, ctx->d_b, ctx->d_sbox, uq,
                       ctx->stride, p, ctx->have_void ? ctx->d_vslot : nullptr, ctx->d_uvoid, G,
                       lamV, lamF, lamG, lamX);
    return 0;
  });
  if (lam && lamV) {
    const uint64_t ntiles = (uint64_t)ctx->ntx * ctx->nty * ctx->ntz;
    hipLaunchKernelGGL(k_tile_keys, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, X, NT,
                       ctx->nn, ctx->ncell, ctx->d_status, p, ctx->d_keys[0]);
    int tbits = 1;
    while ((1ull << tbits) <= ntiles) tbits++;
    size_t tmp = ctx->sort_tmp_bytes;
    ZF_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->d_sort_tmp, tmp, ctx->d_keys[0],
                                                    ctx->d_keys[1], n, 0, 32 + tbits, ctx->stream));
    hipLaunchKernelGGL(k_fill32, dim3(nb(2 * ntiles, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_tile_range, (size_t)(2 * ntiles));
    hipLaunchKernelGGL(k_ranges, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_keys[1], n,
                       ntiles, ctx->d_tile_range);
    with_form(ctx->cfg.storage, [&](auto tag) {
      constexpr int S = decltype(tag)::value;
      using T = typename Store<S>::T;
      if (lam_double)
        hipLaunchKernelGGL((k_interp_adj_gather<S, double>), dim3((unsigned)ntiles), dim3(64), 0,
                           ctx->stream, ctx->d_keys[1], ctx->d_tile_range, X, lamV, NT,
                           ctx->d_b, ctx->d_sbox, static_cast<const T *>(f), ctx->stride, p,
                           static_cast<double *>(lam));
      else
        hipLaunchKernelGGL((k_interp_adj_gather<S, float>), dim3((unsigned)ntiles), dim3(64), 0,
                           ctx->stream, ctx->d_keys[1], ctx->d_tile_range, X, lamV, NT,
                           ctx->d_b, ctx->d_sbox, static_cast<const T *>(f), ctx->stride, p,
                           static_cast<float *>(lam));
      return 0;
    });
  }
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

}  // namespace zf

using namespace zf;

extern "C" {

int zf_ibm_vjp(zf_ctx *ctx, const double *lam_force, const double *lam_vel, const double *node_force,
               double *lam_node_force, double *lam_pos, double *lam_f) {
  if (!ctx || !lam_pos) return ZF_ERR_ARG;
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "zf_ibm_setup not called");
  if (lam_force && !node_force) return fail(ctx, ZF_ERR_ARG, "lam_force needs node_force");
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn, stride = ctx->stride;
  const int n = ctx->ncell * ctx->nn, np = ctx->dommax;
  if (n == 0) return ZF_OK;
  // device scratch: G, lamV, lamG, lamX (3 NT each), lamF (3 stride),
  // lam (19 stride, double), export buffers
  const size_t bytes = 8 * (12 * NT + 3 * stride + (size_t)Q * stride + (size_t)Q * np + 3 * (size_t)n);
  double *buf = nullptr;
  int rc = dev_alloc(ctx, (void **)&buf, bytes);
  if (rc) return rc;
  double *G = buf, *lamV = G + 3 * NT, *lamG = lamV + 3 * NT, *lamX = lamG + 3 * NT,
         *lamF = lamX + 3 * NT, *lam = lamF + 3 * stride, *exp = lam + (size_t)Q * stride,
         *hstage = exp + (size_t)Q * np;
  hipError_t e = hipMemsetAsync(buf, 0, bytes, ctx->stream);
  auto up3 = [&](const double *h, double *d) {
    if (e == hipSuccess) e = hipMemcpyAsync(hstage, h, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      hipLaunchKernelGGL(k_aos_to_soa, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, hstage, n, NT, d);
  };
  if (node_force) up3(node_force, G);
  if (lam_vel) up3(lam_vel, lamV);
  if (lam_force && e == hipSuccess) {
    e = hipMemcpyAsync(exp, lam_force, 24 * (size_t)np, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      hipLaunchKernelGGL(k_p_to_s3, dim3(nb(np, 256)), dim3(256), 0, ctx->stream, exp, np,
                         ctx->d_s_of_p, stride, lamF);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess &&
      (rc = ibm_vjp_device(ctx, ctx->d_x, ctx->d_f[ctx->cur], node_force ? G : nullptr, lam_vel ? lamV : nullptr,
                           lam_force ? lamF : nullptr, lamG, lamX, lam_f ? lam : nullptr, true))) {
    ctx->device_bytes -= (int64_t)bytes;
    dev_free(ctx, (void **)&buf);
    return rc;
  }
  auto down3 = [&](const double *d, double *h) {
    if (e == hipSuccess)
      hipLaunchKernelGGL(k_soa_to_aos, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, d, n, NT, hstage);
    if (e == hipSuccess) e = hipMemcpyAsync(h, hstage, 24 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  };
  if (lam_node_force) down3(lamG, lam_node_force);
  down3(lamX, lam_pos);
  if (lam_f && e == hipSuccess) {
    hipLaunchKernelGGL(k_export_lam_ibm<double>, dim3(nb(np, 256)), dim3(256), 0, ctx->stream, lam,
                       stride, np, ctx->d_s_of_p, exp);
    e = hipMemcpyAsync(lam_f, exp, 8 * (size_t)Q * np, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  ctx->device_bytes -= (int64_t)bytes;
  dev_free(ctx, (void **)&buf);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "ibm vjp: %s", hipGetErrorString(e));
  return ZF_OK;
}

}  // extern "C"
