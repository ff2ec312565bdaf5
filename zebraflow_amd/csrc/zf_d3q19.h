// zf_d3q19.h -- D3Q19 lattice tables, population storage forms and the
// equilibrium, shared by the time step, the adjoint and the wall operators.
// Directions in the reference's order (ZebraflowWT.c:2471-2504): c and c^1
// are opposite, c = 18 is the rest population.
#ifndef ZF_D3Q19_H
#define ZF_D3Q19_H

#include "zf_internal.h"

#include <type_traits>
#include <utility>

namespace zf {
namespace d3q19 {

template <int S>
struct Store;
// F: element of the force field and of the IBM velocity/force buffers.
template <>
struct Store<ZF_STORE_F64> {
  using T = double;
  using R = double;
  using F = double;
  __device__ static R ld(const T *a, size_t i, double w) { return a[i] - w; }
  __device__ static T st(R g, double w) { return g + w; }
};
template <>
struct Store<ZF_STORE_F32DEV> {
  using T = float;
  using R = float;
  using F = float;
  __device__ static R ld(const T *a, size_t i, double) { return a[i]; }
  __device__ static T st(R g, double) { return g; }
};
template <>
struct Store<ZF_STORE_F16DEV> {
  using T = _Float16;
  using R = float;
  using F = float;
  // w applied as hi + lo floats (the float weights alone are biased high)
  __device__ static R ld(const T *a, size_t i, double w) {
    const float h = (float)a[i];
    return __builtin_fmaf((float)w, h, (float)(w - (double)(float)w) * h);
  }
  __device__ static T st(R g, double w) { return (T)(g * (float)(1.0 / w)); }
};

// Compile-time lattice tables for fully unrolled loops.
__device__ constexpr int EX[Q] = {1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0, 0};
__device__ constexpr int EY[Q] = {0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0};
__device__ constexpr int EZ[Q] = {0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0};
__device__ constexpr double WQ[Q] = {
    1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 3};

// w_c * y.  In FP32 the weights 1/18, 1/36 and 1/3 round up, so that the
// float weights sum to 1 + 2.6e-8: multiplying by them makes the equilibrium
// a mass (and momentum) source proportional to drho on every step, a bias
// that accumulates along the flow.  The float form therefore multiplies by
// the weight split in two floats, w = hi + lo, which rounds w * y correctly
// up to 2^-48 and leaves only unbiased round-to-nearest errors.
template <class R>
__device__ __forceinline__ R wmul(int c, R y) { return (R)WQ[c] * y; }
template <>
__device__ __forceinline__ float wmul<float>(int c, float y) {
  const float hi = (float)WQ[c];
  const float lo = (float)(WQ[c] - (double)hi);
  return __builtin_fmaf(hi, y, lo * y);
}

// Equilibrium deviation feq_c - w_c for rho = 1 + drho (ZebraflowWT.c:6139),
// w (drho + X + drho X) with X = 3 c.u + 4.5 (c.u)^2 - 1.5 u^2: the sum 1 +
// drho is never formed, so the deviations keep their own relative precision.
template <class R>
__device__ __forceinline__ R geq(int c, R drho, R ux, R uy, R uz, R usq) {
  const R cu = (R)EX[c] * ux + (R)EY[c] * uy + (R)EZ[c] * uz;
  const R x = (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usq;
  return wmul<R>(c, drho + x + drho * x);
}

// Boundary values, ZebraflowWT.c:6294-6433: all 19 populations of a BC node
// set to feq(D_bc, U_bc) from its neighbour's rho, u.  Written to
// dst[c * stride + e.s - off] (off: the first storage index dst holds).
template <int S, class BP>
__device__ __forceinline__ void write_bc(typename Store<S>::T *__restrict__ dst, size_t stride,
                                         const BcEntry &e, typename Store<S>::R drm,
                                         typename Store<S>::R um, typename Store<S>::R vm,
                                         typename Store<S>::R wm, const BP &bp, bool nt,
                                         int off = 0) {
  using R = typename Store<S>::R;
  R drb, ub, vb, wb;  // density deviation and velocity of the boundary value
  if (e.code == 0 || e.code == 4998) {
    drb = drm; ub = vb = wb = 0;
  } else if (e.code >= 1001 && e.code <= 1009) {
    drb = (R)bp.pdrho[e.code - 1001]; ub = um; vb = vm; wb = wm;
  } else if (e.code == ZF_B_VELOCITY_INLET) {
    drb = drm; ub = (R)bp.inlet[0]; vb = (R)bp.inlet[1]; wb = (R)bp.inlet[2];
  } else {  // ZF_B_PRESSURE_OUTLET
    drb = (R)bp.outlet_drho; ub = um; vb = vm; wb = wm;
  }
  const R usq = ub * ub + vb * vb + wb * wb;
  const size_t at = (size_t)(e.s - off);
#pragma unroll
  for (int c = 0; c < Q; c++) {
    const auto v = Store<S>::st(geq<R>(c, drb, ub, vb, wb, usq), WQ[c]);
    if (nt) __builtin_nontemporal_store(v, &dst[(size_t)c * stride + at]);
    else dst[(size_t)c * stride + at] = v;
  }
}

// Workgroup order: the hardware deals workgroups round-robin over the 8
// XCDs.  order 0: plain.  order 1: XCD x runs one contiguous range of node
// blocks, so the chord boundaries a block shares with its neighbours stay in
// one L2.  order k > 1: the XCDs sweep one common front, XCD x taking chunk
// x of every 8 consecutive chunks of k blocks (blocks past the last whole
// round of 8k keep their place).  Every order is a permutation of [0, nblk).
__device__ __forceinline__ int logical_block(int b, int nblk, int order) {
  if (order == 0) return b;
  if (order > 1) {
    const int x = b % 8, i = b / 8;
    const int round = 8 * order, full = nblk / round * round;
    if (b >= full) return b;
    return (i / order) * round + x * order + i % order;
  }
  return xcd_range_block(b, nblk);
}

// Population load and store with a compile-time cache policy: non-temporal
// accesses do not allocate in the last-level cache.
template <bool NT, class T>
__device__ __forceinline__ T ld_policy(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st_policy(T *p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// f(integral_constant<int, c>) for c = 0..Q-1, each with its own constant c.
template <class Fn, int... C>
__device__ __forceinline__ void for_dirs_seq(Fn &&f, std::integer_sequence<int, C...>) {
  (f(std::integral_constant<int, C>()), ...);
}
template <class Fn>
__device__ __forceinline__ void for_dirs(Fn &&f) {
  for_dirs_seq(f, std::make_integer_sequence<int, Q>());
}

// BGK collision with Guo forcing of one node (ZebraflowWT.c:6127-6143) on
// the pulled deviations g: tau = 0.5 + visc_scale*bclass, half-force
// velocity shift, DT*(1 - 1/(2 tau)) source.  Each post-collision deviation
// (compute type CT) goes to out(integral_constant<int, c>, value) as soon
// as it is formed, so a kernel stores it at once (fewer live registers).
// Returns the density deviation and the shifted velocity the boundary
// values of the same step take.  Both stream kernels (one step per launch,
// two steps per launch) call it, so their arithmetic is the same expression
// by expression.  Which multiply-adds the compiler fuses can still differ
// between the kernels where an operand varies per node (a force field):
// results then agree to the last bit or two.  Forbidding fusion here makes
// them bit-identical but costs the one-step kernel 3 % (FP32) to 12 % (FP16),
// measured.
template <class CT, class Out>
__device__ __forceinline__ void bgk_guo(const CT (&g)[Q], CT fx, CT fy, CT fz, CT dt, CT visc_scale,
                                        int bclass, CT &drho_out, CT &ux_out, CT &uy_out,
                                        CT &uz_out, Out &&out) {
  CT drho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int c = 0; c < Q; c++) {
    drho += g[c];
    if (EX[c]) jx += EX[c] > 0 ? g[c] : -g[c];
    if (EY[c]) jy += EY[c] > 0 ? g[c] : -g[c];
    if (EZ[c]) jz += EZ[c] > 0 ? g[c] : -g[c];
  }
  // half-force velocity shift, ZebraflowWT.c:6132-6135: u = v / (1 + drho)
  // taken as v - v q with q = drho / (1 + drho)
  const CT q = drho / ((CT)1 + drho);
  const CT vx = jx + (CT)0.5 * dt * fx;
  const CT vy = jy + (CT)0.5 * dt * fy;
  const CT vz = jz + (CT)0.5 * dt * fz;
  const CT ux = vx - vx * q;
  const CT uy = vy - vy * q;
  const CT uz = vz - vz * q;
  const CT usq = ux * ux + uy * uy + uz * uz;
  const CT omega = (CT)1 / ((CT)0.5 + visc_scale * (CT)bclass);
  const CT fpref = dt * ((CT)1 - (CT)0.5 * omega);  // DT*(1 - 1/(2 tau))
  const CT uf = ux * fx + uy * fy + uz * fz;
  for_dirs([&](auto cc) {
    constexpr int c = decltype(cc)::value;
    const CT cu = (CT)EX[c] * ux + (CT)EY[c] * uy + (CT)EZ[c] * uz;
    const CT cf = (CT)EX[c] * fx + (CT)EY[c] * fy + (CT)EZ[c] * fz;
    const CT ge = geq<CT>(c, drho, ux, uy, uz, usq);
    // Guo source, ZebraflowWT.c:6140: ((c-u).F)/cs^2 + (c.u)(c.F)/cs^4
    const CT src_c = (CT)WQ[c] * ((CT)3 * (cf - uf) + (CT)9 * cu * cf);
    out(cc, g[c] - omega * (g[c] - ge) + fpref * src_c);
  });
  drho_out = drho;
  ux_out = ux;
  uy_out = uy;
  uz_out = uz;
}

struct StepParams {
  double visc_scale;  // tau = 0.5 + visc_scale*B
  double dt;          // DT of the Guo terms
  double g[3];        // uniform body force
};

struct BcParams {
  double inlet[3];        // inlet velocity vector at this iteration
  double outlet_drho;     // outlet density - 1
  double pdrho[9];        // density - 1 of codes 1001..1009 at this iteration
};

}  // namespace d3q19

// Parameters of iteration `iter` (inlet velocity, pressure waveforms).
void step_params(const zf_ctx *ctx, int iter, d3q19::StepParams &prm, d3q19::BcParams &bp);
// One stream-collide launch from src to dst (no buffer swap, force kept).
int step_into(zf_ctx *ctx, const void *src, void *dst, int iter);

}  // namespace zf

#endif
