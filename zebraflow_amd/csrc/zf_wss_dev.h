// zf_wss_dev.h -- device helpers of the wall shear stress operator
// (PostProWSSandPressure.c:267-369), shared by the operator and the
// adjoint's wall-shear-stress objective.
#ifndef ZF_WSS_DEV_H
#define ZF_WSS_DEV_H

#include "zf_d3q19.h"

namespace zf {
namespace wssdev {

using namespace d3q19;

struct WGrid {
  double x0, y0, z0, dx;
  int im, jm, km;
};

__device__ __forceinline__ size_t wq(const WGrid &g, int i, int j, int k) {
  return ((size_t)(i - 1) * g.jm + (j - 1)) * g.km + (k - 1);
}

__device__ __forceinline__ void wss_window(const WGrid &g, double x, double y, double z, int lo[3],
                                           int hi[3]) {
  const int ii = (int)((x - g.x0) / g.dx + 1.01), jj = (int)((y - g.y0) / g.dx + 1.01),
            kk = (int)((z - g.z0) / g.dx + 1.01);
  lo[0] = ii - 3; hi[0] = ii + 3;
  if (hi[0] > g.im) hi[0] = g.im;
  else if (lo[0] < 3) lo[0] = 3;
  lo[1] = jj - 3; hi[1] = jj + 3;
  if (hi[1] > g.jm) hi[1] = g.jm;
  else if (lo[1] < 1) lo[1] = 1;
  lo[2] = kk - 3; hi[2] = kk + 3;
  if (hi[2] > g.km) hi[2] = g.km;
  else if (lo[2] < 1) lo[2] = 1;
  lo[0] = max(lo[0], 1); lo[1] = max(lo[1], 1); lo[2] = max(lo[2], 1);
}

// drho and momentum of one listed node from its stored populations.
template <int S>
__device__ __forceinline__ void node_moments(const typename Store<S>::T *__restrict__ f, size_t stride,
                                        int s, double &drho, double j[3]) {
  drho = 0;
  j[0] = j[1] = j[2] = 0;
#pragma unroll
  for (int c = 0; c < Q; c++) {
    const double g = (double)Store<S>::ld(f, (size_t)c * stride + s, WQ[c]);
    drho += g;
    j[0] += EX[c] * g;
    j[1] += EY[c] * g;
    j[2] += EZ[c] * g;
  }
}

// Calls fn(storage index, weight) for every lattice node of the probe's
// stencil: |d| <= 2 DX, 0.11 < B < 4997, unnormalised cosine weight.
template <class Fn>
__device__ __forceinline__ void wss_stencil(const WGrid &g, const int32_t *__restrict__ b,
                                            const int32_t *__restrict__ sbox, double xp, double yp,
                                            double zp, Fn &&fn) {
  const double width = 2.0 * g.dx, pi = 3.14159265358979;
  int lo[3], hi[3];
  wss_window(g, xp, yp, zp, lo, hi);
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int k = lo[2]; k <= hi[2]; k++) {
        const double ax = xp - (g.x0 + (i - 1) * g.dx), ay = yp - (g.y0 + (j - 1) * g.dx),
                     az = zp - (g.z0 + (k - 1) * g.dx);
        if (sqrt(ax * ax + ay * ay + az * az) > width) continue;
        const size_t q = wq(g, i, j, k);
        const int bv = b[q];
        if (!(bv >= 1 && bv <= 4996)) continue;
        const double w = (1.0 + cos(pi * fabs(ax) / width)) * (1.0 + cos(pi * fabs(ay) / width)) *
                         (1.0 + cos(pi * fabs(az) / width)) / 64.0;
        fn(sbox[q], w);
      }
}

}  // namespace wssdev
}  // namespace zf

#endif
