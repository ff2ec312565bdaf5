/*
 * zf_oracle_wss.c -- wall shear stress and wall pressure at wall-mesh nodes,
 * FP64 (TEST INFRASTRUCTURE).
 *
 * Restates PostProWSSandPressure.c:
 *   window             :267-279, :325-337  index (int)((x-X0)/DX + 1.01) +- 3,
 *                                          i clipped to [3, IM] (the "else
 *                                          if" pair clips only one side),
 *                                          j, k to [1, JM], [1, KM]
 *   wall pressure      :281-315  nearest lattice node with B >= 0 closer
 *                                than 1.42 DX (strict <, (i,j,k) order);
 *                                p = (D - 1) * LBM_PRESSURE_SCALE - Pgauge;
 *                                no such node -> invalid
 *   wall shear stress  :322-369  probe x + DX n, unnormalised cosine kernel
 *                                over |d| <= 2 DX on nodes 0.11 < B < 4997,
 *                                u_t = u - (u.n) n, tau = -(mu/DX) u_t
 * The post-processor's further rejection of pressures below 1 (:317) is a
 * plotting filter and is not part of the operator.
 */
#include <math.h>
#include <stddef.h>

#include "zf_oracle.h"

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

static void wss_window(const zfo_ibm_params *p, double x, double y, double z, int lo[3], int hi[3]) {
  const int ii = (int)((x - p->x0) / p->dx + 1.01), jj = (int)((y - p->y0) / p->dx + 1.01),
            kk = (int)((z - p->z0) / p->dx + 1.01);
  lo[0] = ii - 3; hi[0] = ii + 3;
  if (hi[0] > p->im) hi[0] = p->im;
  else if (lo[0] < 3) lo[0] = 3;
  lo[1] = jj - 3; hi[1] = jj + 3;
  if (hi[1] > p->jm) hi[1] = p->jm;
  else if (lo[1] < 1) lo[1] = 1;
  lo[2] = kk - 3; hi[2] = kk + 3;
  if (hi[2] > p->km) hi[2] = p->km;
  else if (lo[2] < 1) lo[2] = 1;
  /* indices the reference would read outside its arrays are skipped */
  if (lo[0] < 1) lo[0] = 1;
  if (lo[1] < 1) lo[1] = 1;
  if (lo[2] < 1) lo[2] = 1;
}

void zfo_wss(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid, const double *D,
             const double *U, const double *V, const double *W, int64_t n, const double *xyz,
             const double *nrm, double mu_over_dx, double pressure_scale, double pgauge,
             double *wss, double *press, int32_t *valid) {
  const int jm = p->jm, km = p->km;
  const double PI = 3.14159265358979, width = 2.0 * p->dx;
#pragma omp parallel for schedule(static, 64)
  for (int64_t m = 0; m < n; m++) {
    const double *x = &xyz[3 * m], *nv = &nrm[3 * m];
    int lo[3], hi[3];
    /* wall pressure */
    wss_window(p, x[0], x[1], x[2], lo, hi);
    double mindis = 1.42 * p->dx;
    int bi = 0, bj = 0, bk = 0;
    for (int i = lo[0]; i <= hi[0]; i++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int k = lo[2]; k <= hi[2]; k++) {
          if (b[bq(i, j, k, jm, km)] < 0) continue;
          const double dx = x[0] - (p->x0 + (i - 1) * p->dx), dy = x[1] - (p->y0 + (j - 1) * p->dx),
                       dz = x[2] - (p->z0 + (k - 1) * p->dx);
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d < mindis) { mindis = d; bi = i; bj = j; bk = k; }
        }
    if (bi > 0) {
      valid[m] = 1;
      press[m] = (D[domid[bq(bi, bj, bk, jm, km)] - 1] - 1.0) * pressure_scale - pgauge;
    } else {
      valid[m] = 0;
      press[m] = 0.0;
    }
    /* wall shear stress */
    const double xp = x[0] + p->dx * nv[0], yp = x[1] + p->dx * nv[1], zp = x[2] + p->dx * nv[2];
    wss_window(p, xp, yp, zp, lo, hi);
    double u[3] = {0, 0, 0};
    for (int i = lo[0]; i <= hi[0]; i++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int k = lo[2]; k <= hi[2]; k++) {
          const double ax = xp - (p->x0 + (i - 1) * p->dx), ay = yp - (p->y0 + (j - 1) * p->dx),
                       az = zp - (p->z0 + (k - 1) * p->dx);
          if (sqrt(ax * ax + ay * ay + az * az) > width) continue;
          const int bv = b[bq(i, j, k, jm, km)];
          if (!(bv >= 1 && bv <= 4996)) continue;
          const double w = (1.0 + cos(PI * fabs(ax) / width)) * (1.0 + cos(PI * fabs(ay) / width)) *
                           (1.0 + cos(PI * fabs(az) / width)) / 64.0;
          const int o = domid[bq(i, j, k, jm, km)] - 1;
          u[0] += w * U[o];
          u[1] += w * V[o];
          u[2] += w * W[o];
        }
    const double un = u[0] * nv[0] + u[1] * nv[1] + u[2] * nv[2];
    for (int a = 0; a < 3; a++) wss[3 * m + a] = -mu_over_dx * (u[a] - un * nv[a]);
  }
}
