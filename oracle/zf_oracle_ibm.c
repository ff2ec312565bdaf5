/*
 * zf_oracle_ibm.c -- immersed-boundary coupling of the oracle, FP64
 * (TEST INFRASTRUCTURE).
 *
 * Restates ZebraflowWT.c (ROI branch):
 *   velocity interpolation   :8702-8996  (9^3 window, spherical cut-off
 *                                         |d| <= 2 DX, cosine kernel /64,
 *                                         fluid nodes 0 < B < 4998)
 *   force spreading          :10746-10867 (same kernel onto every B >= 0
 *                                         node, F_fluid = gamma_p v_rep + F)
 *   position update          :11975-11989 (x += v DT SOLVERSTAGGER, cell
 *                                         removed when |v| > vmax or NaN)
 * Coordinates: node (i,j,k) sits at (x0 + (i-1) dx, ...), the reference's
 * X[i][j][k] = X0 + (i-1)*DX (ZebraflowWT.c:2510).  The window index is
 * (int)((x - x0)/dx + 1.01) exactly as :8705-8707.
 */
#include <math.h>
#include <stddef.h>

#include "zf_oracle.h"

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

/* Window and weight shared by both directions; returns 0 when the node is
 * outside the spherical cut-off. */
static inline double ibm_weight(double xc, double yc, double zc, double xp, double yp,
                                double zp, double width) {
  const double dx = xc - xp, dy = yc - yp, dz = zc - zp;
  if (sqrt(dx * dx + dy * dy + dz * dz) > width) return 0.0;
  const double P = 3.14159265358979;
  return (1.0 + cos(P * fabs(dx) / width)) * (1.0 + cos(P * fabs(dy) / width)) *
         (1.0 + cos(P * fabs(dz) / width)) / 64.0;
}

void zfo_ibm_interpolate(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                         const double *U, const double *V, const double *W, int64_t nnode,
                         const double *xyz, double *vel) {
  zfo_ibm_interpolate_wall(p, b, domid, U, V, W, NULL, NULL, nnode, xyz, vel);
}

/* The window's non-fluid nodes: B == 4998 reads Uvoid (:8980-8987), any
 * other with W_index > 0 reads Uvoid with WallStencilWidth (:8988-8995),
 * which equals StencilWidth (2, :147-148), so both take the same weight. */
void zfo_ibm_interpolate_wall(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                              const double *U, const double *V, const double *W,
                              const int32_t *windex, const double *uvoid, int64_t nnode,
                              const double *xyz, double *vel) {
  const int im = p->im, jm = p->jm, km = p->km;
  const size_t box = (size_t)im * jm * km;
  const double width = 2.0 * p->dx; /* StencilWidth*DX */
#pragma omp parallel for schedule(static, 64)
  for (int64_t n = 0; n < nnode; n++) {
    const double xp = xyz[3 * n], yp = xyz[3 * n + 1], zp = xyz[3 * n + 2];
    const int i = (int)((xp - p->x0) / p->dx + 1.01), j = (int)((yp - p->y0) / p->dx + 1.01),
              k = (int)((zp - p->z0) / p->dx + 1.01);
    int i1 = i - 4, i2 = i + 4, j1 = j - 4, j2 = j + 4, k1 = k - 4, k2 = k + 4;
    i1 = i1 < 1 ? 1 : i1;
    i2 = i2 > im ? im : i2;
    j1 = j1 < 1 ? 1 : j1;
    j2 = j2 > jm ? jm : j2;
    k1 = k1 < 1 ? 1 : k1;
    k2 = k2 > km ? km : k2;
    double v[3] = {0, 0, 0};
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k1; kk <= k2; kk++) {
          const size_t q = bq(ii, jj, kk, jm, km);
          const int bv = b[q];
          if (!(bv > 0 && bv < 4998)) {
            if (!uvoid || !(bv == 4998 || windex[q] > 0)) continue;
            const double w = ibm_weight(p->x0 + (ii - 1) * p->dx, p->y0 + (jj - 1) * p->dx,
                                        p->z0 + (kk - 1) * p->dx, xp, yp, zp, width);
            if (w == 0.0) continue;
            for (int a = 0; a < 3; a++) v[a] += w * p->vel_scale * uvoid[(size_t)a * box + q];
            continue;
          }
          const double w = ibm_weight(p->x0 + (ii - 1) * p->dx, p->y0 + (jj - 1) * p->dx,
                                      p->z0 + (kk - 1) * p->dx, xp, yp, zp, width);
          if (w == 0.0) continue;
          const int o = domid[q] - 1;
          v[0] += w * p->vel_scale * U[o];
          v[1] += w * p->vel_scale * V[o];
          v[2] += w * p->vel_scale * W[o];
        }
    vel[3 * n] = v[0];
    vel[3 * n + 1] = v[1];
    vel[3 * n + 2] = v[2];
  }
}

/* Serial in node order (the reference's float atomics give no order). */
void zfo_ibm_spread(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                    int64_t dommax, int64_t nnode, const double *xyz, const double *ffluid,
                    double *force) {
  const int im = p->im, jm = p->jm, km = p->km;
  const double width = 2.0 * p->dx;
  for (int64_t n = 0; n < nnode; n++) {
    const double xp = xyz[3 * n], yp = xyz[3 * n + 1], zp = xyz[3 * n + 2];
    const int i = (int)((xp - p->x0) / p->dx + 1.01), j = (int)((yp - p->y0) / p->dx + 1.01),
              k = (int)((zp - p->z0) / p->dx + 1.01);
    int i1 = i - 4, i2 = i + 4, j1 = j - 4, j2 = j + 4, k1 = k - 4, k2 = k + 4;
    i1 = i1 < 1 ? 1 : i1;
    i2 = i2 > im ? im : i2;
    j1 = j1 < 1 ? 1 : j1;
    j2 = j2 > jm ? jm : j2;
    k1 = k1 < 1 ? 1 : k1;
    k2 = k2 > km ? km : k2;
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k1; kk <= k2; kk++) {
          const size_t q = bq(ii, jj, kk, jm, km);
          if (b[q] < 0) continue;
          const double w = ibm_weight(p->x0 + (ii - 1) * p->dx, p->y0 + (jj - 1) * p->dx,
                                      p->z0 + (kk - 1) * p->dx, xp, yp, zp, width);
          if (w == 0.0) continue;
          const int o = domid[q] - 1;
          for (int a = 0; a < 3; a++) force[a * dommax + o] += w * ffluid[3 * n + a] * p->force_scale;
        }
  }
}

/* The same spreading with the reference's parallel structure, for the CPU
 * baseline's timing only: nodes shared among threads, every lattice-node
 * update an atomic add (ZebraflowWT.c:10746-10867, its float atomics at
 * :10857), so the sums' order is the threads' (results agree with
 * zfo_ibm_spread to rounding). */
void zfo_ibm_spread_omp(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                        int64_t dommax, int64_t nnode, const double *xyz, const double *ffluid,
                        double *force) {
  const int im = p->im, jm = p->jm, km = p->km;
  const double width = 2.0 * p->dx;
#pragma omp parallel for schedule(static, 50)
  for (int64_t n = 0; n < nnode; n++) {
    const double xp = xyz[3 * n], yp = xyz[3 * n + 1], zp = xyz[3 * n + 2];
    const int i = (int)((xp - p->x0) / p->dx + 1.01), j = (int)((yp - p->y0) / p->dx + 1.01),
              k = (int)((zp - p->z0) / p->dx + 1.01);
    int i1 = i - 4, i2 = i + 4, j1 = j - 4, j2 = j + 4, k1 = k - 4, k2 = k + 4;
    i1 = i1 < 1 ? 1 : i1;
    i2 = i2 > im ? im : i2;
    j1 = j1 < 1 ? 1 : j1;
    j2 = j2 > jm ? jm : j2;
    k1 = k1 < 1 ? 1 : k1;
    k2 = k2 > km ? km : k2;
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k1; kk <= k2; kk++) {
          const size_t q = bq(ii, jj, kk, jm, km);
          if (b[q] < 0) continue;
          const double w = ibm_weight(p->x0 + (ii - 1) * p->dx, p->y0 + (jj - 1) * p->dx,
                                      p->z0 + (kk - 1) * p->dx, xp, yp, zp, width);
          if (w == 0.0) continue;
          const int o = domid[q] - 1;
          for (int a = 0; a < 3; a++) {
#pragma omp atomic
            force[a * dommax + o] += w * ffluid[3 * n + a] * p->force_scale;
          }
        }
  }
}

/* Returns the number of cells removed.  status < 0 cells are skipped. */
int zfo_advance(const zfo_ibm_params *p, int ncell, int nn, double *xyz, const double *vel,
                int32_t *status) {
  int removed = 0;
  for (int c = 0; c < ncell; c++) {
    if (status[c] < 0) continue;
    for (int n = 0; n < nn; n++) {
      const double *v = &vel[3 * ((size_t)c * nn + n)];
      const double s = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (s > p->vmax || isnan(s) || isinf(s)) {
        status[c] = -5000;
        removed++;
        break;
      }
      for (int a = 0; a < 3; a++) xyz[3 * ((size_t)c * nn + n) + a] += v[a] * p->dt_advance;
    }
  }
  return removed;
}
