/*
 * zf_oracle_wrap.c -- cells on a parent (feeder) lattice, FP64 (TEST
 * INFRASTRUCTURE).
 *
 * The reference's parent lattices carry their cells through a periodic
 * window along k: a parent's cells keep their node coordinates in
 * [Z0_p, ZM_p] (ZM_p = Z0_p + (KM_p - 1) DX) and each node wraps by
 * L = ZM_p - Z0_p when it leaves (ZebraflowWT.c:12196-12340).  Its stages
 * address the lattice with k taken modulo KM_p - 1.  Restated here, each
 * from its DAparent branch (parents 2 and 3 are the same code with the
 * other direction):
 *   interpolation + wall repulsion  :7541-7732 (status 2), :8101-8297
 *                                   (status > 2: window centred at zpmap)
 *   cell list                       :5898-5950
 *   cell-cell repulsion             :10204-10383
 *   spreading                       :11184-11310, :11547-11669
 *   position update and wrap        :12196-12340
 * The per-cell quantities (span, unwrapping of cells across the wrap,
 * REGRESSFLAG, cloning) are restated in oracle/pyoracle.py.
 */
#include <math.h>
#include <stdlib.h>

#include "zf_oracle.h"

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* kmap of the interpolation and spreading windows (:7574-7576):
 * below 1 add KM-1, above KM subtract KM-1 (plane KM is kept). */
static inline int kmap_ibm(int k, int km) {
  return k < 1 ? k + (km - 1) : k > km ? k - (km - 1) : k;
}

static inline double cosk(double d, double width) {
  const double P = 3.14159265358979;
  return 1.0 + cos(P * fabs(d) / width);
}

/* zpmap (:8110-8115): the crossing cell's nodes behind the middle of the
 * window seen past its exit end. */
static inline double zpmap(const zfo_wrap *w, double zp) {
  const double L = w->zm - w->z0, mid = 0.5 * (w->z0 + w->zm);
  if (w->dir > 0) return zp < mid ? zp + L : zp;
  return zp > mid ? zp - L : zp;
}

void zfo_wrap_interpolate(const zfo_ibm_params *p, const zfo_wrap *w, const int32_t *b,
                          const int32_t *domid, const double *U, const double *V, const double *W,
                          const int32_t *windex, const double *uvoid, int ncell, int nn,
                          const int32_t *status, const double *xyz, double *vel) {
  const int im = p->im, jm = p->jm, km = p->km;
  const size_t box = (size_t)im * jm * km;
  const double width = 2.0 * p->dx; /* StencilWidth*DX = WallStencilWidth*DX */
#pragma omp parallel for schedule(static, 64)
  for (int64_t n = 0; n < (int64_t)ncell * nn; n++) {
    const int c = (int)(n / nn);
    vel[3 * n] = vel[3 * n + 1] = vel[3 * n + 2] = 0.0;
    if (status[c] < 2) continue;
    const double xp = xyz[3 * n], yp = xyz[3 * n + 1], zp = xyz[3 * n + 2];
    const double ze = status[c] > 2 ? zpmap(w, zp) : zp;
    const int i = (int)((xp - p->x0) / p->dx + 1.01), j = (int)((yp - p->y0) / p->dx + 1.01),
              k = (int)((ze - p->z0) / p->dx + 1.01);
    const int i1 = clampi(i - 4, 1, im), i2 = clampi(i + 4, 1, im);
    const int j1 = clampi(j - 4, 1, jm), j2 = clampi(j + 4, 1, jm);
    double v[3] = {0, 0, 0};
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k - 4; kk <= k + 4; kk++) {
          const double xc = p->x0 + (ii - 1) * p->dx, yc = p->y0 + (jj - 1) * p->dx,
                       zc = p->z0 + (kk - 1) * p->dx;
          const size_t q = bq(ii, jj, kmap_ibm(kk, km), jm, km);
          if (b[q] > 0) {
            /* cut-off at zpmap, the z factor at zp (:8145-8147: the same
             * value, the kernel being periodic in L / (2 DX) even) */
            const double dd = sqrt((xc - xp) * (xc - xp) + (yc - yp) * (yc - yp) + (zc - ze) * (zc - ze));
            if (dd > width) continue;
            const double wt = cosk(xc - xp, width) * cosk(yc - yp, width) * cosk(zc - zp, width) / 64.0;
            const int o = domid[q] - 1;
            v[0] += wt * p->vel_scale * U[o];
            v[1] += wt * p->vel_scale * V[o];
            v[2] += wt * p->vel_scale * W[o];
          } else if (uvoid && windex[q] > 0) {
            /* the wall layer's cut-off is taken at zp (:8154), so a node
             * seen through zpmap gets no Uvoid term */
            const double dd = sqrt((xc - xp) * (xc - xp) + (yc - yp) * (yc - yp) + (zc - zp) * (zc - zp));
            if (dd > width) continue;
            const double wt = cosk(xc - xp, width) * cosk(yc - yp, width) * cosk(zc - zp, width) / 64.0;
            for (int a = 0; a < 3; a++) v[a] += wt * p->vel_scale * uvoid[(size_t)a * box + q];
          }
        }
    vel[3 * n] = v[0];
    vel[3 * n + 1] = v[1];
    vel[3 * n + 2] = v[2];
  }
}

void zfo_wrap_wall_repulsion(const zfo_ibm_params *p, const zfo_wrap *w, const zfo_rep_params *r,
                             const int32_t *windex, const double *wall_xyz, const int32_t *wall_tri,
                             const double *wall_normal, const int32_t *wall_node_tri, int ncell,
                             int nn, const int32_t *status, const double *xyz, double *vrep,
                             int32_t *nearest) {
  const int im = p->im, jm = p->jm, km = p->km;
#pragma omp parallel for schedule(static, 16)
  for (int64_t g = 0; g < (int64_t)ncell * nn; g++) {
    const int c = (int)(g / nn);
    nearest[g] = 0;
    if (status[c] < 2) continue;
    const double *x = &xyz[3 * g];
    const double ze = status[c] > 2 ? zpmap(w, x[2]) : x[2];
    const int i = (int)((x[0] - p->x0) / p->dx + 1.01), j = (int)((x[1] - p->y0) / p->dx + 1.01),
              k = (int)((ze - p->z0) / p->dx + 1.01);
    const int i1 = clampi(i - 4, 1, im), i2 = clampi(i + 4, 1, im);
    const int j1 = clampi(j - 4, 1, jm), j2 = clampi(j + 4, 1, jm);
    double mindis = r->wall_search;
    int triwall = 0;
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k - 4; kk <= k + 4; kk++) {
          const int wn = windex[bq(ii, jj, kmap_ibm(kk, km), jm, km)];
          if (wn <= 0) continue;
          const int32_t *row = &wall_node_tri[(size_t)(wn - 1) * 13];
          for (int l = 1; l <= row[0] && l <= 12; l++) {
            const int t = row[l];
            const int32_t *tv = &wall_tri[3 * (size_t)(t - 1)];
            double cen[3] = {0, 0, 0};
            for (int s = 0; s < 3; s++) {
              const double *wp = &wall_xyz[3 * (size_t)(tv[s] - 1)];
              const double d = sqrt((x[0] - wp[0]) * (x[0] - wp[0]) + (x[1] - wp[1]) * (x[1] - wp[1]) +
                                    (x[2] - wp[2]) * (x[2] - wp[2]));
              if (d < mindis) { mindis = d; triwall = t; }
              for (int a = 0; a < 3; a++) cen[a] += wp[a] / 3.0;
            }
            const double d = sqrt((x[0] - cen[0]) * (x[0] - cen[0]) + (x[1] - cen[1]) * (x[1] - cen[1]) +
                                  (x[2] - cen[2]) * (x[2] - cen[2]));
            if (d < mindis) { mindis = d; triwall = t; }
          }
        }
    if (triwall == 0) continue;
    nearest[g] = triwall;
    const double *w1 = &wall_xyz[3 * (size_t)(wall_tri[3 * (size_t)(triwall - 1)] - 1)];
    const double *n = &wall_normal[3 * (size_t)(triwall - 1)];
    const double d = (x[0] - w1[0]) * n[0] + (x[1] - w1[1]) * n[1] + (x[2] - w1[2]) * n[2];
    if (d < r->wall_d0 && fabs(d) < r->wall_dmax) {
      double f = (r->wall_d0 - d) / r->wall_dmax;
      if (d < r->wall_d0 - r->wall_dmax) f = 1.0;
      for (int a = 0; a < 3; a++) vrep[3 * g + a] += f * r->k_wall * n[a];
    }
  }
}

/* Cell list (:5898-5950): a triangle of a cell that is not in status 2
 * whose centroid lies past the exit is binned from inside the window; the
 * window's k wraps with the neighbour translated by L (:5913-5926); the
 * nearest listed node wins with no cut-off. */
void zfo_wrap_cell_list(const zfo_ibm_params *p, const zfo_wrap *w, const int32_t *b,
                        const int32_t *domid, int ncell, int nt, const int32_t *status,
                        const double *tri_centroid, int32_t *bin) {
  const int im = p->im, jm = p->jm, km = p->km;
  const double L = w->zm - w->z0;
#pragma omp parallel for schedule(static, 64)
  for (int64_t q = 0; q < (int64_t)ncell * nt; q++) {
    const int c = (int)(q / nt);
    bin[q] = 0;
    if (status[c] == -5000) continue;
    const double xp = tri_centroid[3 * q], yp = tri_centroid[3 * q + 1];
    double zp = tri_centroid[3 * q + 2];
    if (status[c] != 2) {
      if (w->dir > 0 && zp >= w->zm) zp -= L;
      if (w->dir < 0 && zp <= w->z0) zp += L;
    }
    const int i = (int)((xp - p->x0) / p->dx + 1.01), j = (int)((yp - p->y0) / p->dx + 1.01),
              k = (int)((zp - p->z0) / p->dx + 1.01);
    const int i1 = clampi(i - 4, 1, im), i2 = clampi(i + 4, 1, im);
    const int j1 = clampi(j - 4, 1, jm), j2 = clampi(j + 4, 1, jm);
    double mindis = 5000.0;
    int bi = 0, bj = 0, bk = 0;
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k - 4; kk <= k + 4; kk++) {
          const int kkk = kk > km ? kk - km + 1 : kk < 1 ? kk + km - 1 : kk;
          if (b[bq(ii, jj, kkk, jm, km)] < 0) continue;
          /* the neighbour at its translated place, z0 + (kk - 1) dx */
          const double dx = p->x0 + (ii - 1) * p->dx - xp, dy = p->y0 + (jj - 1) * p->dx - yp,
                       dz = p->z0 + (kk - 1) * p->dx - zp;
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d < mindis) { mindis = d; bi = ii; bj = jj; bk = kkk; }
        }
    if (bi) bin[q] = domid[bq(bi, bj, bk, jm, km)];
  }
}

/* 16 sub-triangle centroids (:10307-10345), as zf_oracle_cells.c. */
static void sub16(const double *p1, const double *p2, const double *p3, double tc[16][3]) {
  for (int v = 0; v < 3; v++) {
    double tp[16];
    tp[1] = p1[v]; tp[2] = p2[v]; tp[3] = p3[v];
    tp[4] = 0.5 * (tp[1] + tp[2]); tp[5] = 0.5 * (tp[2] + tp[3]); tp[6] = 0.5 * (tp[1] + tp[3]);
    tp[7] = 0.5 * (tp[1] + tp[4]); tp[8] = 0.5 * (tp[2] + tp[4]); tp[9] = 0.5 * (tp[2] + tp[5]);
    tp[10] = 0.5 * (tp[5] + tp[3]); tp[11] = 0.5 * (tp[3] + tp[6]); tp[12] = 0.5 * (tp[6] + tp[1]);
    tp[13] = 0.5 * (tp[4] + tp[6]); tp[14] = 0.5 * (tp[4] + tp[5]); tp[15] = 0.5 * (tp[5] + tp[6]);
    const int T[16][3] = {{1, 7, 12},  {7, 13, 12}, {7, 4, 13},   {4, 14, 13},
                          {4, 8, 14},  {8, 9, 14},  {8, 2, 9},    {12, 13, 6},
                          {13, 15, 6}, {13, 14, 15}, {14, 5, 15}, {14, 9, 5},
                          {6, 15, 11}, {15, 10, 11}, {15, 5, 10}, {11, 10, 3}};
    for (int s = 0; s < 16; s++) tc[s][v] = (tp[T[s][0]] + tp[T[s][1]] + tp[T[s][2]]) / 3.0;
  }
}

/* Cell-cell repulsion among the parent's cells (:10204-10383).  The window
 * is centred at the node's own coordinate; its planes k >= KM read plane
 * k - KM + 1 (so plane KM itself is read as plane 1) and planes k < 1 read
 * k + KM - 1 (checks 1 and 2); otherwise a node in the last or first
 * quarter of the window is check 3 or 4.  A candidate's centroid is
 * translated by L across a wrapped plane (check 1: centroid behind the
 * middle, check 2: past it), and its vertices by the quarter rules
 * (checks 1 and 3: vertices in the first quarter +L; 2 and 4: in the last
 * quarter -L). */
void zfo_wrap_cell_repulsion(const zfo_ibm_params *p, const zfo_wrap *w, const zfo_rep_params *r,
                             const int32_t *b, const int32_t *domid, int ncell, int nn, int nt,
                             const int32_t *tri, const int32_t *status, const double *xyz,
                             const double *node_normal, const double *tri_centroid,
                             const double *tri_normal, const int32_t *bin, double *vrep) {
  (void)b;
  const int im = p->im, jm = p->jm, km = p->km;
  const double L = w->zm - w->z0, mid = 0.5 * (w->z0 + w->zm);
  const double q1 = w->z0 + 0.25 * L, q3 = w->z0 + 0.75 * L;
  int dmax = 0;
  for (size_t q = 0; q < (size_t)ncell * nt; q++) dmax = bin[q] > dmax ? bin[q] : dmax;
  int *start = calloc((size_t)dmax + 2, sizeof(int));
  int *ent = malloc(sizeof(int) * ((size_t)ncell * nt + 1));
  int *fill = calloc((size_t)dmax + 1, sizeof(int));
  for (size_t q = 0; q < (size_t)ncell * nt; q++)
    if (bin[q] > 0) start[bin[q] + 1]++;
  for (int d = 0; d <= dmax; d++) start[d + 1] += start[d];
  for (size_t q = 0; q < (size_t)ncell * nt; q++)
    if (bin[q] > 0) ent[start[bin[q]] + fill[bin[q]]++] = (int)q;
#pragma omp parallel for schedule(static, 16)
  for (int g = 0; g < ncell * nn; g++) {
    const int c = g / nn, node = g % nn;
    if (status[c] == -5000) continue;
    const double *x = &xyz[3 * (size_t)g], *e1 = &node_normal[3 * (size_t)g];
    const int ic = (int)((x[0] - p->x0) / p->dx + 1.01), jc = (int)((x[1] - p->y0) / p->dx + 1.01),
              kc = (int)((x[2] - p->z0) / p->dx + 1.01);
    const int i1 = clampi(ic - 4, 1, im), i2 = clampi(ic + 4, 1, im);
    const int j1 = clampi(jc - 4, 1, jm), j2 = clampi(jc + 4, 1, jm);
    double v[3] = {0, 0, 0};
    for (int i = i1; i <= i2; i++)
      for (int j = j1; j <= j2; j++)
        for (int k = kc - 4; k <= kc + 4; k++) {
          int check = 0, kk = k;
          if (k >= km) { kk = k - km + 1; check = 1; }
          else if (k < 1) { kk = k + km - 1; check = 2; }
          else if (x[2] >= q3) check = 3;
          else if (x[2] < q1) check = 4;
          const int o = domid[bq(i, j, kk, jm, km)];
          if (o <= 0 || o > dmax) continue;
          for (int e = start[o]; e < start[o + 1]; e++) {
            const int q = ent[e], m = q / nt, t = q % nt;
            if (status[m] == -5000) continue;
            if (m == c && (tri[3 * t] - 1 == node || tri[3 * t + 1] - 1 == node ||
                           tri[3 * t + 2] - 1 == node))
              continue;
            const double *tcen = &tri_centroid[3 * (size_t)q], *tn = &tri_normal[3 * (size_t)q];
            double zp = tcen[2];
            if (check == 1 && zp < mid) zp += L;
            if (check == 2 && zp > mid) zp -= L;
            if (e1[0] * tn[0] + e1[1] * tn[1] + e1[2] * tn[2] >= 0) continue;
            const double dx = x[0] - tcen[0], dy = x[1] - tcen[1], dz = x[2] - zp;
            if (sqrt(dx * dx + dy * dy + dz * dz) >= r->search) continue;
            double vv[3][3];
            for (int s = 0; s < 3; s++) {
              const double *pv = &xyz[3 * ((size_t)m * nn + tri[3 * t + s] - 1)];
              vv[s][0] = pv[0];
              vv[s][1] = pv[1];
              vv[s][2] = pv[2];
              if ((check == 1 || check == 3) && vv[s][2] < q1) vv[s][2] += L;
              else if ((check == 2 || check == 4) && vv[s][2] >= q3) vv[s][2] -= L;
            }
            double tc[16][3];
            sub16(vv[0], vv[1], vv[2], tc);
            for (int s = 0; s < 16; s++) {
              const double ex = tc[s][0] - x[0], ey = tc[s][1] - x[1], ez = tc[s][2] - x[2];
              const double d = sqrt(ex * ex + ey * ey + ez * ez);
              if (d < r->cutoff) {
                const double f = -r->k_cell * (r->cutoff - d) / r->cutoff;
                v[0] += f * ex / d;
                v[1] += f * ey / d;
                v[2] += f * ez / d;
              }
            }
          }
        }
    for (int a = 0; a < 3; a++) vrep[3 * (size_t)g + a] += v[a];
  }
  free(start);
  free(ent);
  free(fill);
}

/* Spreading (:11184-11310 and, for status > 2, :11547-11669, which centres
 * the window at the node's own coordinate as well): every node of a live
 * cell onto the listed nodes (B >= 0) of its window, k through kmap. */
void zfo_wrap_spread(const zfo_ibm_params *p, const zfo_wrap *w, const int32_t *b,
                     const int32_t *domid, int64_t dommax, int ncell, int nn,
                     const int32_t *status, const double *xyz, const double *ffluid, double *force) {
  (void)w;
  const int im = p->im, jm = p->jm, km = p->km;
  const double width = 2.0 * p->dx;
  for (int64_t n = 0; n < (int64_t)ncell * nn; n++) {
    if (status[n / nn] < 2) continue;
    const double xp = xyz[3 * n], yp = xyz[3 * n + 1], zp = xyz[3 * n + 2];
    const int i = (int)((xp - p->x0) / p->dx + 1.01), j = (int)((yp - p->y0) / p->dx + 1.01),
              k = (int)((zp - p->z0) / p->dx + 1.01);
    const int i1 = clampi(i - 4, 1, im), i2 = clampi(i + 4, 1, im);
    const int j1 = clampi(j - 4, 1, jm), j2 = clampi(j + 4, 1, jm);
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k - 4; kk <= k + 4; kk++) {
          const size_t q = bq(ii, jj, kmap_ibm(kk, km), jm, km);
          if (b[q] < 0) continue;
          const double xc = p->x0 + (ii - 1) * p->dx, yc = p->y0 + (jj - 1) * p->dx,
                       zc = p->z0 + (kk - 1) * p->dx;
          const double dd = sqrt((xc - xp) * (xc - xp) + (yc - yp) * (yc - yp) + (zc - zp) * (zc - zp));
          if (dd > width) continue;
          const double wt = cosk(xc - xp, width) * cosk(yc - yp, width) * cosk(zc - zp, width) / 64.0;
          const int o = domid[q] - 1;
          for (int a = 0; a < 3; a++) force[a * dommax + o] += wt * ffluid[3 * n + a] * p->force_scale;
        }
  }
}

/* Position update of the parent's cells (:12196-12340), membrane nodes then
 * nucleus nodes (xyzn may be NULL).  Status 2: each node moves and wraps,
 * and a node leaving through the exit flags the cell for cloning unless
 * its REGRESSFLAG is set.  Status 4: the same wrap without the flag, its
 * REGRESSFLAG set, and back to status 2 once its membrane lies inside the
 * window's first three quarters and 5 from the entry. */
void zfo_wrap_advance(const zfo_wrap *w, double dt, int ncell, int nn, int nnuc, double *xyz,
                      const double *vel, double *xyzn, const double *veln, int32_t *status,
                      int32_t *regress, int32_t *cloneflag) {
  const double L = w->zm - w->z0;
  for (int c = 0; c < ncell; c++) {
    cloneflag[c] = 0;
    const int st = status[c];
    if (st != 2 && st != 4) continue;
    if (st == 4) regress[c] = 1;
    double nzmax = -5000.0, nzmin = 5000.0;
    for (int part = 0; part < 2; part++) {
      const int np = part ? nnuc : nn;
      double *X = part ? xyzn : xyz;
      const double *Vv = part ? veln : vel;
      if (!X || np <= 0) continue;
      for (int n = 0; n < np; n++) {
        double *x = &X[3 * ((size_t)c * np + n)];
        const double *v = &Vv[3 * ((size_t)c * np + n)];
        for (int a = 0; a < 3; a++) x[a] += v[a] * dt;
        if (w->dir > 0) {
          if (x[2] >= w->zm) {
            if (st == 2 && regress[c] == 0) cloneflag[c] = 1;
            x[2] -= L;
          } else if (x[2] < w->z0) {
            x[2] += L;
          }
        } else {
          if (x[2] <= w->z0) {
            if (st == 2 && regress[c] == 0) cloneflag[c] = 1;
            x[2] += L;
          } else if (x[2] > w->zm) {
            x[2] -= L;
          }
        }
        if (!part) {
          if (x[2] > nzmax) nzmax = x[2];
          if (x[2] < nzmin) nzmin = x[2];
        }
      }
    }
    if (st == 4) {
      if (w->dir > 0 && nzmin > w->z0 + 5.0 && nzmax < w->z0 + 0.75 * L) status[c] = 2;
      if (w->dir < 0 && nzmax < w->zm - 5.0 && nzmin > w->zm - 0.75 * L) status[c] = 2;
    }
  }
}
