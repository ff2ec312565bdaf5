/*
 * zf_oracle_cells.c -- triangle cell list and cell-cell / cell-wall
 * repulsion of the oracle, FP64 (TEST INFRASTRUCTURE).
 *
 * Restates ZebraflowWT.c (ROI branch):
 *   cell list (EulNeigh)      :5851-5900  nearest listed node (B >= 0) to
 *                                          each triangle centroid in the 9^3
 *                                          window, kept if closer than the
 *                                          bin cut-off (1 um = 2 DX)
 *   cell-cell repulsion       :9913-10052 per node, over the window's bins:
 *                                          skip own adjacent triangles, keep
 *                                          facing ones (n_node . n_tri < 0)
 *                                          within the search radius, 16
 *                                          sub-triangle centroids, v_rep +=
 *                                          -k (r - d)/r e for d < r
 *   cell-wall repulsion       :8998-9142  nearest wall triangle by vertex or
 *                                          centroid distance (< 1.5) through
 *                                          the W_index map, signed plane
 *                                          distance d: v_rep += k (d0 - d) n
 *                                          if d < d0 and |d| < dmax
 * The reference's per-bin order depends on thread scheduling (atomic
 * capture, :5888); the oracle appends in (cell, triangle) order, and lists
 * are compared as per-bin multisets.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "zf_oracle.h"

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

static void window(const zfo_ibm_params *p, double x, double y, double z, int w, int lo[3], int hi[3]) {
  const int i = (int)((x - p->x0) / p->dx + 1.01), j = (int)((y - p->y0) / p->dx + 1.01),
            k = (int)((z - p->z0) / p->dx + 1.01);
  lo[0] = i - w < 1 ? 1 : i - w; hi[0] = i + w > p->im ? p->im : i + w;
  lo[1] = j - w < 1 ? 1 : j - w; hi[1] = j + w > p->jm ? p->jm : j + w;
  lo[2] = k - w < 1 ? 1 : k - w; hi[2] = k + w > p->km ? p->km : k + w;
}

/* bin[c*nt + t] = 1-based DOMID of the triangle's bin, or 0. */
void zfo_cell_list(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid, int ncell,
                   int nt, const int32_t *status, const double *tri_centroid, double bin_cutoff,
                   int32_t *bin) {
  /* each triangle's bin is its own nearest-node search: parallel over
   * triangles, as the reference's per-cell OpenMP loop (:5851-5900) */
#pragma omp parallel for schedule(static, 64)
  for (int64_t qq = 0; qq < (int64_t)ncell * nt; qq++) {
      const int c = (int)(qq / nt);
      const size_t q = (size_t)qq;
      bin[q] = 0;
      if (status[c] < 0) continue;
      const double *x = &tri_centroid[3 * q];
      int lo[3], hi[3];
      window(p, x[0], x[1], x[2], 4, lo, hi);
      double mindis = 5000.0;
      int bi = 0, bj = 0, bk = 0;
      for (int i = lo[0]; i <= hi[0]; i++)
        for (int j = lo[1]; j <= hi[1]; j++)
          for (int k = lo[2]; k <= hi[2]; k++) {
            if (b[bq(i, j, k, p->jm, p->km)] < 0) continue;
            const double dx = p->x0 + (i - 1) * p->dx - x[0], dy = p->y0 + (j - 1) * p->dx - x[1],
                         dz = p->z0 + (k - 1) * p->dx - x[2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            if (d < mindis) { mindis = d; bi = i; bj = j; bk = k; }
          }
      if (mindis < bin_cutoff) bin[q] = domid[bq(bi, bj, bk, p->jm, p->km)];
  }
}

/* 16 sub-triangle centroids of the 4-times refined triangle (:9990-10025). */
static void sub_centroids(const double *p1, const double *p2, const double *p3, double tc[16][3]) {
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

void zfo_cell_repulsion(const zfo_ibm_params *p, const zfo_rep_params *r, const int32_t *b,
                        const int32_t *domid, int ncell, int nn, int nt, const int32_t *tri,
                        const int32_t *status, const double *xyz, const double *node_normal,
                        const double *tri_centroid, const double *tri_normal, const int32_t *bin,
                        double *vrep) {
  /* bins as per-node lists (CSR over DOMID) in (cell, triangle) order */
  int dmax = 0;
  for (size_t q = 0; q < (size_t)ncell * nt; q++) dmax = bin[q] > dmax ? bin[q] : dmax;
  int *start = calloc((size_t)dmax + 2, sizeof(int));
  int *ent = malloc(sizeof(int) * ((size_t)ncell * nt + 1));
  for (size_t q = 0; q < (size_t)ncell * nt; q++)
    if (bin[q] > 0) start[bin[q] + 1]++;
  for (int d = 0; d <= dmax; d++) start[d + 1] += start[d];
  int *fill = calloc((size_t)dmax + 1, sizeof(int));
  for (size_t q = 0; q < (size_t)ncell * nt; q++)
    if (bin[q] > 0) ent[start[bin[q]] + fill[bin[q]]++] = (int)q;
#pragma omp parallel for schedule(static, 16)
  for (int g = 0; g < ncell * nn; g++) {
    const int c = g / nn, node = g % nn;
    if (status[c] <= 0) continue;
    const double *x = &xyz[3 * (size_t)g], *e1 = &node_normal[3 * (size_t)g];
    int lo[3], hi[3];
    window(p, x[0], x[1], x[2], 4, lo, hi);
    double v[3] = {0, 0, 0};
    for (int i = lo[0]; i <= hi[0]; i++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int k = lo[2]; k <= hi[2]; k++) {
          const int o = domid[bq(i, j, k, p->jm, p->km)];
          if (o <= 0 || o > dmax) continue;
          for (int e = start[o]; e < start[o + 1]; e++) {
            const int q = ent[e], m = q / nt, t = q % nt;
            if (m == c && (tri[3 * t] - 1 == node || tri[3 * t + 1] - 1 == node ||
                           tri[3 * t + 2] - 1 == node))
              continue;
            const double *tn = &tri_normal[3 * (size_t)q], *tcen = &tri_centroid[3 * (size_t)q];
            if (e1[0] * tn[0] + e1[1] * tn[1] + e1[2] * tn[2] >= 0) continue;
            const double dx = x[0] - tcen[0], dy = x[1] - tcen[1], dz = x[2] - tcen[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) >= r->search) continue;
            const double *q1 = &xyz[3 * ((size_t)m * nn + tri[3 * t] - 1)];
            const double *q2 = &xyz[3 * ((size_t)m * nn + tri[3 * t + 1] - 1)];
            const double *q3 = &xyz[3 * ((size_t)m * nn + tri[3 * t + 2] - 1)];
            double tc[16][3];
            sub_centroids(q1, q2, q3, tc);
            for (int s = 0; s < 16; s++) {
              const double ax = tc[s][0] - x[0], ay = tc[s][1] - x[1], az = tc[s][2] - x[2];
              const double d = sqrt(ax * ax + ay * ay + az * az);
              if (d < r->cutoff) {
                const double f = -r->k_cell * (r->cutoff - d) / r->cutoff;
                v[0] += f * ax / d;
                v[1] += f * ay / d;
                v[2] += f * az / d;
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

/* Wall connectivity (ZebraflowWT.c:3900-3935): rows of 13 ints (count, up
 * to 12 triangles, 1-based, in triangle order). */
void zfo_wall_node_tri(int nwn, int nwt, const int32_t *wall_tri, int32_t *wall_node_tri) {
  memset(wall_node_tri, 0, sizeof(int32_t) * 13 * (size_t)nwn);
  for (int m = 1; m <= nwt; m++)
    for (int s = 0; s < 3; s++) {
      int32_t *row = &wall_node_tri[13 * (size_t)(wall_tri[3 * (size_t)(m - 1) + s] - 1)];
      row[0]++;
      if (row[0] <= 12) row[row[0]] = m;
    }
}

/* W_index (ZebraflowWT.c:3672-3826): every lattice node within the 9^3
 * window of a wall node takes the nearest wall node; distances compared as
 * FP32 values, ties to the lower wall node (the reference's serial strict <
 * over FP32 sqrtf). */
void zfo_windex(const zfo_ibm_params *p, int nwn, const double *wall_xyz, int32_t *windex) {
  const size_t box = (size_t)p->im * p->jm * p->km;
  float *wd = malloc(sizeof(float) * box);
  zfo_windex_dist(p, nwn, wall_xyz, windex, wd);
  free(wd);
}

void zfo_windex_dist(const zfo_ibm_params *p, int nwn, const double *wall_xyz, int32_t *windex,
                     float *wd) {
  const size_t box = (size_t)p->im * p->jm * p->km;
  for (size_t q = 0; q < box; q++) { wd[q] = 5000.0f; windex[q] = 0; }
  for (int l = 1; l <= nwn; l++) {
    const double *w = &wall_xyz[3 * (size_t)(l - 1)];
    int lo[3], hi[3];
    window(p, w[0], w[1], w[2], 4, lo, hi);
    for (int i = lo[0]; i <= hi[0]; i++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int k = lo[2]; k <= hi[2]; k++) {
          const double dx = w[0] - (p->x0 + (i - 1) * p->dx), dy = w[1] - (p->y0 + (j - 1) * p->dx),
                       dz = w[2] - (p->z0 + (k - 1) * p->dx);
          const float d = (float)sqrt(dx * dx + dy * dy + dz * dz);
          const size_t q = bq(i, j, k, p->jm, p->km);
          if (d < wd[q]) { wd[q] = d; windex[q] = l; }
        }
  }
}

/* ZebraflowWT.c:4788-4851, serially in (i, j, k) order over k = k_lo ..
 * KM - k_hi.  The reference's float coordinates: distances from the
 * selection point to the FP32 wall node coordinates and to the FP32
 * centroid (x1 + x2 + x3)/3 (:4660), strict <, mindis from 5000; the ramp
 * weight1 of its FP32 Wdistance. */
void zfo_uvoid(const zfo_ibm_params *p, const int32_t *b, const int32_t *windex, const float *wdist,
               const double *wall_xyz, const int32_t *wall_tri, const double *wall_normal,
               const int32_t *wall_node_tri, double speed, double d0, double d1, int k_lo, int k_hi,
               int select, const double *sp, double *uvoid) {
  const size_t box = (size_t)p->im * p->jm * p->km;
  memset(uvoid, 0, sizeof(double) * 3 * box);
  for (int i = 1; i <= p->im; i++)
    for (int j = 1; j <= p->jm; j++)
      for (int k = k_lo; k <= p->km - k_hi; k++) {
        const size_t q = bq(i, j, k, p->jm, p->km);
        if (windex[q] <= 0 || b[q] > 0) continue;
        float xp, yp, zp;
        if (select) { xp = (float)sp[0]; yp = (float)sp[1]; zp = (float)sp[2]; }
        else {
          xp = (float)(p->x0 + (i - 1) * p->dx);
          yp = (float)(p->y0 + (j - 1) * p->dx);
          zp = (float)(p->z0 + (k - 1) * p->dx);
        }
        const int32_t *row = &wall_node_tri[13 * (size_t)(windex[q] - 1)];
        const int cnt = row[0] < 12 ? row[0] : 12;
        float mindis = 5000.0f;
        int triwall = 0;
        for (int l = 1; l <= cnt; l++) {
          const int tri = row[l];
          float c[3] = {0, 0, 0};
          for (int s = 0; s < 3; s++) {
            const double *w = &wall_xyz[3 * (size_t)(wall_tri[3 * (size_t)(tri - 1) + s] - 1)];
            const float x = (float)w[0], y = (float)w[1], z = (float)w[2];
            const float d = sqrtf((xp - x) * (xp - x) + (yp - y) * (yp - y) + (zp - z) * (zp - z));
            if (d < mindis) { mindis = d; triwall = tri; }
            c[0] += x; c[1] += y; c[2] += z;
          }
          for (int a = 0; a < 3; a++) c[a] = (float)(c[a] / 3.);
          const float d = sqrtf((xp - c[0]) * (xp - c[0]) + (yp - c[1]) * (yp - c[1]) +
                                (zp - c[2]) * (zp - c[2]));
          if (d < mindis) { mindis = d; triwall = tri; }
        }
        if (!triwall) continue;
        const double wd = wdist[q];
        const double w1 = wd <= d0 ? 0.0 : wd <= d1 ? (wd - d0) / (d1 - d0) : 1.0;
        for (int a = 0; a < 3; a++)
          uvoid[(size_t)a * box + q] = w1 * speed * wall_normal[3 * (size_t)(triwall - 1) + a];
      }
}

void zfo_wall_repulsion(const zfo_ibm_params *p, const zfo_rep_params *r, const int32_t *windex,
                        int nwn, const double *wall_xyz, int nwt, const int32_t *wall_tri,
                        const double *wall_normal, const int32_t *wall_node_tri, int ncell, int nn,
                        const int32_t *status, const double *xyz, double *vrep, int32_t *nearest) {
  (void)nwn;
  (void)nwt;
#pragma omp parallel for schedule(static, 16)
  for (int g = 0; g < ncell * nn; g++) {
    const int c = g / nn;
    nearest[g] = 0;
    if (status[c] <= 0) continue;
    const double *x = &xyz[3 * (size_t)g];
    int lo[3], hi[3];
    window(p, x[0], x[1], x[2], 4, lo, hi);
    double mindis = r->wall_search;
    int triwall = 0;
    for (int i = lo[0]; i <= hi[0]; i++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int k = lo[2]; k <= hi[2]; k++) {
          const int wn = windex[bq(i, j, k, p->jm, p->km)];
          if (wn <= 0) continue;
          const int32_t *row = &wall_node_tri[(size_t)(wn - 1) * 13];
          for (int l = 1; l <= row[0] && l <= 12; l++) {
            const int t = row[l];
            const int32_t *tv = &wall_tri[3 * (size_t)(t - 1)];
            double cen[3] = {0, 0, 0};
            for (int s = 0; s < 3; s++) {
              const double *w = &wall_xyz[3 * (size_t)(tv[s] - 1)];
              const double d = sqrt((x[0] - w[0]) * (x[0] - w[0]) + (x[1] - w[1]) * (x[1] - w[1]) +
                                    (x[2] - w[2]) * (x[2] - w[2]));
              if (d < mindis) { mindis = d; triwall = t; }
              for (int a = 0; a < 3; a++) cen[a] += w[a] / 3.0;
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
      for (int a = 0; a < 3; a++) vrep[3 * (size_t)g + a] += f * r->k_wall * n[a];
    }
  }
}
