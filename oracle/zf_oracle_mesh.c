/*
 * zf_oracle_mesh.c -- RBC membrane of the oracle, FP64 (TEST INFRASTRUCTURE).
 *
 * Restates ZebraflowWT.c:
 *   mesh topology            CGSM_ZFISHRBC_DATA_GEN  :1632-1901
 *   equilibrium parameters   CGSM_FILAMENT_PARAMETERS :1930-2053,
 *                            CGSM_EQUILIBRIUM_METRICS :2056-2223, loop :2338-2346,
 *                            rest areas :2387-2460
 *   triangle geometry        CGSM_AREA :5749-5851, anti-sliver :6042-6086,
 *                            node normals :6089-6093
 *   cell area, centroid      :6717-6796
 *   local forces             :6800-7384 (WLC+POW, bending, local/global area,
 *                            volume gradient, per-term guards)
 *   volume constraint        :7512-7535
 * Units follow the reference: coordinates in a length unit that len_scale
 * converts to the force formulas' unit (1e-6 m per um in the reference).
 * Indices in topology arrays are 1-based like the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "zf_oracle.h"

/* ---- a7: edge list by first appearance over triangles -------------------- */
int zfo_mesh_topology(int nn, int nt, const int32_t *tri, int32_t *edge, int ecap,
                      int32_t *nbr_node, int32_t *nbr_tri, int nlist) {
  int ne = 0;
  /* edges, pairs (1,2),(2,3),(3,1) of each triangle in order (:1655-1718) */
  for (int t = 0; t < nt; t++) {
    const int a[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
    for (int s = 0; s < 3; s++) {
      const int p = a[s], q = a[(s + 1) % 3];
      int dup = 0;
      for (int e = 0; e < ne && !dup; e++)
        dup = (edge[6 * e] == p && edge[6 * e + 1] == q) || (edge[6 * e] == q && edge[6 * e + 1] == p);
      if (dup) continue;
      if (ne >= ecap) return -1;
      memset(&edge[6 * ne], 0, 6 * sizeof(int32_t));
      edge[6 * ne] = p;
      edge[6 * ne + 1] = q;
      ne++;
    }
  }
  /* rhombus apexes n3/n4 and triangles t1/t2: the first and second triangle,
   * in triangle order, holding the edge (:1725-1805) */
  for (int e = 0; e < ne; e++) {
    const int p = edge[6 * e], q = edge[6 * e + 1];
    for (int t = 0; t < nt; t++) {
      const int a[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
      int third = 0;
      for (int s = 0; s < 3; s++) {
        const int u = a[s], v = a[(s + 1) % 3], w = a[(s + 2) % 3];
        if ((u == p && v == q) || (u == q && v == p)) { third = w; break; }
      }
      if (!third) continue;
      if (edge[6 * e + 2] == 0) { edge[6 * e + 2] = third; edge[6 * e + 4] = t + 1; }
      else if (edge[6 * e + 3] == 0) { edge[6 * e + 3] = third; edge[6 * e + 5] = t + 1; }
      if (edge[6 * e + 2] && edge[6 * e + 3]) break;
    }
  }
  /* node -> node (:1836-1856) and node -> triangle (:1873-1891) lists;
   * row = count, then up to nlist-1 entries */
  for (int n = 1; n <= nn; n++) {
    int32_t *r = &nbr_node[(size_t)(n - 1) * nlist];
    memset(r, 0, nlist * sizeof(int32_t));
    int c = 0;
    for (int e = 0; e < ne; e++) {
      int other = 0;
      if (edge[6 * e] == n) other = edge[6 * e + 1];
      else if (edge[6 * e + 1] == n) other = edge[6 * e];
      if (!other) continue;
      c++;
      if (c < nlist) r[c] = other;
    }
    r[0] = c;
    int32_t *rt = &nbr_tri[(size_t)(n - 1) * nlist];
    memset(rt, 0, nlist * sizeof(int32_t));
    c = 0;
    for (int t = 0; t < nt; t++)
      if (tri[3 * t] == n || tri[3 * t + 1] == n || tri[3 * t + 2] == n) {
        c++;
        if (c < nlist) rt[c] = t + 1;
      }
    rt[0] = c;
  }
  return ne;
}

/* ---- a8: equilibrium parameters (51 fixed-point passes) ------------------ */
void zfo_mesh_equilibrium(int nn, int ne, const double *xyz, const int32_t *edge,
                          const zfo_memb_params *mp, double es, int passes, double *lmax,
                          double *plength, double *kp, double *theta0, double *sin0,
                          double *cos0) {
  (void)nn;
  const double L = mp->len_scale, kT = mp->kT, x0 = mp->x0;
  /* beta, ZebraflowWT.c:1942 */
  const double beta = 0.5 * x0 / ((1 - x0) * (1 - x0) * (1 - x0)) - 0.25 / ((1 - x0) * (1 - x0)) + 0.25;
#define X(n, d) xyz[3 * ((n) - 1) + (d)]
  for (int e = 0; e < ne; e++) plength[e] = 1e-9 * (L / 1e-6);
  for (int pass = 1; pass <= passes; pass++) {
    for (int e = 0; e < ne; e++) {
      const int n1 = edge[6 * e], n2 = edge[6 * e + 1];
      const double dx = X(n1, 0) - X(n2, 0), dy = X(n1, 1) - X(n2, 1), dz = X(n1, 2) - X(n2, 2);
      const double l = sqrt(dx * dx + dy * dy + dz * dz);
      /* :1980-1989 (m = 2 in the POW exponent) */
      const double alpha = sqrt(3.0) * 0.25 * kT / (l * L);
      const double eta = sqrt(3.0) * 0.25 * (2.0 + 1.0) / (l * l * l * L * L * L);
      if (pass == 1) plength[e] = 1e-9 * (L / 1e-6);
      kp[e] = (es - alpha * beta / plength[e]) / eta;
      plength[e] = (l * l * L * L) * kT * (x0 - 0.25 + 1.0 / (4.0 * (1 - x0) * (1 - x0))) / kp[e];
    }
  }
  /* LMAX and the rest dihedral angle (:2081-2210) */
  for (int e = 0; e < ne; e++) {
    const int n1 = edge[6 * e], n2 = edge[6 * e + 1], n3 = edge[6 * e + 2], n4 = edge[6 * e + 3];
    double p1[3], p2[3], p3[3], p4[3];
    for (int d = 0; d < 3; d++) { p1[d] = X(n1, d); p2[d] = X(n2, d); p3[d] = X(n3, d); p4[d] = X(n4, d); }
    const double l12 = sqrt((p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]) +
                            (p1[2] - p2[2]) * (p1[2] - p2[2]));
    lmax[e] = L * l12 / x0;
    /* outward area vectors of (n1,n2,n3) and (n2,n1,n4), normalised */
    double a1[3], a2[3], c1[3], c2[3];
    const double u1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double v1[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    const double u2[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
    const double v2[3] = {p4[0] - p2[0], p4[1] - p2[1], p4[2] - p2[2]};
    a1[0] = u1[1] * v1[2] - u1[2] * v1[1]; a1[1] = u1[2] * v1[0] - u1[0] * v1[2]; a1[2] = u1[0] * v1[1] - u1[1] * v1[0];
    a2[0] = u2[1] * v2[2] - u2[2] * v2[1]; a2[1] = u2[2] * v2[0] - u2[0] * v2[2]; a2[2] = u2[0] * v2[1] - u2[1] * v2[0];
    const double m1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    const double m2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    double ct = 0, chk = 0;
    for (int d = 0; d < 3; d++) {
      a1[d] /= m1; a2[d] /= m2;
      c1[d] = (p1[d] + p2[d] + p3[d]) / 3.0;
      c2[d] = (p1[d] + p2[d] + p4[d]) / 3.0;
      ct += a1[d] * a2[d];
      chk += (a1[d] - a2[d]) * (c1[d] - c2[d]);
    }
    if (ct > 1) ct = 1; else if (ct < -1) ct = -1;
    const double st = chk >= 0 ? sqrt(1 - ct * ct) : -sqrt(1 - ct * ct);
    sin0[e] = st;
    cos0[e] = ct;
    theta0[e] = asin(st);
  }
#undef X
}

/* ---- a9-a12: one membrane force evaluation for ncell cells --------------- */
static int bad(double v, double lim) { return isnan(v) || isinf(v) || fabs(v) > lim; }

void zfo_membrane_forces(int ncell, int nn, int nt, int ne, const int32_t *tri,
                         const int32_t *edge, const double *lmax, const double *plength,
                         const double *kp, const double *sin0, const double *cos0,
                         const double *tri_area_eq, const zfo_memb_params *mp, const double *xyz,
                         double *force, double *node_normal, double *tri_area, double *tri_normal,
                         double *tri_centroid, double *cell_area, double *cell_vol,
                         double *cell_centroid, int *guard_fail) {
  const double L = mp->len_scale, lim = mp->force_limit;
  /* cells are independent: one thread per cell, as the reference's
   * per-cell OpenMP loops (ZebraflowWT.c:6800-7535), each with its own
   * scratch, so every cell's sums keep their order */
#pragma omp parallel
  {
  double *fvol = calloc((size_t)3 * nn, sizeof(double));
  double *narea = calloc((size_t)nn, sizeof(double));
  int *owned = calloc((size_t)nt, sizeof(int));
#pragma omp for schedule(static, 1)
  for (int c = 0; c < ncell; c++) {
    const double *x = xyz + (size_t)3 * nn * c;
    double *F = force + (size_t)3 * nn * c;
    double *NN = node_normal + (size_t)3 * nn * c;
    double *TA = tri_area + (size_t)nt * c, *TN = tri_normal + (size_t)3 * nt * c,
           *TC = tri_centroid + (size_t)3 * nt * c;
    int fails = 0;
    memset(F, 0, sizeof(double) * 3 * nn);
    memset(NN, 0, sizeof(double) * 3 * nn);
    memset(fvol, 0, sizeof(double) * 3 * nn);
    memset(narea, 0, sizeof(double) * nn);
    memset(owned, 0, sizeof(int) * nt);
    /* CGSM_AREA (:5757-5851) and anti-sliver (:6042-6086) */
    for (int t = 0; t < nt; t++) {
      const int n1 = tri[3 * t] - 1, n2 = tri[3 * t + 1] - 1, n3 = tri[3 * t + 2] - 1;
      double p1[3], p2[3], p3[3], cen[3];
      for (int d = 0; d < 3; d++) {
        p1[d] = x[3 * n1 + d]; p2[d] = x[3 * n2 + d]; p3[d] = x[3 * n3 + d];
        cen[d] = (p1[d] + p2[d] + p3[d]) / 3.0;
        TC[3 * t + d] = cen[d];
      }
      const double e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
      const double e3[3] = {p1[0] - p3[0], p1[1] - p3[1], p1[2] - p3[2]};
      const double Ax = 0.5 * (e1[1] * e3[2] - e1[2] * e3[1]) * L * L;
      const double Ay = 0.5 * (e1[2] * e3[0] - e1[0] * e3[2]) * L * L;
      const double Az = 0.5 * (e1[0] * e3[1] - e1[1] * e3[0]) * L * L;
      const double A = sqrt(Ax * Ax + Ay * Ay + Az * Az);
      TA[t] = A;
      const double nrm[3] = {-Ax / A, -Ay / A, -Az / A};
      for (int d = 0; d < 3; d++) TN[3 * t + d] = nrm[d];
      const int ns[3] = {n1, n2, n3};
      for (int s = 0; s < 3; s++) {
        narea[ns[s]] += A / 3.0;
        for (int d = 0; d < 3; d++) NN[3 * ns[s] + d] += nrm[d] * A / 3.0;
      }
      const double *ps[3] = {p1, p2, p3};
      for (int s = 0; s < 3; s++) {
        double ev[3], l = 0;
        for (int d = 0; d < 3; d++) { ev[d] = ps[s][d] - cen[d]; l += ev[d] * ev[d]; }
        l = sqrt(l);
        if (l < mp->sliver_len) {
          const double mag = l > mp->sliver_len_min ? mp->relax2 * mp->sliver_f1 / l
                                                    : mp->relax2 * mp->sliver_f2;
          for (int d = 0; d < 3; d++) F[3 * ns[s] + d] += mag * ev[d] / l;
        }
      }
    }
    for (int n = 0; n < nn; n++)
      for (int d = 0; d < 3; d++) NN[3 * n + d] /= narea[n];
    /* cell area and centroid (:6729-6791) in FP64 */
    double area = 0, cc[3] = {0, 0, 0};
    for (int n = 0; n < nn; n++)
      for (int d = 0; d < 3; d++) cc[d] += x[3 * n + d] / nn;
    for (int t = 0; t < nt; t++) area += TA[t];
    cell_area[c] = area;
    for (int d = 0; d < 3; d++) cell_centroid[3 * c + d] = cc[d];
    const double f_area_global = -mp->ka * mp->relax * (area - mp->area0) / mp->area0;
    /* local forces, edge by edge (:6810-7384) */
    for (int e = 0; e < ne; e++) {
      const int n1 = edge[6 * e] - 1, n2 = edge[6 * e + 1] - 1, n3 = edge[6 * e + 2] - 1,
                n4 = edge[6 * e + 3] - 1, t1 = edge[6 * e + 4] - 1, t2 = edge[6 * e + 5] - 1;
      double p1[3], p2[3], p3[3], p4[3];
      for (int d = 0; d < 3; d++) { p1[d] = x[3 * n1 + d]; p2[d] = x[3 * n2 + d]; p3[d] = x[3 * n3 + d]; p4[d] = x[3 * n4 + d]; }
      double fn1[3] = {0}, fn2[3] = {0}, fn3[3] = {0}, fn4[3] = {0};
      /* WLC + POW (:6888-6924) */
      const double d12[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
      double len = sqrt(d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2]);
      const double e1v[3] = {d12[0] / len, d12[1] / len, d12[2] / len};
      double ratio = len * L / lmax[e];
      if (ratio > mp->x_max) ratio = mp->x_max;
      if (ratio < mp->x_floor) len = mp->x_floor * lmax[e] / L;
      const double wlc = -mp->kT / plength[e] * (ratio - 0.25 + 1.0 / (4.0 * (1 - ratio) * (1 - ratio)));
      const double comp = kp[e] / (len * len * L * L);
      for (int d = 0; d < 3; d++) {
        fn1[d] = mp->relax2 * e1v[d] * wlc + mp->relax2 * e1v[d] * comp;
        fn2[d] = -mp->relax2 * e1v[d] * wlc - mp->relax2 * e1v[d] * comp;
      }
      /* geometry of the two triangles, clamped areas (:6929-6951) */
      double a1 = TA[t1], a2 = TA[t2];
      if (a1 < mp->area_lo * tri_area_eq[t1]) a1 = mp->area_lo * tri_area_eq[t1];
      else if (a1 > mp->area_hi * tri_area_eq[t1]) a1 = mp->area_hi * tri_area_eq[t1];
      if (a2 < mp->area_lo * tri_area_eq[t2]) a2 = mp->area_lo * tri_area_eq[t2];
      else if (a2 > mp->area_hi * tri_area_eq[t2]) a2 = mp->area_hi * tri_area_eq[t2];
      const double *m1 = &TN[3 * t1], *m2 = &TN[3 * t2];
      double ct = m1[0] * m2[0] + m1[1] * m2[1] + m1[2] * m2[2];
      if (ct > 1) ct = 1; else if (ct < -1) ct = -1;
      double chk = 0;
      for (int d = 0; d < 3; d++) chk += (m1[d] - m2[d]) * (TC[3 * t1 + d] - TC[3 * t2 + d]) * L;
      const double st = chk >= 0 ? sqrt(1 - ct * ct) : -sqrt(1 - ct * ct);
      /* bending (:6984-7081) */
      double bb = fabs(ct) > mp->cos_cut ? 0.0 : mp->kbend * (st * cos0[e] - ct * sin0[e]) / st;
      if (mp->cos_blend > 0 && fabs(ct) <= mp->cos_cut && fabs(ct) > mp->cos_cut - mp->cos_blend) {
        /* smooth variant, not in the reference: smoothstep fade to zero at cos_cut */
        const double t = (mp->cos_cut - fabs(ct)) * (1.0 / mp->cos_blend);
        bb = bb * (t * t * (3.0 - 2.0 * t));
      }
      double dp21[3], dp24[3], dp32[3], dp41[3], dp13[3];
      for (int d = 0; d < 3; d++) {
        dp21[d] = L * (p2[d] - p1[d]); dp24[d] = L * (p2[d] - p4[d]); dp32[d] = L * (p3[d] - p2[d]);
        dp41[d] = L * (p4[d] - p1[d]); dp13[d] = L * (p1[d] - p3[d]);
      }
#define CROSS(o, s, n, v) do { o[0] = s * (n[1] * v[2] - n[2] * v[1]); o[1] = s * (n[2] * v[0] - n[0] * v[2]); o[2] = s * (n[0] * v[1] - n[1] * v[0]); } while (0)
      double c1[3], c2[3], c3[3], c4[3], c5[3], c6[3], c7[3], c8[3], c9[3], c10[3];
      CROSS(c1, a1, m1, dp21); CROSS(c2, a2, m2, dp21); CROSS(c3, a1, m1, dp24); CROSS(c4, a2, m2, dp24);
      CROSS(c5, a1, m1, dp32); CROSS(c6, a2, m2, dp32); CROSS(c7, a1, m1, dp41); CROSS(c8, a2, m2, dp41);
      CROSS(c9, a1, m1, dp13); CROSS(c10, a2, m2, dp13);
      double lc1[3], lc2[3];
      for (int d = 0; d < 3; d++) {
        lc1[d] = (TC[3 * t1 + d] - cc[d]) * L;
        lc2[d] = (TC[3 * t2 + d] - cc[d]) * L;
      }
      double c1v[3], c5v[3], c9v[3], c4v[3], c8v[3], c12v[3];
      CROSS(c1v, 1.0, lc1, dp21); CROSS(c5v, 1.0, lc1, dp32); CROSS(c9v, 1.0, lc1, dp13);
      CROSS(c4v, 1.0, lc2, dp24); CROSS(c8v, 1.0, lc2, dp41); CROSS(c12v, -1.0, lc2, dp21);
#undef CROSS
      const double b11 = -bb * ct / (a1 * a1), b12 = bb / (a1 * a2), b22 = -bb * ct / (a2 * a2);
      double fb[4][3];
      for (int d = 0; d < 3; d++) {
        fb[0][d] = b11 * c5[d] + b12 * (c3[d] + c6[d]) + b22 * c4[d];
        fb[1][d] = b11 * c9[d] + b12 * (c7[d] + c10[d]) + b22 * c8[d];
        fb[2][d] = b11 * c1[d] + b12 * c2[d];
        fb[3][d] = b12 * -c1[d] + b22 * -c2[d];
      }
      double *fn[4] = {fn1, fn2, fn3, fn4};
      for (int r = 0; r < 4; r++) {
        int k = 0;
        for (int d = 0; d < 3; d++) k += bad(fb[r][d], lim);
        fails += k;   /* counted per bad component, as CAPS_REMOVE (:7041-7070) */
        if (!k) for (int d = 0; d < 3; d++) fn[r][d] += fb[r][d];
      }
      /* area and volume terms, once per triangle: its first edge (:7126-7359) */
      for (int side = 0; side < 2; side++) {
        const int t = side ? t2 : t1;
        if (owned[t]) continue;
        owned[t] = 1;
        const double a = side ? a2 : a1;
        const double *m = side ? m2 : m1;
        const double alpha = -mp->kd * mp->relax * (a - tri_area_eq[t]) / (4.0 * tri_area_eq[t] * a);
        const double *g[3], *gv[3];
        double neg2[3];
        if (!side) { g[0] = c5; g[1] = c9; g[2] = c1; gv[0] = c5v; gv[1] = c9v; gv[2] = c1v; }
        else {
          for (int d = 0; d < 3; d++) neg2[d] = -c2[d];
          g[0] = c4; g[1] = c8; g[2] = neg2; gv[0] = c4v; gv[1] = c8v; gv[2] = c12v;
        }
        double *tgt[3] = {fn1, fn2, side ? fn4 : fn3};
        const int vn[3] = {n1, n2, side ? n4 : n3};
        double fa[3][3], faa[3][3], fv[3][3];
        for (int r = 0; r < 3; r++)
          for (int d = 0; d < 3; d++) {
            fa[r][d] = alpha * g[r][d];
            faa[r][d] = 0.25 * f_area_global / a * g[r][d];
            fv[r][d] = 1.0 / 6.0 * (a * m[d] / 3.0 + gv[r][d]);
          }
        for (int r = 0; r < 3; r++) {
          int k = 0;
          for (int d = 0; d < 3; d++) k += bad(fa[r][d], lim);
          fails += k;
          if (!k) for (int d = 0; d < 3; d++) tgt[r][d] += fa[r][d];
        }
        for (int r = 0; r < 3; r++) {
          int k = 0;
          for (int d = 0; d < 3; d++) k += bad(faa[r][d], lim);
          fails += k;
          if (!k) for (int d = 0; d < 3; d++) tgt[r][d] += faa[r][d];
        }
        for (int r = 0; r < 3; r++) {
          int k = 0;
          for (int d = 0; d < 3; d++) k |= bad(fv[r][d], mp->vol_limit);
          if (!k) for (int d = 0; d < 3; d++) fvol[3 * vn[r] + d] += fv[r][d];
        }
      }
      const int ns[4] = {n1, n2, n3, n4};
      for (int r = 0; r < 4; r++)
        for (int d = 0; d < 3; d++) F[3 * ns[r] + d] += fn[r][d];
    }
    /* volume constraint (:7517-7531) in FP64 */
    double vol = 0;
    for (int t = 0; t < nt; t++)
      vol += TA[t] * (TC[3 * t] * TN[3 * t] + TC[3 * t + 1] * TN[3 * t + 1] + TC[3 * t + 2] * TN[3 * t + 2]) * L;
    vol /= 3.0;
    cell_vol[c] = vol;
    const double fvs = -mp->kv * (vol - mp->vol0) / mp->vol0;
    for (int n = 0; n < nn; n++)
      for (int d = 0; d < 3; d++) F[3 * n + d] += mp->relax * fvs * fvol[3 * n + d];
    guard_fail[c] = fails;
  }
  free(fvol);
  free(narea);
  free(owned);
  }
}

/* ---- nucleus network (ZebraflowWT.c:1903-1922, :2003-2050, :2203-2221,
 * :7388-7466, :9519-9730) ------------------------------------------------- */

/* Filament parameters of the rest network: the 51-pass fixed point of
 * :2003-2050 with es (the caller passes 0.1 Es), LMAX_N of :2203-2221. */
void zfo_nucleus_equilibrium(int nnuc, int ne, const double *xyz, const int32_t *edge,
                             const zfo_memb_params *mp, double es, int passes, double *lmax,
                             double *plength, double *kp) {
  (void)nnuc;
  const double L = mp->len_scale, x0 = mp->x0;
  const double beta = 0.5 * x0 / ((1. - x0) * (1. - x0) * (1. - x0)) - 0.25 / ((1. - x0) * (1. - x0)) + 0.25;
  for (int e = 0; e < ne; e++) {
    const double *p1 = &xyz[3 * (size_t)(edge[2 * e] - 1)], *p2 = &xyz[3 * (size_t)(edge[2 * e + 1] - 1)];
    const double l = sqrt((p1[0] - p2[0]) * (p1[0] - p2[0]) + (p1[1] - p2[1]) * (p1[1] - p2[1]) +
                          (p1[2] - p2[2]) * (p1[2] - p2[2]));
    const double alpha = sqrt(3.) * 0.25 * mp->kT / (l * L);
    const double eta = sqrt(3.) * 0.25 * (2. + 1.) / (l * l * l * L * L * L);
    double pl = 1e-9 * (L / 1e-6), k = 0;
    for (int it = 0; it < passes; it++) {
      k = (es - alpha * beta / pl) / eta;
      pl = (l * l * L * L) * mp->kT * (x0 - 0.25 + 1. / (4. * (1. - x0) * (1. - x0))) / k;
    }
    lmax[e] = L * l / x0;
    plength[e] = pl;
    kp[e] = k;
  }
}

/* F_Total_N (:7388-7466): per cell, edges in order, the WLC + POW spring
 * added to both end nodes. */
void zfo_nucleus_forces(int ncell, int nnuc, int ne, const int32_t *edge, const double *lmax,
                        const double *plength, const double *kp, const zfo_memb_params *mp,
                        const double *xyzn, double *force) {
  const double L = mp->len_scale;
  for (int c = 0; c < ncell; c++) {
    const double *x = xyzn + (size_t)3 * nnuc * c;
    double *F = force + (size_t)3 * nnuc * c;
    memset(F, 0, sizeof(double) * 3 * nnuc);
    for (int e = 0; e < ne; e++) {
      const int n1 = edge[2 * e] - 1, n2 = edge[2 * e + 1] - 1;
      double d[3], len = 0;
      for (int a = 0; a < 3; a++) { d[a] = x[3 * n1 + a] - x[3 * n2 + a]; len += d[a] * d[a]; }
      len = sqrt(len);
      double e1v[3], e2v[3];
      for (int a = 0; a < 3; a++) { e1v[a] = d[a] / len; e2v[a] = -d[a] / len; }
      double ratio = len * L / lmax[e];
      if (ratio > mp->x_max) ratio = mp->x_max;
      if (ratio < mp->x_floor) len = mp->x_floor * lmax[e] / L;
      const double wlc = -mp->kT / plength[e] * (ratio - 0.25 + 1. / (4. * (1. - ratio) * (1. - ratio)));
      const double comp = kp[e] * (1. / (len * len * L * L));
      for (int a = 0; a < 3; a++) {
        double f1 = mp->relax2 * e1v[a] * wlc;
        f1 = f1 + mp->relax2 * e1v[a] * comp;
        double f2 = mp->relax2 * e2v[a] * wlc;
        f2 = f2 + mp->relax2 * e2v[a] * comp;
        F[3 * n1 + a] += f1;
        F[3 * n2 + a] += f2;
      }
    }
  }
}

/* Nucleus-membrane contact of the cells with status >= 0 (:9524-9730):
 * each nucleus node lists its cell's triangles with centroid closer than
 * list_r (in order); per nucleus node in order, per listed triangle whose
 * centroid is closer than near_r, the 19 probe points of :9662-9699, and
 * every point within `contact` adds k (contact - d)/contact n/3 to the
 * repulsion velocity of the triangle's three nodes.  List entries beyond
 * list_cap are dropped (counted in the return value). */
int zfo_nucleus_contact(int ncell, int nnuc, int nn, int nt, const int32_t *tri,
                        const int32_t *status, const double *xyzn, const double *xyz,
                        const double *tri_centroid, const double *tri_normal, double list_r,
                        double near_r, double contact, double k, int list_cap, double *vrep) {
  int lost = 0;
  int *lst = malloc(sizeof(int) * (size_t)(list_cap > 0 ? list_cap : 1));
  for (int c = 0; c < ncell; c++) {
    if (status[c] < 0) continue;
    const double *xn = xyzn + (size_t)3 * nnuc * c, *x = xyz + (size_t)3 * nn * c;
    const double *TC = tri_centroid + (size_t)3 * nt * c, *TN = tri_normal + (size_t)3 * nt * c;
    double *V = vrep + (size_t)3 * nn * c;
    for (int node = 0; node < nnuc; node++) {
      const double xc = xn[3 * node], yc = xn[3 * node + 1], zc = xn[3 * node + 2];
      int cnt = 0;
      for (int m = 0; m < nt; m++) {
        const double dx = xc - TC[3 * m], dy = yc - TC[3 * m + 1], dz = zc - TC[3 * m + 2];
        if (sqrt(dx * dx + dy * dy + dz * dz) < list_r) {
          if (cnt < list_cap) lst[cnt++] = m;
          else lost++;
        }
      }
      for (int l = 0; l < cnt; l++) {
        const int t = lst[l];
        const double dx = xc - TC[3 * t], dy = yc - TC[3 * t + 1], dz = zc - TC[3 * t + 2];
        if (!(sqrt(dx * dx + dy * dy + dz * dz) < near_r)) continue;
        const int n1 = tri[3 * t] - 1, n2 = tri[3 * t + 1] - 1, n3 = tri[3 * t + 2] - 1;
        double ep[10][3], tp[20][3];
        for (int v = 0; v < 3; v++) {
          const double a = x[3 * n1 + v], b = x[3 * n2 + v], cc = x[3 * n3 + v];
          ep[1][v] = (a + b) / 2.;
          ep[2][v] = (b + cc) / 2.;
          ep[3][v] = (cc + a) / 2.;
          tp[1][v] = (a + b + cc) / 3.;
          tp[2][v] = (ep[3][v] + a + ep[1][v]) / 3.;
          tp[3][v] = (ep[1][v] + b + ep[2][v]) / 3.;
          tp[4][v] = (ep[2][v] + cc + ep[3][v]) / 3.;
          tp[5][v] = (ep[1][v] + ep[3][v]) / 2.;
          tp[6][v] = (ep[1][v] + ep[2][v]) / 2.;
          tp[7][v] = (ep[2][v] + ep[3][v]) / 2.;
          tp[8][v] = (a + tp[2][v]) / 2.;
          tp[9][v] = (ep[1][v] + tp[2][v]) / 2.;
          tp[10][v] = (ep[3][v] + tp[2][v]) / 2.;
          tp[11][v] = (ep[1][v] + tp[3][v]) / 2.;
          tp[12][v] = (b + tp[3][v]) / 2.;
          tp[13][v] = (ep[2][v] + tp[3][v]) / 2.;
          tp[14][v] = (ep[3][v] + tp[4][v]) / 2.;
          tp[15][v] = (ep[2][v] + tp[4][v]) / 2.;
          tp[16][v] = (cc + tp[4][v]) / 2.;
          tp[17][v] = (ep[1][v] + tp[1][v]) / 2.;
          tp[18][v] = (ep[2][v] + tp[1][v]) / 2.;
          tp[19][v] = (ep[3][v] + tp[1][v]) / 2.;
        }
        for (int v = 1; v <= 19; v++) {
          const double d = sqrt((tp[v][0] - xc) * (tp[v][0] - xc) + (tp[v][1] - yc) * (tp[v][1] - yc) +
                                (tp[v][2] - zc) * (tp[v][2] - zc));
          if (d <= contact) {
            const double rf = (contact - d) / contact * k;
            for (int i = 0; i < 3; i++) {
              V[3 * n1 + i] += rf * TN[3 * t + i] / 3.;
              V[3 * n2 + i] += rf * TN[3 * t + i] / 3.;
              V[3 * n3 + i] += rf * TN[3 * t + i] / 3.;
            }
          }
        }
      }
    }
  }
  free(lst);
  return lost;
}
