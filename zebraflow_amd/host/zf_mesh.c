/*
 * zf_mesh.c -- RBC membrane meshes and cell placement.
 *
 * The reference reads one fixed membrane mesh (ZFISHRBCNODES/TRIANGLES.dat,
 * ZebraflowWT.c:1601-1629).  The benchmark cells are Evans-Fung biconcave
 * discs meshed from a subdivided icosahedron: level 3 gives 642 nodes, 1280
 * triangles and 1920 edges.  Triangles are wound counter-clockwise seen
 * from outside, so (n2-n1) x (n3-n1) is the outward normal, the orientation
 * CGSM_AREA relies on (ZebraflowWT.c:5808-5831).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "zf_host.h"

typedef struct { int a, b, mid; } emid_t;

static int find_or_add_mid(emid_t *tab, int cap, int a, int b, double *xyz, int *nv) {
  if (a > b) { int t = a; a = b; b = t; }
  unsigned h = ((unsigned)a * 73856093u ^ (unsigned)b * 19349663u) % (unsigned)cap;
  while (tab[h].mid >= 0) {
    if (tab[h].a == a && tab[h].b == b) return tab[h].mid;
    h = (h + 1) % (unsigned)cap;
  }
  const int m = (*nv)++;
  double p[3], n = 0;
  for (int d = 0; d < 3; d++) { p[d] = 0.5 * (xyz[3 * a + d] + xyz[3 * b + d]); n += p[d] * p[d]; }
  n = sqrt(n);
  for (int d = 0; d < 3; d++) xyz[3 * m + d] = p[d] / n;
  tab[h].a = a; tab[h].b = b; tab[h].mid = m;
  return m;
}

int zfh_mesh_icosphere(int level, double radius, int biconcave, double *xyz, int32_t *tri,
                       int node_cap, int tri_cap, int32_t *n_tri) {
  if (level < 0 || level > 6 || !xyz || !tri || !n_tri) return -1;
  int nv = 12, nt = 20;
  int need_v = 10 * (1 << (2 * level)) + 2, need_t = 20 * (1 << (2 * level));
  if (node_cap < need_v || tri_cap < need_t) return -1;
  const double t = (1.0 + sqrt(5.0)) / 2.0;
  const double v0[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                            {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                            {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const int f0[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                         {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                         {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                         {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  for (int i = 0; i < 12; i++) {
    const double n = sqrt(v0[i][0] * v0[i][0] + v0[i][1] * v0[i][1] + v0[i][2] * v0[i][2]);
    for (int d = 0; d < 3; d++) xyz[3 * i + d] = v0[i][d] / n;
  }
  int32_t *cur = malloc(sizeof(int32_t) * 3 * need_t), *nxt = malloc(sizeof(int32_t) * 3 * need_t);
  for (int f = 0; f < 20; f++)
    for (int s = 0; s < 3; s++) cur[3 * f + s] = f0[f][s];
  for (int l = 0; l < level; l++) {
    const int cap = 8 * nt + 1;
    emid_t *tab = malloc(sizeof(emid_t) * cap);
    for (int i = 0; i < cap; i++) tab[i].mid = -1;
    for (int f = 0; f < nt; f++) {
      const int a = cur[3 * f], b = cur[3 * f + 1], c = cur[3 * f + 2];
      const int ab = find_or_add_mid(tab, cap, a, b, xyz, &nv);
      const int bc = find_or_add_mid(tab, cap, b, c, xyz, &nv);
      const int ca = find_or_add_mid(tab, cap, c, a, xyz, &nv);
      const int q[4][3] = {{a, ab, ca}, {b, bc, ab}, {c, ca, bc}, {ab, bc, ca}};
      for (int k = 0; k < 4; k++)
        for (int s = 0; s < 3; s++) nxt[3 * (4 * f + k) + s] = q[k][s];
    }
    free(tab);
    nt *= 4;
    int32_t *sw = cur; cur = nxt; nxt = sw;
  }
  /* Evans-Fung biconcave profile on the unit sphere (z = +-sqrt(1-r^2)):
   * z' = 0.5 z (c0 + c1 r^2 + c2 r^4), then scale by the radius. */
  const double c0 = 0.2072, c1 = 2.0026, c2 = -1.1228;
  for (int i = 0; i < nv; i++) {
    double *p = &xyz[3 * i];
    if (biconcave) {
      const double r2 = p[0] * p[0] + p[1] * p[1];
      p[2] = 0.5 * p[2] * (c0 + c1 * r2 + c2 * r2 * r2);
    }
    for (int d = 0; d < 3; d++) p[d] *= radius;
  }
  for (int f = 0; f < nt; f++)
    for (int s = 0; s < 3; s++) tri[3 * f + s] = cur[3 * f + s] + 1; /* 1-based */
  free(cur);
  free(nxt);
  *n_tri = nt;
  return nv;
}

/* Rotation from three seeded angles, then translate. */
static void place(const double *ref, int nn, const double c[3], double a, double b, double g,
                  double *out) {
  const double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
  const double R[3][3] = {{ca * cb, ca * sb * sg - sa * cg, ca * sb * cg + sa * sg},
                          {sa * cb, sa * sb * sg + ca * cg, sa * sb * cg - ca * sg},
                          {-sb, cb * sg, cb * cg}};
  for (int n = 0; n < nn; n++)
    for (int d = 0; d < 3; d++)
      out[3 * n + d] = c[d] + R[d][0] * ref[3 * n] + R[d][1] * ref[3 * n + 1] + R[d][2] * ref[3 * n + 2];
}

static uint64_t sm(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double ur(uint64_t *s) { return (sm(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Random non-overlapping placement: each cell's bounding sphere (radius
 * r_cell + gap) lies inside interior fluid (every lattice node within it has
 * 0<B<10) and does not meet another cell's.  Returns cells placed. */
int zfh_place_cells(const int32_t *b, int im, int jm, int km, const double *ref, int nn,
                    double r_cell, double gap, int ncell, uint64_t seed, int max_tries,
                    double *xyz_out) {
  uint64_t s = seed;
  double *cen = malloc(sizeof(double) * 3 * (ncell > 0 ? ncell : 1));
  int placed = 0;
  const double rr = r_cell + gap;
  const int ri = (int)ceil(rr);
  for (int tries = 0; tries < max_tries && placed < ncell; tries++) {
    const double c[3] = {rr + 1 + ur(&s) * (im - 3 - 2 * rr), rr + 1 + ur(&s) * (jm - 3 - 2 * rr),
                         rr + 1 + ur(&s) * (km - 3 - 2 * rr)};
    int ok = 1;
    for (int q = 0; q < placed && ok; q++) {
      const double dx = c[0] - cen[3 * q], dy = c[1] - cen[3 * q + 1], dz = c[2] - cen[3 * q + 2];
      if (dx * dx + dy * dy + dz * dz < 4 * rr * rr) ok = 0;
    }
    for (int i = (int)floor(c[0]) - ri; i <= (int)ceil(c[0]) + ri && ok; i++)
      for (int j = (int)floor(c[1]) - ri; j <= (int)ceil(c[1]) + ri && ok; j++)
        for (int k = (int)floor(c[2]) - ri; k <= (int)ceil(c[2]) + ri && ok; k++) {
          const double dx = i - c[0], dy = j - c[1], dz = k - c[2];
          if (dx * dx + dy * dy + dz * dz > (rr + 1) * (rr + 1)) continue;
          if (i < 0 || i >= im || j < 0 || j >= jm || k < 0 || k >= km) { ok = 0; break; }
          const int v = b[((size_t)i * jm + j) * km + k];
          if (!(v > 0 && v < 10)) ok = 0;
        }
    if (!ok) continue;
    for (int d = 0; d < 3; d++) cen[3 * placed + d] = c[d];
    place(ref, nn, c, 2 * M_PI * ur(&s), acos(2 * ur(&s) - 1), 2 * M_PI * ur(&s),
          xyz_out + (size_t)3 * nn * placed);
    placed++;
  }
  free(cen);
  return placed;
}

/* Dense random placement for high haematocrit: centres drawn from the
 * interior fluid nodes, random orientations, and the test is on the mesh
 * itself rather than a bounding sphere: every node must sit on a lattice
 * point neighbourhood (its nearest node and that node's 6 neighbours) of
 * interior fluid, and no node may come within `gap` of a node of a cell
 * already placed (uniform-grid hash of placed nodes).  Returns cells placed. */
int zfh_place_cells_dense(const int32_t *b, int im, int jm, int km, const double *ref, int nn,
                          double gap, int ncell, uint64_t seed, int max_tries, double *xyz_out) {
  if (!b || !ref || nn <= 0 || ncell < 0 || gap <= 0 || !xyz_out) return -1;
  const size_t box = (size_t)im * jm * km;
  int64_t nf = 0;
  for (size_t q = 0; q < box; q++) nf += b[q] > 0 && b[q] < 10;
  if (nf == 0 || ncell == 0) return 0;
  int64_t *fl = malloc(sizeof(int64_t) * nf);
  nf = 0;
  for (size_t q = 0; q < box; q++)
    if (b[q] > 0 && b[q] < 10) fl[nf++] = (int64_t)q;
  /* hash grid of placed nodes, cell size h >= gap */
  const double h = gap > 1.0 ? gap : 1.0;
  const int gi = (int)(im / h) + 2, gj = (int)(jm / h) + 2, gk = (int)(km / h) + 2;
  const size_t ng = (size_t)gi * gj * gk;
  int32_t *head = malloc(sizeof(int32_t) * ng);
  int32_t *next = malloc(sizeof(int32_t) * (size_t)ncell * nn);
  for (size_t q = 0; q < ng; q++) head[q] = -1;
  double *cand = malloc(sizeof(double) * 3 * nn);
  uint64_t s = seed;
  int placed = 0;
  const int off[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int tries = 0; tries < max_tries && placed < ncell; tries++) {
    const int64_t q = fl[(int64_t)(ur(&s) * nf) % nf];
    const double c[3] = {(double)(q / ((int64_t)jm * km)) + ur(&s) - 0.5,
                         (double)((q / km) % jm) + ur(&s) - 0.5, (double)(q % km) + ur(&s) - 0.5};
    place(ref, nn, c, 2 * M_PI * ur(&s), acos(2 * ur(&s) - 1), 2 * M_PI * ur(&s), cand);
    int ok = 1;
    for (int n = 0; n < nn && ok; n++) {
      const double *x = &cand[3 * n];
      const int i = (int)floor(x[0] + 0.5), j = (int)floor(x[1] + 0.5), k = (int)floor(x[2] + 0.5);
      for (int o = 0; o < 7 && ok; o++) {
        const int a = i + off[o][0], bb = j + off[o][1], cc = k + off[o][2];
        if (a < 0 || a >= im || bb < 0 || bb >= jm || cc < 0 || cc >= km) { ok = 0; break; }
        const int v = b[((size_t)a * jm + bb) * km + cc];
        if (!(v > 0 && v < 10)) ok = 0;
      }
      if (!ok) break;
      const int hi = (int)(x[0] / h), hj = (int)(x[1] / h), hk = (int)(x[2] / h);
      for (int a = hi - 1; a <= hi + 1 && ok; a++)
        for (int bb = hj - 1; bb <= hj + 1 && ok; bb++)
          for (int cc = hk - 1; cc <= hk + 1 && ok; cc++) {
            if (a < 0 || a >= gi || bb < 0 || bb >= gj || cc < 0 || cc >= gk) continue;
            for (int e = head[((size_t)a * gj + bb) * gk + cc]; e >= 0; e = next[e]) {
              const double *y = &xyz_out[3 * (size_t)e];
              const double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
              if (dx * dx + dy * dy + dz * dz < gap * gap) { ok = 0; break; }
            }
          }
    }
    if (!ok) continue;
    memcpy(&xyz_out[3 * (size_t)placed * nn], cand, sizeof(double) * 3 * nn);
    for (int n = 0; n < nn; n++) {
      const int e = placed * nn + n;
      const double *x = &cand[3 * n];
      const size_t g = ((size_t)(int)(x[0] / h) * gj + (size_t)(int)(x[1] / h)) * gk + (size_t)(int)(x[2] / h);
      next[e] = head[g];
      head[g] = e;
    }
    placed++;
  }
  free(fl);
  free(head);
  free(next);
  free(cand);
  return placed;
}

void zfh_membrane_config_lattice(const double *xyz, int nn, const int32_t *tri, int nt, double es,
                                 double kb, int max_cells, zf_membrane_config *mc) {
  zf_membrane_config_reference(mc);
  /* mean rest edge length over the unique edges (each interior edge of a
   * closed mesh is shared by two triangles, counted once by a < b) */
  double lsum = 0;
  int64_t ne = 0;
  double area = 0, vol = 0;
  for (int t = 0; t < nt; t++) {
    const int v[3] = {tri[3 * t] - 1, tri[3 * t + 1] - 1, tri[3 * t + 2] - 1};
    for (int s = 0; s < 3; s++) {
      const int a = v[s], b = v[(s + 1) % 3];
      if (a < b) {
        const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
                     dz = xyz[3 * a + 2] - xyz[3 * b + 2];
        lsum += sqrt(dx * dx + dy * dy + dz * dz);
        ne++;
      }
    }
    const double *p0 = &xyz[3 * v[0]], *p1 = &xyz[3 * v[1]], *p2 = &xyz[3 * v[2]];
    const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double av[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                          0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    area += sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
    vol += (av[0] * (p0[0] + p1[0] + p2[0]) + av[1] * (p0[1] + p1[1] + p2[1]) +
            av[2] * (p0[2] + p1[2] + p2[2])) / 9.0;
  }
  (void)nn;
  const double l0 = ne > 0 ? lsum / (double)ne : 1.0;
  mc->len_scale = 1.0;
  mc->es = es;
  mc->kT = 0.0036 * es * l0 * l0;
  mc->kbend = 2.0 / sqrt(3.0) * kb;
  mc->kd = 12.0 * es;
  mc->ka = 25.67 * es;
  mc->kv = 9.17 * es;
  mc->area0 = area;
  mc->vol0 = vol;
  mc->force_limit = 60.0 * es * l0;
  mc->vol_limit = 1e30;
  mc->sliver_len = 0.1;
  mc->sliver_len_min = 0.002;
  mc->sliver_f1 = 0.01 * es * l0;
  mc->sliver_f2 = 6.0 * es * l0;
  mc->max_cells = max_cells;
}
