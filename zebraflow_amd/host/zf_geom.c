/*
 * zf_geom.c -- synthetic vessel geometries written as a B_index box.
 *
 * The reference reads its lumen as B_index values from
 * FLUIDdomainBindex.csv (ZebraflowWT.c:2574-2619); these generators produce
 * the same kind of field for the benchmark configurations: B=1 fluid, B=0 a
 * one-node wall layer (non-fluid nodes with a D3Q19 fluid neighbour, the
 * layer LBM_BC_SETUP treats as walls, ZebraflowWT.c:1527), B=-1 outside, and
 * inlet/outlet codes on the i faces.  Node (i,j,k) sits at lattice position
 * (i-1, j-1, k-1).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "zf_host.h"

static const int E18[18][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 0, 1},  {-1, 0, -1},
    {-1, 0, 1}, {1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {0, -1, 1}, {0, 1, -1}};

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

static uint64_t splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double urand(uint64_t *s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Mark every node within distance r of segment a-b with `tag` (if untagged). */
static void raster_segment(uint8_t *tag, int im, int jm, int km, const double a[3],
                           const double b[3], double r, uint8_t t) {
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = fmin(a[d], b[d]) - r - 1;
    hi[d] = fmax(a[d], b[d]) + r + 1;
  }
  const int i0 = (int)fmax(1, floor(lo[0]) + 1), i1 = (int)fmin(im, ceil(hi[0]) + 1);
  const int j0 = (int)fmax(1, floor(lo[1]) + 1), j1 = (int)fmin(jm, ceil(hi[1]) + 1);
  const int k0 = (int)fmax(1, floor(lo[2]) + 1), k1 = (int)fmin(km, ceil(hi[2]) + 1);
  const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double L2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
  for (int i = i0; i <= i1; i++)
    for (int j = j0; j <= j1; j++)
      for (int k = k0; k <= k1; k++) {
        const double p[3] = {i - 1.0, j - 1.0, k - 1.0};
        double t_ = 0;
        if (L2 > 0)
          t_ = ((p[0] - a[0]) * ab[0] + (p[1] - a[1]) * ab[1] + (p[2] - a[2]) * ab[2]) / L2;
        t_ = t_ < 0 ? 0 : t_ > 1 ? 1 : t_;
        const double dx = p[0] - a[0] - t_ * ab[0], dy = p[1] - a[1] - t_ * ab[1],
                     dz = p[2] - a[2] - t_ * ab[2];
        if (dx * dx + dy * dy + dz * dz <= r * r) {
          uint8_t *q = &tag[bq(i, j, k, jm, km)];
          if (*q == 0) *q = t;
        }
      }
}

/* Fluid from tags, then the wall layer (wrapping along periodic axes). */
static int64_t finish(int32_t *b, const uint8_t *tag, int im, int jm, int km, const int per[3]) {
  const size_t n = (size_t)im * jm * km;
  int64_t nf = 0;
  for (size_t q = 0; q < n; q++) {
    b[q] = tag[q] ? 1 : -1;
    nf += tag[q] ? 1 : 0;
  }
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++) {
        const size_t q = bq(i, j, k, jm, km);
        if (tag[q]) continue;
        for (int c = 0; c < 18; c++) {
          int ii = i + E18[c][0], jj = j + E18[c][1], kk = k + E18[c][2];
          if (per[0]) ii = (ii - 1 + im) % im + 1;
          if (per[1]) jj = (jj - 1 + jm) % jm + 1;
          if (per[2]) kk = (kk - 1 + km) % km + 1;
          if (ii < 1 || ii > im || jj < 1 || jj > jm || kk < 1 || kk > km) continue;
          if (tag[bq(ii, jj, kk, jm, km)]) { b[q] = 0; break; }
        }
      }
  return nf;
}

int64_t zfh_geom_tube(int32_t *b, int im, int jm, int km, int axis, double r, int periodic,
                      int inlet_code, int outlet_code) {
  if (!b || im < 3 || jm < 3 || km < 3 || axis < 0 || axis > 2 || r <= 0) return -1;
  const int dims[3] = {im, jm, km};
  uint8_t *tag = calloc((size_t)im * jm * km, 1);
  if (!tag) return -1;
  double a[3], c[3];
  for (int d = 0; d < 3; d++) a[d] = c[d] = (dims[d] - 1) / 2.0;
  a[axis] = -r - 2;
  c[axis] = dims[axis] + r + 1;
  /* a straight cylinder: rasterise with the end caps outside the box */
  raster_segment(tag, im, jm, km, a, c, r, 1);
  int per[3] = {0, 0, 0};
  per[axis] = periodic ? 1 : 0;
  const int64_t nf = finish(b, tag, im, jm, km, per);
  if (!periodic) {
    for (int i = 1; i <= im; i++)
      for (int j = 1; j <= jm; j++)
        for (int k = 1; k <= km; k++) {
          const int x[3] = {i, j, k};
          const size_t q = bq(i, j, k, jm, km);
          if (b[q] != 1) continue;
          if (x[axis] == 1) b[q] = inlet_code ? inlet_code : 0;
          else if (x[axis] == dims[axis]) b[q] = outlet_code ? outlet_code : 0;
        }
  }
  free(tag);
  return nf;
}

int64_t zfh_geom_vessel_pair(int32_t *b, int im, int jm, int km, double r_main, double r_conn,
                             double amp, int n_conn, uint64_t seed) {
  if (!b || im < 16 || jm < 16 || km < 16 || r_main <= 0 || n_conn < 0) return -1;
  uint8_t *tag = calloc((size_t)im * jm * km, 1);
  if (!tag) return -1;
  uint64_t s = seed;
  const double jda = 0.25 * (jm - 1), jpcv = 0.75 * (jm - 1), kc = 0.5 * (km - 1);
  /* seeded wavelength and phase of the two centrelines (oscillating in k) */
  const double lam_da = im * (0.5 + 0.5 * urand(&s)), ph_da = 2 * M_PI * urand(&s);
  const double lam_pv = im * (0.5 + 0.5 * urand(&s)), ph_pv = 2 * M_PI * urand(&s);
#define KDA(x) (kc + amp * sin(2 * M_PI * (x) / lam_da + ph_da))
#define KPV(x) (kc + amp * sin(2 * M_PI * (x) / lam_pv + ph_pv))
  const double pcv_end = (im - 1) - r_main - 4.0;
  /* DA runs through both i faces; the PCV is capped before the far face */
  for (double x = -r_main - 2; x < im + r_main + 1; x += 1.0) {
    const double a[3] = {x, jda, KDA(x)}, c[3] = {x + 1, jda, KDA(x + 1)};
    raster_segment(tag, im, jm, km, a, c, r_main, 1);
  }
  for (double x = -r_main - 2; x < pcv_end; x += 1.0) {
    const double xe = fmin(x + 1, pcv_end);
    const double a[3] = {x, jpcv, KPV(x)}, c[3] = {xe, jpcv, KPV(xe)};
    raster_segment(tag, im, jm, km, a, c, r_main, 2);
  }
  /* connectors at seeded, well-separated i positions in the middle 70 % */
  for (int n = 0; n < n_conn; n++) {
    const double lo = 0.15 * im + n * 0.7 * im / n_conn;
    const double x = lo + (0.25 + 0.5 * urand(&s)) * 0.7 * im / n_conn;
    const double a[3] = {x, jda, KDA(x)}, c[3] = {x, jpcv, KPV(x)};
    raster_segment(tag, im, jm, km, a, c, r_conn, 3);
  }
#undef KDA
#undef KPV
  const int per[3] = {0, 0, 0};
  const int64_t nf = finish(b, tag, im, jm, km, per);
  for (int j = 1; j <= jm; j++)
    for (int k = 1; k <= km; k++) {
      const size_t q0 = bq(1, j, k, jm, km), q1 = bq(im, j, k, jm, km);
      if (b[q0] == 1) b[q0] = tag[q0] == 2 ? ZF_B_PRESSURE_OUTLET : ZF_B_VELOCITY_INLET;
      if (b[q1] == 1) b[q1] = tag[q1] == 1 ? ZF_B_PRESSURE_OUTLET : 0;
    }
  free(tag);
  return nf;
}

/* One tree segment list entry. */
typedef struct { double a[3], dir[3], r, len; int gen; } branch_t;

int64_t zfh_geom_tree(int32_t *b, int im, int jm, int km, double r_root, double ratio,
                      int generations, double angle_lo, double angle_hi, uint64_t seed) {
  if (!b || im < 16 || jm < 16 || km < 16 || r_root <= 0 || generations < 1) return -1;
  uint8_t *tag = calloc((size_t)im * jm * km, 1);
  if (!tag) return -1;
  uint64_t s = seed;
  const int maxb = (1 << (generations + 1));
  branch_t *br = calloc(maxb, sizeof *br);
  int nb = 0;
  /* root along +i from beyond the inlet face */
  const double seglen = (double)(im - 1) / (generations + 1);
  br[nb++] = (branch_t){{-r_root - 2, 0.5 * (jm - 1), 0.5 * (km - 1)}, {1, 0, 0}, r_root,
                        seglen + r_root + 2, 0};
  for (int q = 0; q < nb; q++) {
    const branch_t p = br[q];
    double e[3];
    for (int d = 0; d < 3; d++) e[d] = p.a[d] + p.dir[d] * p.len;
    const int leaf = p.gen == generations;
    double len = p.len;
    if (leaf) {
      /* leaves run on to the outlet face along their direction */
      const double need = p.dir[0] > 1e-6 ? (im + p.r + 1 - p.a[0]) / p.dir[0] : p.len;
      len = fmax(p.len, need);
      for (int d = 0; d < 3; d++) e[d] = p.a[d] + p.dir[d] * len;
    }
    raster_segment(tag, im, jm, km, p.a, e, p.r, 1);
    if (leaf) continue;
    /* bifurcate in a seeded plane containing the parent direction */
    const double phi = 2 * M_PI * urand(&s);
    for (int side = -1; side <= 1; side += 2) {
      const double ang = (angle_lo + (angle_hi - angle_lo) * urand(&s)) * M_PI / 180.0;
      double d[3] = {cos(ang), side * sin(ang) * cos(phi), side * sin(ang) * sin(phi)};
      /* keep children inside the cross-section: reflect toward the axis */
      const double yc = e[1] + d[1] * seglen, zc = e[2] + d[2] * seglen;
      const double rr = p.r * ratio + 2;
      if (yc < rr || yc > jm - 1 - rr) d[1] = -d[1];
      if (zc < rr || zc > km - 1 - rr) d[2] = -d[2];
      br[nb++] = (branch_t){{e[0], e[1], e[2]}, {d[0], d[1], d[2]}, p.r * ratio, seglen, p.gen + 1};
    }
  }
  /* branches that reach a side face (j or k) are cut one node short of it,
   * so the wall layer closes them inside the box */
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++)
        if (j == 1 || j == jm || k == 1 || k == km) tag[bq(i, j, k, jm, km)] = 0;
  const int per[3] = {0, 0, 0};
  const int64_t nf = finish(b, tag, im, jm, km, per);
  for (int j = 1; j <= jm; j++)
    for (int k = 1; k <= km; k++) {
      const size_t q0 = bq(1, j, k, jm, km), q1 = bq(im, j, k, jm, km);
      if (b[q0] == 1) b[q0] = ZF_B_VELOCITY_INLET;
      if (b[q1] == 1) b[q1] = ZF_B_PRESSURE_OUTLET;
    }
  free(br);
  free(tag);
  return nf;
}
