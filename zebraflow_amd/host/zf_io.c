/*
 * zf_io.c -- the reference's text formats for the lattice and fluid state.
 *   FLUIDdomainBindex.csv   read at ZebraflowWT.c:2574-2619
 *   FLUIDdomain-%09d.csv    written by OUTPUT(), ZebraflowWT.c:359-373
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zf_host.h"

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

int zfh_write_bindex_csv(const char *path, const int32_t *b, int im, int jm, int km, double x0,
                         double y0, double z0, double dx) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "Bindex,X,Y,Z\n");
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++) {
        const int v = b[bq(i, j, k, jm, km)];
        if (v < 0) continue;
        fprintf(f, "%d,%f,%f,%f\n", v, x0 + (i - 1) * dx, y0 + (j - 1) * dx, z0 + (k - 1) * dx);
      }
  return fclose(f) == 0 ? 0 : -1;
}

/* Same parse as the reference: count lines, skip the header, then
 * "%d%c%f%c%f%c%f" per row and (int)((x-X0)/DX + 1.01) per axis.  Rows that
 * land outside the box are rejected (the reference would write out of
 * bounds). */
int64_t zfh_read_bindex_csv(const char *path, int32_t *b, int im, int jm, int km, double x0,
                            double y0, double z0, double dx) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  const size_t n = (size_t)im * jm * km;
  for (size_t q = 0; q < n; q++) b[q] = -1;
  char line[256];
  if (!fgets(line, sizeof line, f)) { fclose(f); return -1; }
  int64_t rows = 0;
  int bv;
  float xp, yp, zp;
  char c1, c2, c3;
  while (fscanf(f, "%d%c%f%c%f%c%f", &bv, &c1, &xp, &c2, &yp, &c3, &zp) == 7) {
    const int i = (int)((xp - x0) / dx + 1.01), j = (int)((yp - y0) / dx + 1.01),
              k = (int)((zp - z0) / dx + 1.01);
    if (i < 1 || i > im || j < 1 || j > jm || k < 1 || k > km) { fclose(f); return -1; }
    b[bq(i, j, k, jm, km)] = bv;
    rows++;
  }
  fclose(f);
  return rows;
}

int zfh_write_fluid_csv(const char *path, const int32_t *b, int im, int jm, int km, double x0,
                        double y0, double z0, double dx, const double *d, const double *u,
                        const double *v, const double *w, const double *force, int64_t dommax) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "X,Y,Z,Bindex,U,V,W,D,Bx,By,Bz,EulNeigh,Windex\n");
  int64_t o = 0;
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++) {
        const int bv = b[bq(i, j, k, jm, km)];
        if (bv < 0) continue;
        if (o >= dommax) { fclose(f); return -1; }
        const double fx = force ? force[o] : 0, fy = force ? force[dommax + o] : 0,
                     fz = force ? force[2 * dommax + o] : 0;
        fprintf(f, "%f,%f,%f,%d,%.7e,%.7e,%.7e,%.7e,%.7e,%.7e,%.7e,%d,%d\n", x0 + (i - 1) * dx,
                y0 + (j - 1) * dx, z0 + (k - 1) * dx, bv, (float)u[o], (float)v[o], (float)w[o],
                (float)d[o], (float)fx, (float)fy, (float)fz, 0, 0);
        o++;
      }
  return fclose(f) == 0 ? 0 : -1;
}
