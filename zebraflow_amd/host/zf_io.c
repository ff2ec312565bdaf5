/*
 * zf_io.c -- the reference's text formats for the lattice and fluid state.
 *   FLUIDdomainBindex.csv   read at ZebraflowWT.c:2574-2619
 *   FLUIDdomainStart.*      read at ZebraflowWT.c:2626-3387 (and the
 *                           parents' start files, :3394-3620)
 *   FLUIDdomain-%09d.csv    written by OUTPUT(), ZebraflowWT.c:359-373
 *   AllCellMemb-%09d.csv    written by OUTPUT(), ZebraflowWT.c:435-446
 *   VolumeAreaHistory.dat   appended by OUTPUT(), ZebraflowWT.c:346-351
 *   RBC and wall meshes     ZebraflowWT.c:1601-1626, the post-processor's
 *                           WallMesh*.dat readers
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
                        const double *v, const double *w, const double *force,
                        const int32_t *eulneigh, const int32_t *windex, int64_t dommax) {
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
                (float)d[o], (float)fx, (float)fy, (float)fz, eulneigh ? eulneigh[o] : 0,
                windex ? windex[bq(i, j, k, jm, km)] : 0);
        o++;
      }
  return fclose(f) == 0 ? 0 : -1;
}

/* ---- the fluid start state ----------------------------------------------- */

/* The start state in OUTPUT()'s 13-column layout, read as the reference
 * reads FLUIDdomainStart.aa .. .aj (ZebraflowWT.c:2626-3387, the pieces of
 * one CSV split by lines, so only the first holds the header row) and the
 * parents' *parentdomainStart.csv (:3394-3620).  Per row
 * "%f,%f,%f,%d,%f,%f,%f,%f,%f,%f,%f,%d,%d" = X,Y,Z,Bindex,U,V,W,D,Bx,By,Bz,
 * EulNeigh,Windex; a row with Bindex >= 0 sets D, U, V, W and the force of
 * the node (int)((X-X0)/DX + 1.01), ... at its list position p = DOMID - 1 of
 * the box b (lexicographic over b >= 0, i outer, k inner, ZebraflowWT.c:
 * 1481-1502); the EulNeigh and Windex columns are not used (the cell list
 * and W_index are rebuilt from the cells and the wall mesh).  Rows outside
 * the box, or on a node that b does not list, are an error (the reference
 * would write outside its arrays).  Returns the rows applied or < 0. */
int64_t zfh_read_fluid_start(int32_t npaths, const char *const *paths, const int32_t *b, int im,
                             int jm, int km, double x0, double y0, double z0, double dx,
                             int64_t dommax, double *d, double *u, double *v, double *w,
                             double *force) {
  if (npaths <= 0 || !paths || !b || !d || !u || !v || !w) return -1;
  const size_t n = (size_t)im * jm * km;
  int32_t *pos = malloc(sizeof(int32_t) * (n ? n : 1));
  if (!pos) return -1;
  int64_t nl = 0;
  for (size_t q = 0; q < n; q++) pos[q] = b[q] >= 0 ? (int32_t)nl++ : -1;
  if (nl != dommax) { free(pos); return -2; }
  int64_t applied = 0;
  for (int32_t fi = 0; fi < npaths; fi++) {
    FILE *f = fopen(paths[fi], "r");
    if (!f) { free(pos); return -1; }
    char line[512];
    if (fi == 0 && !fgets(line, sizeof line, f)) { fclose(f); free(pos); return -1; }
    float xp, yp, zp, uu, vv, ww, dd, bx, by, bz;
    int bv, eul, win;
    char c[12];
    while (fscanf(f, "%f%c%f%c%f%c%d%c%f%c%f%c%f%c%f%c%f%c%f%c%f%c%d%c%d", &xp, &c[0], &yp, &c[1],
                  &zp, &c[2], &bv, &c[3], &uu, &c[4], &vv, &c[5], &ww, &c[6], &dd, &c[7], &bx,
                  &c[8], &by, &c[9], &bz, &c[10], &eul, &c[11], &win) == 25) {
      if (bv < 0) continue;
      const int i = (int)((xp - x0) / dx + 1.01), j = (int)((yp - y0) / dx + 1.01),
                k = (int)((zp - z0) / dx + 1.01);
      if (i < 1 || i > im || j < 1 || j > jm || k < 1 || k > km) { fclose(f); free(pos); return -3; }
      const int32_t p = pos[bq(i, j, k, jm, km)];
      if (p < 0) { fclose(f); free(pos); return -3; }
      d[p] = dd; u[p] = uu; v[p] = vv; w[p] = ww;
      if (force) { force[p] = bx; force[dommax + p] = by; force[2 * dommax + p] = bz; }
      applied++;
    }
    fclose(f);
  }
  free(pos);
  return applied;
}

/* Populations at equilibrium with the given moments, as the reference
 * initialises them from its start state (ZebraflowWT.c:5253-5262:
 * F = F_OLD = feq(D, U), C_LBM = 1), uploaded with the force field
 * (NULL or all zero: none). */
int zfh_init_from_macro(zf_ctx *ctx, const double *d, const double *u, const double *v,
                        const double *w, const double *force) {
  if (!ctx || !d || !u || !v || !w) return ZF_ERR_ARG;
  zf_lattice_info info;
  int err = zf_get_lattice_info(ctx, &info);
  if (err) return err;
  const size_t n = (size_t)info.dommax;
  double *f = malloc(sizeof(double) * 19 * (n ? n : 1));
  if (!f) return ZF_ERR_NOMEM;
  static const int E[19][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0}, {0, 0, 1},
                               {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0},
                               {1, 0, 1},  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {0, 1, 1},
                               {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 0, 0}};
  for (size_t p = 0; p < n; p++) {
    const double usq = u[p] * u[p] + v[p] * v[p] + w[p] * w[p];
    for (int c = 0; c < 19; c++) {
      const double wc = c < 6 ? 1.0 / 18.0 : c < 18 ? 1.0 / 36.0 : 1.0 / 3.0;
      const double cu = u[p] * E[c][0] + v[p] * E[c][1] + w[p] * E[c][2];
      f[(size_t)c * n + p] = d[p] * wc * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    }
  }
  err = zf_upload_populations(ctx, f);
  free(f);
  if (err) return err;
  int any = 0;
  for (size_t i = 0; force && i < 3 * n && !any; i++) any = force[i] != 0.0;
  return zf_upload_force(ctx, any ? force : NULL);
}

/* ---- meshes: the reference's VTK-style .dat files ------------------------- */

/* Skips `lines` header lines. */
static int skip_lines(FILE *f, int lines) {
  char buf[512];
  for (int l = 0; l < lines; l++)
    if (!fgets(buf, sizeof buf, f)) return -1;
  return 0;
}

/* "POINTS n type" then whitespace-separated coordinates
 * (ZFISHRBCNODES.dat, ZebraflowWT.c:1602-1613; WallMeshNodeCoord.dat,
 * PostProWSSandPressure.c:57-65). */
int64_t zfh_read_points(const char *path, double *xyz, int64_t cap) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char kw[32];
  long long n = 0;
  if (fscanf(f, "%31s %lld", kw, &n) != 2 || strcmp(kw, "POINTS") != 0 || n < 0) {
    fclose(f);
    return -1;
  }
  if (skip_lines(f, 1)) { fclose(f); return -1; }
  if (!xyz) { fclose(f); return n; }
  if (n > cap) { fclose(f); return -2; }
  for (long long i = 0; i < 3 * n; i++)
    if (fscanf(f, "%lf", &xyz[i]) != 1) { fclose(f); return -1; }
  fclose(f);
  return n;
}

/* "CELLS n m" then n triples of 1-based node ids (ZFISHRBCTRIANGLES.dat,
 * ZebraflowWT.c:1617-1626). */
int64_t zfh_read_rbc_triangles(const char *path, int32_t *tri, int64_t cap) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char kw[32];
  long long n = 0, m = 0;
  if (fscanf(f, "%31s %lld %lld", kw, &n, &m) != 3 || strcmp(kw, "CELLS") != 0 || n < 0) {
    fclose(f);
    return -1;
  }
  if (!tri) { fclose(f); return n; }
  if (n > cap) { fclose(f); return -2; }
  for (long long i = 0; i < 3 * n; i++)
    if (fscanf(f, "%d", &tri[i]) != 1) { fclose(f); return -1; }
  fclose(f);
  return n;
}

/* "POLYGONS n m" then n rows "3 a b c" of 0-based ids, stored 1-based
 * (WallMeshCells.dat, PostProWSSandPressure.c:84-96). */
int64_t zfh_read_wall_cells(const char *path, int32_t *tri, int64_t cap) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char kw[32];
  long long n = 0, m = 0;
  if (fscanf(f, "%31s %lld %lld", kw, &n, &m) != 3 || strcmp(kw, "POLYGONS") != 0 || n < 0) {
    fclose(f);
    return -1;
  }
  if (!tri) { fclose(f); return n; }
  if (n > cap) { fclose(f); return -2; }
  for (long long t = 0; t < n; t++) {
    int k, a, b, c;
    if (fscanf(f, "%d %d %d %d", &k, &a, &b, &c) != 4 || k != 3) { fclose(f); return -1; }
    tri[3 * t] = a + 1;
    tri[3 * t + 1] = b + 1;
    tri[3 * t + 2] = c + 1;
  }
  fclose(f);
  return n;
}

/* "CELL_DATA n", "NORMALS Normals float", then n normals
 * (WallMeshCellNormals.dat). */
int64_t zfh_read_wall_normals(const char *path, double *nrm, int64_t cap) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char kw[32];
  long long n = 0;
  if (fscanf(f, "%31s %lld", kw, &n) != 2 || strcmp(kw, "CELL_DATA") != 0 || n < 0) {
    fclose(f);
    return -1;
  }
  if (skip_lines(f, 2)) { fclose(f); return -1; }
  if (!nrm) { fclose(f); return n; }
  if (n > cap) { fclose(f); return -2; }
  for (long long i = 0; i < 3 * n; i++)
    if (fscanf(f, "%lf", &nrm[i]) != 1) { fclose(f); return -1; }
  fclose(f);
  return n;
}

int zfh_write_points(const char *path, const double *xyz, int64_t n) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "POINTS %lld double\n", (long long)n);
  for (int64_t i = 0; i < n; i++)
    fprintf(f, "%.15g %.15g %.15g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  return fclose(f) == 0 ? 0 : -1;
}

int zfh_write_rbc_triangles(const char *path, const int32_t *tri, int64_t n) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "CELLS %lld %lld\n", (long long)n, (long long)(4 * n));
  for (int64_t t = 0; t < n; t++) fprintf(f, "%d %d %d\n", tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]);
  return fclose(f) == 0 ? 0 : -1;
}

int zfh_write_wall_mesh(const char *nodes_path, const char *cells_path, const char *normals_path,
                        const double *xyz, int64_t nwn, const int32_t *tri, const double *nrm,
                        int64_t nwt) {
  FILE *f = fopen(nodes_path, "w");
  if (!f) return -1;
  fprintf(f, "POINTS %lld float\n", (long long)nwn);
  for (int64_t i = 0; i < nwn; i++) fprintf(f, "%.9g %.9g %.9g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  if (fclose(f)) return -1;
  if (!(f = fopen(cells_path, "w"))) return -1;
  fprintf(f, "POLYGONS %lld %lld\n", (long long)nwt, (long long)(4 * nwt));
  for (int64_t t = 0; t < nwt; t++) fprintf(f, "3 %d %d %d\n", tri[3 * t] - 1, tri[3 * t + 1] - 1, tri[3 * t + 2] - 1);
  if (fclose(f)) return -1;
  if (!(f = fopen(normals_path, "w"))) return -1;
  fprintf(f, "CELL_DATA %lld\nNORMALS Normals float\n", (long long)nwt);
  for (int64_t t = 0; t < nwt; t++) fprintf(f, "%.9g %.9g %.9g\n", nrm[3 * t], nrm[3 * t + 1], nrm[3 * t + 2]);
  return fclose(f) == 0 ? 0 : -1;
}

/* ---- cells: the AllCellMemb layout (ZebraflowWT.c:435-446, read back as
 * AllCellMembStart at :5327-5430) -------------------------------------------- */

static const char *kCellHeader =
    "X,Y,Z,Cell,CellStatus,Vx,Vy,Vz,WallTri,WallRepulfX,WallRepulfY,WallRepulfZ,regressflag,"
    "Aratio,Fx,Fy,Fz,FVx,FVy,FVz,OutNeighFlag\n";

/* Reads node rows (header, then nn rows per cell in node order): X, Y, Z,
 * Cell (1-based), CellStatus and 16 further columns.  xyz: cap_cells*nn*3,
 * status: cap_cells.  Returns the number of cells (rows / nn) or < 0. */
int zfh_read_cells_csv(const char *path, int nn, double *xyz, int32_t *status, int cap_cells) {
  FILE *f = fopen(path, "r");
  if (!f || nn <= 0) { if (f) fclose(f); return -1; }
  if (skip_lines(f, 1)) { fclose(f); return -1; }
  char line[1024];
  int64_t row = 0;
  while (fgets(line, sizeof line, f)) {
    double x, y, z;
    int cell, st;
    if (sscanf(line, "%lf,%lf,%lf,%d,%d", &x, &y, &z, &cell, &st) != 5) break;
    const int64_t c = row / nn, node = row % nn;
    if (c >= cap_cells) { fclose(f); return -2; }
    xyz[3 * (c * nn + node)] = x;
    xyz[3 * (c * nn + node) + 1] = y;
    xyz[3 * (c * nn + node) + 2] = z;
    if (status) status[c] = st;
    row++;
  }
  fclose(f);
  if (row % nn) return -1;
  return (int)(row / nn);
}

/* Writes the cells with status > -5000.  vel, vrep, force: ncell*nn*3;
 * wall_tri: ncell*nn (any of these may be NULL: zeros).  The reference's
 * regress flag, area ratio, volume-force split and neighbour flag columns
 * have no counterpart here and are written as 0. */
int zfh_write_cells_csv(const char *path, int ncell, int nn, const double *xyz, const double *vel,
                        const int32_t *status, const int32_t *wall_tri, const double *vrep,
                        const double *force) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fputs(kCellHeader, f);
  for (int c = 0; c < ncell; c++) {
    if (status && status[c] <= -5000) continue;
    for (int n = 0; n < nn; n++) {
      const size_t q = (size_t)c * nn + n;
      const double *x = &xyz[3 * q];
      const double v0 = vel ? vel[3 * q] : 0, v1 = vel ? vel[3 * q + 1] : 0, v2 = vel ? vel[3 * q + 2] : 0;
      const double r0 = vrep ? vrep[3 * q] : 0, r1 = vrep ? vrep[3 * q + 1] : 0, r2 = vrep ? vrep[3 * q + 2] : 0;
      const double f0 = force ? force[3 * q] : 0, f1 = force ? force[3 * q + 1] : 0,
                   f2 = force ? force[3 * q + 2] : 0;
      fprintf(f, "%f,%f,%f,%d,%d,%.7e,%.7e,%.7e,%d,%.7e,%.7e,%.7e,%d,%.7e,%.7e,%.7e,%.7e,%.7e,%.7e,%.7e,%d\n",
              x[0], x[1], x[2], c + 1, status ? status[c] : 1, v0, v1, v2, wall_tri ? wall_tri[q] : 0,
              r0, r1, r2, 0, 0.0, f0, f1, f2, 0.0, 0.0, 0.0, 0);
    }
  }
  return fclose(f) == 0 ? 0 : -1;
}

/* "LINES n m" then n pairs of 1-based node ids (NucleusEdges.dat,
 * ZebraflowWT.c:1903-1911).  Returns n (edge NULL: count only), -2 past cap. */
int64_t zfh_read_lines(const char *path, int32_t *edge, int64_t cap) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char kw[32];
  long long n = 0, m = 0;
  if (fscanf(f, "%31s %lld %lld", kw, &n, &m) != 3 || strcmp(kw, "LINES") != 0 || n < 0) {
    fclose(f);
    return -1;
  }
  if (!edge) { fclose(f); return n; }
  if (n > cap) { fclose(f); return -2; }
  for (long long i = 0; i < 2 * n; i++)
    if (fscanf(f, "%d", &edge[i]) != 1) { fclose(f); return -1; }
  fclose(f);
  return n;
}

/* The nuclei of the start cells, AllCellNucStart.csv (ZebraflowWT.c:5530-5585;
 * the layout OUTPUT() writes as AllCellNucleus, :453-463): a header, then
 * nnuc rows per cell, X,Y,Z first.  Returns the cells read. */
int zfh_read_nuclei_csv(const char *path, int nnuc, double *xyz, int cap_cells) {
  FILE *f = fopen(path, "r");
  if (!f || nnuc <= 0) { if (f) fclose(f); return -1; }
  if (skip_lines(f, 1)) { fclose(f); return -1; }
  char line[1024];
  int64_t row = 0;
  while (fgets(line, sizeof line, f)) {
    double x, y, z;
    if (sscanf(line, "%lf,%lf,%lf", &x, &y, &z) != 3) break;
    const int64_t c = row / nnuc, node = row % nnuc;
    if (c >= cap_cells) { fclose(f); return -2; }
    xyz[3 * (c * nnuc + node)] = x;
    xyz[3 * (c * nnuc + node) + 1] = y;
    xyz[3 * (c * nnuc + node) + 2] = z;
    row++;
  }
  fclose(f);
  if (row % nnuc) return -1;
  return (int)(row / nnuc);
}

/* AllCellNucleus-%09d.csv (ZebraflowWT.c:453-463): X,Y,Z,Cell,CellStatus,
 * Vx,Vy,Vz,Fx,Fy,Fz,Neigh,Neighflag per nucleus node of every cell with
 * status > -5000; vel and force ncell*nnuc*3 (NULL: zeros); the neighbour
 * columns (the reference's list bookkeeping) are written as 0. */
int zfh_write_nuclei_csv(const char *path, int ncell, int nnuc, const double *xyz,
                         const double *vel, const double *force, const int32_t *status) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "X,Y,Z,Cell,CellStatus,Vx,Vy,Vz,Fx,Fy,Fz,Neigh,Neighflag\n");
  for (int c = 0; c < ncell; c++) {
    if (status && status[c] <= -5000) continue;
    for (int n = 0; n < nnuc; n++) {
      const size_t q = (size_t)c * nnuc + n;
      const double *x = &xyz[3 * q];
      const double v0 = vel ? vel[3 * q] : 0, v1 = vel ? vel[3 * q + 1] : 0, v2 = vel ? vel[3 * q + 2] : 0;
      const double f0 = force ? force[3 * q] : 0, f1 = force ? force[3 * q + 1] : 0,
                   f2 = force ? force[3 * q + 2] : 0;
      fprintf(f, "%f,%f,%f,%d,%d,%.7e,%.7e,%.7e,%.7e,%.7e,%.7e,%d,%d\n", x[0], x[1], x[2], c + 1,
              status ? status[c] : 1, v0, v1, v2, f0, f1, f2, 0, 0);
    }
  }
  return fclose(f) == 0 ? 0 : -1;
}

/* Appends one row per live cell to VolumeAreaHistory.dat
 * (ZebraflowWT.c:346-351): Iteration#,Cell,Vol,Area,CenX,CenY,CenZ,
 * CellState,RegressFlag; the header is written when the file is new. */
int zfh_append_volume_area(const char *path, int iter, int ncell, const double *vol,
                           const double *area, const double *centroid, const int32_t *status) {
  FILE *t = fopen(path, "r");
  const int fresh = t == NULL;
  if (t) fclose(t);
  FILE *f = fopen(path, "a");
  if (!f) return -1;
  if (fresh) fprintf(f, "Iteration#,Cell,Vol,Area,CenX,CenY,CenZ,CellState,RegressFlag\n");
  for (int c = 0; c < ncell; c++)
    if (!status || status[c] > -5000)
      fprintf(f, "%d,%d,%.7e,%.7e,%.7e,%.7e,%.7e,%d,%d\n", iter, c + 1, vol[c], area[c],
              centroid[3 * c], centroid[3 * c + 1], centroid[3 * c + 2], status ? status[c] : 1, 0);
  return fclose(f) == 0 ? 0 : -1;
}
