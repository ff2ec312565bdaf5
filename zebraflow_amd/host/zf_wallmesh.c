/*
 * zf_wallmesh.c -- lumen wall mesh of a B_index field.
 *
 * The reference reads its vessel walls as VTK-legacy triangle meshes
 * (WallMeshNodeCoord/Cells/CellNormals.dat, ZebraflowWT.c:3715-4643) and
 * uses them for cell-wall repulsion.  For generated geometries the wall is
 * the boundary between fluid (B > 0) and the rest, triangulated by surface
 * nets: one vertex per lattice cube whose corners straddle the boundary, at
 * the mean of its crossing-edge midpoints (the halfway points where
 * bounce-back puts the wall), and one quad (two triangles) per crossing
 * lattice edge.  Triangle normals point into the fluid, as the repulsion
 * (v_rep += k (d0 - d) n) requires.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "zf_host.h"

static inline size_t bq(int i, int j, int k, int jm, int km) {
  return ((size_t)i * jm + (size_t)j) * km + (size_t)k;
}

int zfh_wall_mesh(const int32_t *b, int im, int jm, int km, double x0, double y0, double z0,
                  double dx, double *xyz, int64_t node_cap, int32_t *tri, double *normal,
                  int64_t tri_cap, int64_t *n_node, int64_t *n_tri) {
  if (!b || im < 2 || jm < 2 || km < 2 || !n_node || !n_tri) return -1;
  const int cm = im - 1, cj = jm - 1, ck = km - 1;
  int32_t *vid = malloc(sizeof(int32_t) * (size_t)cm * cj * ck);
  if (!vid) return -1;
  int64_t nv = 0, nt = 0;
#define IN(i, j, k) (b[bq(i, j, k, jm, km)] > 0)
  /* vertices */
  for (int i = 0; i < cm; i++)
    for (int j = 0; j < cj; j++)
      for (int k = 0; k < ck; k++) {
        const size_t c = ((size_t)i * cj + j) * ck + k;
        vid[c] = -1;
        int cnt = 0;
        double s[3] = {0, 0, 0};
        /* 12 cube edges: 4 along each axis */
        for (int a = 0; a < 3; a++)
          for (int u = 0; u < 2; u++)
            for (int v = 0; v < 2; v++) {
              int p0[3], p1[3];
              p0[a] = 0; p1[a] = 1;
              p0[(a + 1) % 3] = p1[(a + 1) % 3] = u;
              p0[(a + 2) % 3] = p1[(a + 2) % 3] = v;
              const int in0 = IN(i + p0[0], j + p0[1], k + p0[2]);
              const int in1 = IN(i + p1[0], j + p1[1], k + p1[2]);
              if (in0 == in1) continue;
              for (int d = 0; d < 3; d++) s[d] += 0.5 * (p0[d] + p1[d]);
              cnt++;
            }
        if (!cnt) continue;
        if (xyz) {
          if (nv >= node_cap) { free(vid); return -2; }
          xyz[3 * nv] = x0 + (i + s[0] / cnt) * dx;
          xyz[3 * nv + 1] = y0 + (j + s[1] / cnt) * dx;
          xyz[3 * nv + 2] = z0 + (k + s[2] / cnt) * dx;
        }
        vid[c] = (int32_t)nv++;
      }
  /* quads around every crossing lattice edge whose four cubes exist */
  for (int i = 0; i < im; i++)
    for (int j = 0; j < jm; j++)
      for (int k = 0; k < km; k++)
        for (int a = 0; a < 3; a++) {
          int e[3] = {0, 0, 0};
          e[a] = 1;
          const int ii = i + e[0], jj = j + e[1], kk = k + e[2];
          if (ii >= im || jj >= jm || kk >= km) continue;
          const int ina = IN(i, j, k), inb = IN(ii, jj, kk);
          if (ina == inb) continue;
          /* cubes around the edge, counter-clockwise about +axis a in the
           * plane of the axes (a+1, a+2) */
          const int o[4][2] = {{-1, -1}, {0, -1}, {0, 0}, {-1, 0}};
          int32_t q[4];
          int ok = 1;
          for (int r = 0; r < 4 && ok; r++) {
            int c[3] = {i, j, k};
            c[(a + 1) % 3] += o[r][0];
            c[(a + 2) % 3] += o[r][1];
            if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= cm || c[1] >= cj || c[2] >= ck) { ok = 0; break; }
            q[r] = vid[((size_t)c[0] * cj + c[1]) * ck + c[2]];
            if (q[r] < 0) ok = 0;
          }
          if (!ok) continue;
          /* counter-clockwise order has normal +axis; flip when the fluid
           * is on the low side so the normal points into the fluid */
          if (ina) { int32_t t = q[1]; q[1] = q[3]; q[3] = t; }
          const int32_t tr[2][3] = {{q[0], q[1], q[2]}, {q[0], q[2], q[3]}};
          for (int s = 0; s < 2; s++) {
            if (tri) {
              if (nt >= tri_cap) { free(vid); return -2; }
              const double *p0 = &xyz[3 * tr[s][0]], *p1 = &xyz[3 * tr[s][1]], *p2 = &xyz[3 * tr[s][2]];
              const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
              const double v[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
              double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
              const double m = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
              for (int d = 0; d < 3; d++) {
                tri[3 * nt + d] = tr[s][d] + 1; /* 1-based, as the reference's mesh */
                normal[3 * nt + d] = m > 0 ? n[d] / m : 0.0;
              }
            }
            nt++;
          }
        }
#undef IN
  free(vid);
  *n_node = nv;
  *n_tri = nt;
  return 0;
}
