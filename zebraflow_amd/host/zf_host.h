/*
 * zf_host.h -- host side of the engine (C): synthetic geometry, the
 * reference's input/output file formats and its step/output loop.
 * Built into libzfhost.so and the zf_host program; calls libzfgpu.so only
 * through include/zf.h.
 */
#ifndef ZF_HOST_H
#define ZF_HOST_H

#include <stdint.h>

#include "zf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- synthetic geometry -----------------------------------------------------
 * All generators fill a dense box b[(i-1)*jm*km + (j-1)*km + (k-1)] with
 * -1 (outside), 0 (wall layer: non-fluid nodes with a D3Q19 fluid
 * neighbour), 1 (fluid) and the inlet/outlet codes of zf.h.  They return the
 * number of fluid (B==1) nodes, or a negative value on bad arguments. */

/* Straight tube of radius r along `axis` (0=i, 1=j, 2=k), centred in the
 * cross-section.  periodic: the tube runs through both end faces; otherwise
 * its end faces carry inlet_code / outlet_code (0 = closed wall). */
int64_t zfh_geom_tube(int32_t *b, int im, int jm, int km, int axis, double r, int periodic,
                      int inlet_code, int outlet_code);

/* Zebrafish-trunk vessel pair along i: a dorsal-aorta tube (j centre
 * 0.25*jm) and a posterior-cardinal-vein tube (j centre 0.75*jm), both of
 * radius r_main with seeded sinusoidal centrelines of amplitude `amp`,
 * joined by n_conn segmental connector tubes of radius r_conn at seeded i
 * positions.  DA: velocity inlet at i=1, pressure outlet at i=im.  PCV:
 * pressure outlet at i=1, closed (capped) before i=im. */
int64_t zfh_geom_vessel_pair(int32_t *b, int im, int jm, int km, double r_main, double r_conn,
                             double amp, int n_conn, uint64_t seed);

/* Branching tree along i: root of radius r_root from i=1 (velocity inlet),
 * `generations` levels of bifurcations, child radius = ratio * parent,
 * branch half-angles drawn in [angle_lo, angle_hi] degrees; every leaf that
 * reaches the i=im face is a pressure outlet, others are capped. */
int64_t zfh_geom_tree(int32_t *b, int im, int jm, int km, double r_root, double ratio,
                      int generations, double angle_lo, double angle_hi, uint64_t seed);

/* ---- membrane meshes ---------------------------------------------------------
 * Icosphere of `level` subdivisions (level 3: 642 nodes, 1280 triangles),
 * optionally mapped to the Evans-Fung biconcave disc, scaled to `radius`.
 * xyz: 3*node_cap node-major; tri: 3*tri_cap, 1-based, outward winding.
 * Returns the node count (triangles in *n_tri) or -1. */
int zfh_mesh_icosphere(int level, double radius, int biconcave, double *xyz, int32_t *tri,
                       int node_cap, int tri_cap, int32_t *n_tri);
/* Seeded non-overlapping placement of up to ncell copies of the reference
 * shape `ref` (nn nodes, centred at the origin) inside interior fluid of b.
 * Positions are lattice coordinates, node (i,j,k) at (i-1,j-1,k-1).
 * xyz_out: ncell*nn*3.  Returns the number placed. */
int zfh_place_cells(const int32_t *b, int im, int jm, int km, const double *ref, int nn,
                    double r_cell, double gap, int ncell, uint64_t seed, int max_tries,
                    double *xyz_out);
/* Denser placement for cell-laden runs: centres drawn from interior fluid
 * nodes, the overlap test on the mesh nodes themselves (no node within gap
 * of another cell's node, every node's lattice neighbourhood interior
 * fluid).  Returns cells placed, or -1 on bad arguments. */
int zfh_place_cells_dense(const int32_t *b, int im, int jm, int km, const double *ref, int nn,
                          double gap, int ncell, uint64_t seed, int max_tries, double *xyz_out);

/* Lumen wall mesh of a B field by surface nets on the fluid indicator B > 0:
 * vertices at lattice-cube crossings, one quad per crossing lattice edge,
 * normals into the fluid.  Coordinates x0 + (i-1) dx.  With xyz and tri NULL
 * only counts.  tri: 1-based.  Returns 0, -2 if a capacity is too small. */
int zfh_wall_mesh(const int32_t *b, int im, int jm, int km, double x0, double y0, double z0,
                  double dx, double *xyz, int64_t node_cap, int32_t *tri, double *normal,
                  int64_t tri_cap, int64_t *n_node, int64_t *n_tri);

/* ---- reference file formats ----------------------------------------------- */
/* FLUIDdomainBindex.csv (ZebraflowWT.c:2574-2619): header, then
 * "bvalue,x,y,z" for every node with B >= 0; x = x0 + (i-1)*dx. */
int zfh_write_bindex_csv(const char *path, const int32_t *b, int im, int jm, int km,
                         double x0, double y0, double z0, double dx);
/* Reads it into b (prefilled with -1), index (int)((x-x0)/dx + 1.01)
 * exactly as ZebraflowWT.c:2611-2613.  Returns rows read or -1. */
int64_t zfh_read_bindex_csv(const char *path, int32_t *b, int im, int jm, int km, double x0,
                            double y0, double z0, double dx);
/* FLUIDdomain-%09d.csv in the OUTPUT() layout (ZebraflowWT.c:362-368):
 * X,Y,Z,Bindex,U,V,W,D,Bx,By,Bz,EulNeigh,Windex for every B >= 0 node.
 * u,v,w,d,force are in list order p (force may be NULL). */
int zfh_write_fluid_csv(const char *path, const int32_t *b, int im, int jm, int km, double x0,
                        double y0, double z0, double dx, const double *d, const double *u,
                        const double *v, const double *w, const double *force,
                        const int32_t *eulneigh, const int32_t *windex, int64_t dommax);

/* Meshes (the reference's VTK-style .dat files).  With a NULL array the
 * readers return the count only; -2 when it exceeds cap, -1 on a bad file.
 *   RBC nodes        "POINTS n type", coordinates   (ZebraflowWT.c:1602-1613)
 *   RBC triangles    "CELLS n m", 1-based triples   (ZebraflowWT.c:1617-1626)
 *   wall nodes       "POINTS n type", coordinates
 *   wall triangles   "POLYGONS n m", rows "3 a b c" 0-based (stored 1-based)
 *   wall normals     "CELL_DATA n", "NORMALS ...", one normal per triangle */
int64_t zfh_read_points(const char *path, double *xyz, int64_t cap);
int64_t zfh_read_rbc_triangles(const char *path, int32_t *tri, int64_t cap);
int64_t zfh_read_wall_cells(const char *path, int32_t *tri, int64_t cap);
int64_t zfh_read_wall_normals(const char *path, double *nrm, int64_t cap);
int zfh_write_points(const char *path, const double *xyz, int64_t n);
int zfh_write_rbc_triangles(const char *path, const int32_t *tri, int64_t n);
int zfh_write_wall_mesh(const char *nodes_path, const char *cells_path, const char *normals_path,
                        const double *xyz, int64_t nwn, const int32_t *tri, const double *nrm,
                        int64_t nwt);

/* Cells in the AllCellMemb layout (ZebraflowWT.c:435-446, read back as
 * AllCellMembStart, :5327-5430): 21 columns per node row. */
int zfh_read_cells_csv(const char *path, int nn, double *xyz, int32_t *status, int cap_cells);
/* Nucleus files: NucleusEdges.dat ("LINES n m", ZebraflowWT.c:1903-1911;
 * the nodes use zfh_read_points), the start nuclei AllCellNucStart.csv
 * (:5530-5585) and the output AllCellNucleus-%09d.csv (:453-463). */
int64_t zfh_read_lines(const char *path, int32_t *edge, int64_t cap);
int zfh_read_nuclei_csv(const char *path, int nnuc, double *xyz, int cap_cells);
int zfh_write_nuclei_csv(const char *path, int ncell, int nnuc, const double *xyz,
                         const double *vel, const double *force, const int32_t *status);
int zfh_write_cells_csv(const char *path, int ncell, int nn, const double *xyz, const double *vel,
                        const int32_t *status, const int32_t *wall_tri, const double *vrep,
                        const double *force);
/* VolumeAreaHistory.dat rows (ZebraflowWT.c:346-351), appended. */
int zfh_append_volume_area(const char *path, int iter, int ncell, const double *vol,
                           const double *area, const double *centroid, const int32_t *status);

/* Membrane constants in lattice units for a rest mesh (1-based tri): the
 * reference's moduli as ratios to the shear modulus es (kd = 12 es,
 * ka = 25.67 es, kv = 9.17 es, kT = 0.0036 es l0^2 with l0 the mean rest
 * edge, kBend = 2/sqrt(3) kb), rest area and volume as the global targets. */
void zfh_membrane_config_lattice(const double *xyz, int nn, const int32_t *tri, int nt, double es,
                                 double kb, int max_cells, zf_membrane_config *mc);

/* ---- the step/output loop -------------------------------------------------- */
/* ---- the fluid start state -------------------------------------------------
 * The reference's start files in OUTPUT()'s 13-column layout
 * (FLUIDdomainStart.aa .. .aj, ZebraflowWT.c:2626-3387; the parents'
 * *parentdomainStart.csv, :3394-3620): the first of the npaths files holds
 * the header row.  Rows with Bindex >= 0 fill d, u, v, w (dommax each) and
 * force (3*dommax SoA, may be NULL) at the list position of their node in b.
 * Returns the rows applied, or -1 (file), -2 (dommax does not match b),
 * -3 (a row outside the box or on a node b does not list). */
int64_t zfh_read_fluid_start(int32_t npaths, const char *const *paths, const int32_t *b, int im,
                             int jm, int km, double x0, double y0, double z0, double dx,
                             int64_t dommax, double *d, double *u, double *v, double *w,
                             double *force);
/* F = feq(D, U) at every listed node (ZebraflowWT.c:5253-5262), uploaded,
 * and the force field (NULL or all zero: cleared). */
int zfh_init_from_macro(zf_ctx *ctx, const double *d, const double *u, const double *v,
                        const double *w, const double *force);

/* ---- exact restart files ------------------------------------------------------
 * The state as the engine stores it (populations, the force field of the
 * last RBC iteration, cell positions and status), so that a run continues
 * bit-identically; the reference's own restart re-equilibrates the
 * populations from its CSV output (ZebraflowWT.c:5253-5262).  ncell, nn:
 * the cells on the context (0, 0 for none).  zfh_load_state checks that the
 * file matches the context (storage form, fluid nodes, cells) and returns
 * the iteration it was written at in *iter. */
int zfh_save_state(zf_ctx *ctx, const char *path, int32_t iter, int32_t ncell, int32_t nn);
int zfh_load_state(zf_ctx *ctx, const char *path, int32_t *iter, int32_t ncell, int32_t nn);

typedef struct zfh_run_config {
  const char *out_dir;    /* SimulationOutputData */
  int32_t im, jm, km;     /* the box of b */
  int32_t iter0, steps;   /* iterations iter0+1 .. iter0+steps */
  int32_t output_every;   /* OUTPUT() when iter % output_every == 0 (0: never) */
  double x0, y0, z0, dx;  /* node (i,j,k) at (x0 + (i-1) dx, ...) in the CSVs */
  int32_t ncell, nn, nt;  /* cells on the context (0: none) */
  int32_t have_rep, have_wall;
  int32_t checkpoint_every; /* restart file ZFstate-%09d.bin when iter % checkpoint_every == 0 */
  int32_t nnuc;           /* nucleus nodes per cell (0: none): AllCellNucleus output */
} zfh_run_config;

/* The reference's time loop (ZebraflowWT.c:5646-12412): zf_step per
 * iteration, then on output iterations the fluid CSV, the cell CSV and the
 * volume/area history, as OUTPUT() writes them (ZebraflowWT.c:282-446).
 * b: the box (for the CSV row order).  Returns 0 or the failing code. */
int zfh_run(zf_ctx *ctx, const int32_t *b, const zfh_run_config *rc);

#ifdef __cplusplus
}
#endif
#endif
