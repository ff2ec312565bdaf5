/*
 * zf_oracle.h -- CPU restatement of the Zebraflow time step (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the MI355X engine.  Only tests/, the smoke
 * entry point and bench.py's CPU-baseline leg may load it.  The product path
 * (libzfgpu.so, zf_host) never links or calls it.
 *
 * Every function restates the algorithm of ZebraflowWT.c (the reference, a
 * single C99/OpenMP file) and cites the lines it follows.  Conventions kept
 * from the reference:
 *   - lattice indices i,j,k are 1-based; the box array b_index is dense and
 *     0-based, element (i,j,k) at ((i-1)*jm + (j-1))*km + (k-1)  (i outer,
 *     k inner, as B_index[i][j][k] in ZebraflowWT.c:180);
 *   - list entries (DOMLIST, BCLIST, BCNLIST, DOMID values) are 1-based;
 *   - populations are SoA f[c*n + p] with c = 0..18 for the reference's
 *     directions 1..19 and p = DOMID-1.
 *
 * Pinning status of each function is stated in its comment and in the
 * oracle test files (tests/test_oracle_*.py).
 */
#ifndef ZF_ORACLE_H
#define ZF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* D3Q19 table, ZebraflowWT.c:2471-2504 (Xlattice) and :2329-2333 (Wcoeff). */
extern const int ZFO_E[19][3];
extern const double ZFO_W[19];

/* ---- a1: fluid-node list, LBM_DOM_SETUP ZebraflowWT.c:1481-1502 ----------
 * domid: box-sized, receives 1-based l or -1.  domlist: 3*cap ints.
 * Returns DOMMAX, or -1 if cap is exceeded.  Pinned by reference goldens. */
int zfo_dom_setup(const int32_t *b_index, int im, int jm, int km,
                  int32_t *domid, int32_t *domlist, int cap);

/* ---- a3: BC list, LBM_BC_SETUP ZebraflowWT.c:1506-1589 -------------------
 * bclist: 4*cap (i,j,k,B); bcnlist: 3*cap (i,j,k of the chosen neighbour).
 * periodic[3] extends the reference's out-of-box skip (ZebraflowWT.c:1541)
 * with a wrap on periodic axes; all-zero reproduces the reference exactly.
 * A node with no neighbour gets (0,0,0) and is counted in *n_fail (the
 * reference leaves stale values, ZebraflowWT.c:1575-1581).
 * Returns BCMAX or -1 on overflow.  Pinned by reference goldens. */
int zfo_bc_setup(const int32_t *b_index, int im, int jm, int km,
                 const int periodic[3], int32_t *bclist, int32_t *bcnlist,
                 int cap, int *n_fail);

/* ---- parent (feeder) lattices: BC list, LBM_PARENTBC_SETUP
 * ZebraflowWT.c:1015-1350.  As zfo_bc_setup with the parent's thresholds;
 * codes 10+p (boundary copied from the main domain) and 20+p (upstream)
 * map to the main-domain node (i + off[0], j + off[1], k_over + off[2]);
 * vbcn: 3*cap.  Returns BCMAX or -1 on overflow.  Pinned by reference
 * goldens. */
int zfo_parent_bc_setup(const int32_t *b, int im, int jm, int km, int p, int k_over,
                        const int off[3], int32_t *bclist, int32_t *bcnlist, int32_t *vbcn,
                        int cap, int *n_fail);

/* ---- a2: pull-neighbour list, implicit in ZebraflowWT.c:6113-6125 --------
 * nbr[c*dommax + (l-1)] = DOMID[(i,j,k) - e_c] for c = 0..17 (1-based, or
 * -1 if unlisted/outside the box; periodic axes wrap).  Pinned by brute
 * force and the pairing invariant nbr[opp c][nbr[c][l]] == l. */
void zfo_neighbours(const int32_t *domid, const int32_t *domlist, int dommax,
                    int im, int jm, int km, const int periodic[3],
                    int32_t *nbr);

/* ---- a4-a6: one LBM time step ---------------------------------------------
 * Follows the reference loop structure: stream+collide for interior nodes
 * (0<B<10, ZebraflowWT.c:6101-6145), boundary conditions (:6292-6433),
 * macro recompute + copy (:6601-6626).  Arrays F, F_OLD, FEQ as in the
 * reference (:178), DOMID lookups per direction.
 *
 * Extensions (not in the reference, all default-off):
 *   wall_mode 1   : halfway bounce-back on links whose source has B==0
 *                   instead of the equilibrium wall (B==0 nodes untouched);
 *   body_force[3] : uniform force added to the per-node force field;
 *   B==2001       : velocity inlet feq(D[m], u_in(iter)*inlet_dir);
 *   B==3001       : pressure outlet feq(outlet_rho, U[m]);
 *   periodic[3]   : wrap on the pull.
 */
typedef struct zfo_lbm_params {
  int im, jm, km;
  int periodic[3];
  int wall_mode;            /* 0 = equilibrium wall (reference), 1 = halfway BB */
  double visc_scale;        /* tau = 0.5 + visc_scale*B; reference MU_SCALE*MU = 0.72 */
  double force_dt;          /* DT in the Guo terms; reference 0.5 (ZebraflowWT.c:41) */
  double body_force[3];
  double inlet_u0, inlet_amp, inlet_period, inlet_dir[3];
  double outlet_rho;
  double t_scale;           /* seconds per step for the 1001-1009 waveforms (5e-7) */
  double pressure_scale;    /* LBM_PRESSURE_SCALE (3333.33..) */
} zfo_lbm_params;

typedef struct zfo_lattice {
  int im, jm, km, dommax, bcmax;
  const int32_t *b_index, *domid, *domlist, *bclist, *bcnlist;
} zfo_lattice;

/* F, F_OLD: 19*dommax.  D,U,V,W: dommax.  force: 3*dommax (may be NULL).
 * feq: 19*dommax scratch (the reference's FEQ store) or NULL.
 * On return F_OLD holds the new populations (the reference's copy). */
void zfo_lbm_step_f64(const zfo_lattice *lat, const zfo_lbm_params *p,
                      double *F, double *F_OLD, double *FEQ, double *D,
                      double *U, double *V, double *W, double *force,
                      int iter, int rbc_step, int nthreads);
void zfo_lbm_step_f32(const zfo_lattice *lat, const zfo_lbm_params *p,
                      float *F, float *F_OLD, float *FEQ, float *D, float *U,
                      float *V, float *W, float *force, int iter,
                      int rbc_step, int nthreads);

/* Equilibrium, ZebraflowWT.c:2568 with C_LBM = 1 (EQ_A=3, EQ_B=4.5, EQ_C=-1.5). */
/* zfo_lbm_step split into its phases (0 stream + collide, 1 boundary,
 * 2 macro + copy), and the parent lattices' boundary pass
 * (ZebraflowWT.c:6436-6594) that runs between the phases of a main lattice
 * and its parent: main phase 0, parent phase 0, main phase 1,
 * zfo_parent_boundary, then phase 2 of both. */
void zfo_lbm_phase_f64(const zfo_lattice *lat, const zfo_lbm_params *p, int phase, double *F,
                       double *F_OLD, double *FEQ, double *D, double *U, double *V, double *W,
                       double *force, int iter, int rbc_step, int nthreads);
void zfo_parent_boundary_f64(const zfo_lattice *par, const zfo_lbm_params *pp, int p,
                             const int32_t *vbcn, double *F, const double *D_p,
                             const zfo_lattice *mainl, const double *D, const double *U,
                             const double *V, const double *W, int nthreads);
double zfo_feq(int c, double rho, double u, double v, double w);

/* Pressure waveform of the BC codes 1001-1009, ZebraflowWT.c:6318-6421
 * (pressure in the reference's units, before /LBM_PRESSURE_SCALE). */
float zfo_bc_pressure(int code, float simtime);

/* ---- RBC membrane (zf_oracle_mesh.c) -------------------------------------- */
typedef struct zfo_memb_params {
  double len_scale;          /* force-formula length per coordinate unit (ref 1e-6) */
  double kT;                 /* BOLTZ*TEMPERATURE */
  double x0;                 /* e_ratio_eq, rest length / max length */
  double kbend, kd, ka, kv;  /* bending, local area, global area, volume */
  double area0, vol0;        /* global area and volume targets (RBC_SURF, MCV) */
  double x_max, x_floor;     /* e_ratio clamp 0.95, length floor 0.02435 */
  double cos_cut;            /* |cos theta| above which bending is off (0.99) */
  double cos_blend;          /* > 0: smoothstep fade over [cos_cut - w, cos_cut] (not in the reference) */
  double area_lo, area_hi;   /* triangle area clamp factors (0.1, 10) */
  double force_limit;        /* per-component guard of force terms (1e-10) */
  double vol_limit;          /* guard of the volume-gradient terms (1e-10) */
  double sliver_len, sliver_len_min, sliver_f1, sliver_f2; /* 0.05, 0.001, 1e-14, 1e-11 */
  double relax, relax2;      /* relaxfactor, relaxfactor2 (1, 1) */
} zfo_memb_params;

/* a7.  tri: 3*nt (1-based).  edge: 6*ecap rows n1 n2 n3 n4 t1 t2.
 * nbr_node / nbr_tri: nn rows of nlist ints (count, entries).  Returns the
 * edge count or -1.  Pinned by reference goldens (zebrafish mesh). */
int zfo_mesh_topology(int nn, int nt, const int32_t *tri, int32_t *edge, int ecap,
                      int32_t *nbr_node, int32_t *nbr_tri, int nlist);
/* a8.  xyz: 3*nn node-major rest coordinates. */
void zfo_mesh_equilibrium(int nn, int ne, const double *xyz, const int32_t *edge,
                          const zfo_memb_params *mp, double es, int passes, double *lmax,
                          double *plength, double *kp, double *theta0, double *sin0,
                          double *cos0);
/* a9-a12.  xyz: ncell*nn*3 (cell, node, axis).  Outputs per cell. */
void zfo_membrane_forces(int ncell, int nn, int nt, int ne, const int32_t *tri,
                         const int32_t *edge, const double *lmax, const double *plength,
                         const double *kp, const double *sin0, const double *cos0,
                         const double *tri_area_eq, const zfo_memb_params *mp, const double *xyz,
                         double *force, double *node_normal, double *tri_area, double *tri_normal,
                         double *tri_centroid, double *cell_area, double *cell_vol,
                         double *cell_centroid, int *guard_fail);

/* Nucleus network (ZebraflowWT.c:2003-2050 / :2203-2221 parameters,
 * :7388-7466 spring forces, :9519-9730 contact); edge 2*ne 1-based. */
void zfo_nucleus_equilibrium(int nnuc, int ne, const double *xyz, const int32_t *edge,
                             const zfo_memb_params *mp, double es, int passes, double *lmax,
                             double *plength, double *kp);
void zfo_nucleus_forces(int ncell, int nnuc, int ne, const int32_t *edge, const double *lmax,
                        const double *plength, const double *kp, const zfo_memb_params *mp,
                        const double *xyzn, double *force);
int zfo_nucleus_contact(int ncell, int nnuc, int nn, int nt, const int32_t *tri,
                        const int32_t *status, const double *xyzn, const double *xyz,
                        const double *tri_centroid, const double *tri_normal, double list_r,
                        double near_r, double contact, double k, int list_cap, double *vrep);

/* ---- immersed boundary (zf_oracle_ibm.c) ---------------------------------- */
typedef struct zfo_ibm_params {
  int im, jm, km;
  double x0, y0, z0, dx;   /* node (i,j,k) at (x0+(i-1)dx, ...); ref X0,Y0,Z0, DX */
  double vel_scale;        /* VEL_SCALE */
  double force_scale;      /* 1/UNITMASS/LBM_BFORCE_SCALE */
  double dt_advance;       /* DT*SOLVERSTAGGER */
  double vmax;             /* removal speed, 5e-3 */
} zfo_ibm_params;

/* a13.  U,V,W per listed node (list order); xyz, vel: nnode*3. */
void zfo_ibm_interpolate(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                         const double *U, const double *V, const double *W, int64_t nnode,
                         const double *xyz, double *vel);
/* zfo_ibm_interpolate plus the wall layer: window nodes that are not fluid
 * and have W_index > 0 add w * vel_scale * Uvoid (ZebraflowWT.c:8980-8996);
 * windex and uvoid (3*box SoA) may be NULL. */
void zfo_ibm_interpolate_wall(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                              const double *U, const double *V, const double *W,
                              const int32_t *windex, const double *uvoid, int64_t nnode,
                              const double *xyz, double *vel);
/* a16.  ffluid: nnode*3; force: 3*dommax, accumulated (+=). */
void zfo_ibm_spread(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                    int64_t dommax, int64_t nnode, const double *xyz, const double *ffluid,
                    double *force);
/* zfo_ibm_spread with the reference's OpenMP atomics (CPU-baseline timing). */
void zfo_ibm_spread_omp(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid,
                        int64_t dommax, int64_t nnode, const double *xyz, const double *ffluid,
                        double *force);
/* Wall shear stress and wall pressure at n wall-mesh points (xyz, unit
 * normals into the fluid nrm: n*3); D,U,V,W per listed node. */
void zfo_wss(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid, const double *D,
             const double *U, const double *V, const double *W, int64_t n, const double *xyz,
             const double *nrm, double mu_over_dx, double pressure_scale, double pgauge,
             double *wss, double *press, int32_t *valid);
/* a17.  xyz, vel: ncell*nn*3; status per cell. */
int zfo_advance(const zfo_ibm_params *p, int ncell, int nn, double *xyz, const double *vel,
                int32_t *status);

/* ---- cell list and repulsion (zf_oracle_cells.c) --------------------------- */
typedef struct zfo_rep_params {
  double bin_cutoff;       /* triangle kept in its nearest node's bin below this (1 um) */
  double k_cell, cutoff, search;          /* 8.333e-7, 1 um, 2 um */
  double k_wall, wall_d0, wall_dmax, wall_search;  /* 1e-4, 0.25, 1.0, 1.5 */
} zfo_rep_params;

/* a9 (cell list).  tri_centroid: ncell*nt*3; bin: ncell*nt, 1-based DOMID or 0. */
void zfo_cell_list(const zfo_ibm_params *p, const int32_t *b, const int32_t *domid, int ncell,
                   int nt, const int32_t *status, const double *tri_centroid, double bin_cutoff,
                   int32_t *bin);
/* a15.  Adds to vrep (ncell*nn*3). */
void zfo_cell_repulsion(const zfo_ibm_params *p, const zfo_rep_params *r, const int32_t *b,
                        const int32_t *domid, int ncell, int nn, int nt, const int32_t *tri,
                        const int32_t *status, const double *xyz, const double *node_normal,
                        const double *tri_centroid, const double *tri_normal, const int32_t *bin,
                        double *vrep);
void zfo_wall_node_tri(int nwn, int nwt, const int32_t *wall_tri, int32_t *wall_node_tri);
void zfo_windex(const zfo_ibm_params *p, int nwn, const double *wall_xyz, int32_t *windex);
/* zfo_windex with Wdistance, the FP32 distance to the chosen wall node
 * (5000 where none), box-sized. */
void zfo_windex_dist(const zfo_ibm_params *p, int nwn, const double *wall_xyz, int32_t *windex,
                     float *wdist);
/* Wall-layer velocity Uvoid over the box (ZebraflowWT.c:4788-4851), SoA
 * uvoid[a*box + q]: at W_index > 0, B <= 0, k_lo <= k <= KM - k_hi,
 * w(Wdistance) speed n(triwall), triwall the wall node's triangle nearest
 * to the selection point (select: 0 the node, 1 the point sp). */
void zfo_uvoid(const zfo_ibm_params *p, const int32_t *b, const int32_t *windex, const float *wdist,
               const double *wall_xyz, const int32_t *wall_tri, const double *wall_normal,
               const int32_t *wall_node_tri, double speed, double d0, double d1, int k_lo, int k_hi,
               int select, const double *sp, double *uvoid);
/* a14.  Adds to vrep; nearest: ncell*nn wall triangle (1-based) or 0. */
void zfo_wall_repulsion(const zfo_ibm_params *p, const zfo_rep_params *r, const int32_t *windex,
                        int nwn, const double *wall_xyz, int nwt, const int32_t *wall_tri,
                        const double *wall_normal, const int32_t *wall_node_tri, int ncell, int nn,
                        const int32_t *status, const double *xyz, double *vrep, int32_t *nearest);

/* ---- cells on a parent lattice (zf_oracle_wrap.c) -------------------------
 * The parent's periodic cell window along k: Z0_p = z0, ZM_p = zm =
 * z0 + (km - 1) dx; dir +1: cells leave at zm (parent 1), -1: at z0
 * (parents 2 and 3).  Node coordinates are the reference's (wrapped into
 * the window node by node); status 2 (inside) or 4 (crossing), -5000 free. */
typedef struct zfo_wrap {
  int dir;
  double z0, zm;
} zfo_wrap;

/* Interpolation (:7541-7732, :8101-8297); vel: ncell*nn*3 (set, not added);
 * windex and uvoid may be NULL. */
void zfo_wrap_interpolate(const zfo_ibm_params *p, const zfo_wrap *w, const int32_t *b,
                          const int32_t *domid, const double *U, const double *V, const double *W,
                          const int32_t *windex, const double *uvoid, int ncell, int nn,
                          const int32_t *status, const double *xyz, double *vel);
/* Wall repulsion of the interpolation loop (same lines); adds to vrep. */
void zfo_wrap_wall_repulsion(const zfo_ibm_params *p, const zfo_wrap *w, const zfo_rep_params *r,
                             const int32_t *windex, const double *wall_xyz, const int32_t *wall_tri,
                             const double *wall_normal, const int32_t *wall_node_tri, int ncell,
                             int nn, const int32_t *status, const double *xyz, double *vrep,
                             int32_t *nearest);
/* Cell list (:5898-5950); tri_centroid as the membrane stage computes it. */
void zfo_wrap_cell_list(const zfo_ibm_params *p, const zfo_wrap *w, const int32_t *b,
                        const int32_t *domid, int ncell, int nt, const int32_t *status,
                        const double *tri_centroid, int32_t *bin);
/* Cell-cell repulsion (:10204-10383); adds to vrep. */
void zfo_wrap_cell_repulsion(const zfo_ibm_params *p, const zfo_wrap *w, const zfo_rep_params *r,
                             const int32_t *b, const int32_t *domid, int ncell, int nn, int nt,
                             const int32_t *tri, const int32_t *status, const double *xyz,
                             const double *node_normal, const double *tri_centroid,
                             const double *tri_normal, const int32_t *bin, double *vrep);
/* Spreading (:11184-11310, :11547-11669); force 3*dommax, added. */
void zfo_wrap_spread(const zfo_ibm_params *p, const zfo_wrap *w, const int32_t *b,
                     const int32_t *domid, int64_t dommax, int ncell, int nn,
                     const int32_t *status, const double *xyz, const double *ffluid, double *force);
/* Position update with the wrap and clone flags (:12196-12340). */
void zfo_wrap_advance(const zfo_wrap *w, double dt, int ncell, int nn, int nnuc, double *xyz,
                      const double *vel, double *xyzn, const double *veln, int32_t *status,
                      int32_t *regress, int32_t *cloneflag);

/* Wall-clock helper for the CPU baseline. */
double zfo_wtime(void);
int zfo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
