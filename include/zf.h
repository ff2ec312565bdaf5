/*
 * zf.h -- C ABI of libzfgpu.so, the MI355X engine for the Zebraflow time step.
 *
 * The reference (ZebraflowWT.c) has no operator API: its time step is a
 * sequence of inline OpenMP regions in main() over static, 1-based,
 * structure-of-arrays globals (ZebraflowWT.c:5646-12412).  The entry points
 * below replace those regions one for one; each comment names the region.
 *
 * Conventions
 *  - Every call returns ZF_OK (0) or a negative ZF_ERR_* code; the message of
 *    the last failure is returned by zf_last_error().  Nothing aborts.
 *  - All host arrays are caller-owned; the library never keeps a host pointer.
 *  - The box array b_index is dense, element (i,j,k) (1-based lattice
 *    indices) at ((i-1)*jm + (j-1))*km + (k-1): the reference's
 *    B_index[i][j][k] without its padding row (ZebraflowWT.c:180).
 *  - List outputs are 1-based, as the reference stores them.
 *  - Per-node host arrays are indexed by p = DOMID-1 (the reference's list
 *    order); populations are SoA f[c*dommax + p], c = 0..18 for the
 *    reference's directions 1..19 (ZebraflowWT.c:2471-2504).
 *  - One host thread per context.  All device work is ordered on the
 *    context's single HIP stream; calls that return host data synchronise it.
 */
#ifndef ZF_H
#define ZF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ZF_OK = 0,
  ZF_ERR_ARG = -1,      /* bad argument or size */
  ZF_ERR_HIP = -2,      /* a HIP runtime call failed */
  ZF_ERR_STATE = -3,    /* call out of order (e.g. step before lattice) */
  ZF_ERR_GEOMETRY = -4, /* lattice rejected (interior node next to B<0) */
  ZF_ERR_NOMEM = -5     /* device allocation failed */
};

/* Population storage forms. */
enum {
  ZF_STORE_F64 = 0,    /* FP64 populations f_i (the plain comparator form) */
  ZF_STORE_F32DEV = 1, /* FP32 deviations f_i - w_i, FP32 arithmetic */
  ZF_STORE_F16DEV = 2  /* FP16 (f_i - w_i)/w_i, FP32 arithmetic */
};

/* Wall treatment for B==0 nodes. */
enum {
  ZF_WALL_EQUILIBRIUM = 0, /* reference: wall node = feq(D[nbr], 0), ZebraflowWT.c:6309-6315 */
  ZF_WALL_HALFWAY_BB = 1   /* halfway bounce-back on links from B==0 nodes */
};

/* Kernel options (zf_config.flags); results do not depend on them except
 * ZF_FLAG_F64_ARITH, which computes the FP32/FP16 storage forms in FP64,
 * and ZF_FLAG_F32_CELLS.
 * By default the stream kernel sets its cache policy from the population
 * footprint: every direction cached when both population buffers fit the
 * 256 MB last-level cache, otherwise as many directions as take about
 * 120 MB of one buffer, the rest loaded and stored non-temporally. */
#define ZF_FLAG_LINEAR_BLOCKS 1u    /* plain block order instead of the XCD-aware one */
#define ZF_FLAG_NT_POPULATIONS 2u   /* every population load and store non-temporal */
#define ZF_FLAG_F64_ARITH 4u        /* FP64 arithmetic on FP32/FP16 storage */
/* FP32 arithmetic in the membrane's node forces of the FP32/FP16 storage
 * forms (positions relative to each node; triangle geometry, the per-cell
 * area and volume reductions and the cell-cell and wall repulsion stay
 * FP64).  Changes results at the FP32 level. */
#define ZF_FLAG_F32_CELLS 8u
/* Keep directions 0..k-1 cached (k = 19: all; an odd k < 19 counts as k-1),
 * the others non-temporal, instead of the footprint rule. */
#define ZF_FLAG_CACHED_DIRS(k) ((uint32_t)((k) + 1) << 8)
/* ZF_FLAG_PAIRED_STEPS: zf_run / zf_lbm_run advance the lattice two steps
 * per launch (the state between them kept in a ring of slots in the caches
 * instead of the population buffers; the same results bit for bit) when no
 * parent lattice is linked.  Off by default: measured slower than one
 * launch per step.  ZF_FLAG_SINGLE_STEPS: always one launch per step. */
#define ZF_FLAG_SINGLE_STEPS 16u
#define ZF_FLAG_PAIRED_STEPS 32u

/* Boundary codes in b_index beyond the reference's (0 wall, 1..9 fluid with
 * tau = 0.5 + visc_scale*B, 1001..1009 pressure waveforms). */
#define ZF_B_VELOCITY_INLET 2001 /* feq(D[nbr], u_in(t)*inlet_dir) */
#define ZF_B_PRESSURE_OUTLET 3001 /* feq(outlet_rho, U[nbr]) */

typedef struct zf_config {
  int32_t im, jm, km;     /* box (ZebraflowWT.c:81-83 IM, JM, KM) */
  int32_t periodic[3];    /* wrap the pull along i, j, k */
  int32_t storage;        /* ZF_STORE_* */
  int32_t wall_mode;      /* ZF_WALL_* */
  int32_t device;         /* HIP device ordinal */
  int32_t flags;          /* ZF_FLAG_* kernel options */
  double visc_scale;      /* tau = 0.5 + visc_scale*B; reference MU_SCALE*MU = 0.72 */
  double force_dt;        /* DT in the Guo force terms; reference 0.5 */
  double body_force[3];   /* uniform force added to the per-node field */
  double inlet_u0, inlet_amp, inlet_period; /* u_in = u0*(1+amp*sin(2 pi iter/period)) */
  double inlet_dir[3];
  double outlet_rho;
  double t_scale;         /* seconds per step for codes 1001-1009 (T_SCALE) */
  double pressure_scale;  /* LBM_PRESSURE_SCALE */
} zf_config;

/* Reference units and constants (ZebraflowWT.c:34-74): visc_scale 0.72,
 * force_dt 0.5, equilibrium walls, T_SCALE 5e-7 s. */
void zf_config_reference(zf_config *cfg, int32_t im, int32_t jm, int32_t km);
/* Lattice units: tau for B==1, force_dt 1, halfway bounce-back. */
void zf_config_lattice(zf_config *cfg, int32_t im, int32_t jm, int32_t km, double tau);

typedef struct zf_ctx zf_ctx;

int zf_create(const zf_config *cfg, zf_ctx **out);
int zf_destroy(zf_ctx *ctx);
const char *zf_last_error(const zf_ctx *ctx);
const char *zf_version(void);

/* ---- index lists ---------------------------------------------------------
 * Builds, on the device, the fluid-node list (LBM_DOM_SETUP,
 * ZebraflowWT.c:1481-1502), the BC list and BC-neighbour list
 * (LBM_BC_SETUP, ZebraflowWT.c:1506-1589) and the pull-neighbour list
 * (ZebraflowWT.c:6113-6125), then the kernel's run-length form of it.
 * Counts are returned through the (nullable) out-params. */
int zf_upload_lattice(zf_ctx *ctx, const int32_t *b_index, int32_t *dommax,
                      int32_t *bcmax, int32_t *n_interior);
/* DOMLIST dommax*3, BCLIST bcmax*4, BCNLIST bcmax*3 (any may be NULL). */
int zf_download_lists(zf_ctx *ctx, int32_t *domlist, int32_t *bclist, int32_t *bcnlist);
/* DOMID over the box (1-based l or -1). */
int zf_download_domid(zf_ctx *ctx, int32_t *domid);
/* nbr[c*dommax + p] = DOMID of (i,j,k) - e_c, c = 0..17 (1-based, -1 none). */
int zf_download_neighbours(zf_ctx *ctx, int32_t *nbr);

typedef struct zf_lattice_info {
  int32_t dommax, bcmax, n_interior, n_chords, n_bc_active, n_bb_links;
  int32_t n_macro_nodes, stride;
  int64_t device_bytes;
} zf_lattice_info;
/* The configuration the context was created with (storage form etc.). */
int zf_get_config(const zf_ctx *ctx, zf_config *cfg);
int zf_get_lattice_info(zf_ctx *ctx, zf_lattice_info *info);

/* ---- LBM state -------------------------------------------------------------
 * Populations: f = feq(rho, u) everywhere, as the reference initialises them
 * (ZebraflowWT.c:5253-5262). */
int zf_init_equilibrium(zf_ctx *ctx, double rho, const double u[3]);
int zf_upload_populations(zf_ctx *ctx, const double *f);
int zf_download_populations(zf_ctx *ctx, double *f);
/* The populations as stored, without conversion, in the host layout
 * [c*dommax + p]: double f (ZF_STORE_F64), float f - w (ZF_STORE_F32DEV) or
 * half (f - w)/w (ZF_STORE_F16DEV); bytes must be 19*dommax times 8, 4 or
 * 2.  The double layout above rounds the FP32 deviations of f near w (below
 * about 5e-10); these keep every bit, for restart files that continue
 * bit-identically.  No counterpart in the reference, which restarts from
 * its macroscopic output (ZebraflowWT.c:5253-5262). */
int zf_download_populations_raw(zf_ctx *ctx, void *f, int64_t bytes);
int zf_upload_populations_raw(zf_ctx *ctx, const void *f, int64_t bytes);
/* Per-node force field BODYFORCE (ZebraflowWT.c:183), 3*dommax SoA; NULL
 * clears it. */
int zf_upload_force(zf_ctx *ctx, const double *force);
int zf_download_force(zf_ctx *ctx, double *force);

/* ---- the time step -------------------------------------------------------
 * One LBM iteration: stream + macro + BGK-Guo collide (ZebraflowWT.c:6101-6145),
 * boundary conditions (:6292-6433) and the copy (:6601-6626, a buffer swap
 * here).  rbc_step != 0 clears the force field afterwards, as :6622-6624. */
int zf_lbm_step(zf_ctx *ctx, int32_t iter, int32_t rbc_step);
/* nsteps iterations iter0+1 .. iter0+nsteps with rbc_step = 0 (two per
 * launch where the lattice takes them, see ZF_FLAG_PAIRED_STEPS). */
int zf_lbm_run(zf_ctx *ctx, int32_t iter0, int32_t nsteps);
/* The two-step launch on this lattice (its tables are built by this call
 * if not yet): info[0] = 1 if zf_run / zf_lbm_run take it with the current
 * flags, info[1] = 1 if the lattice admits it at all, info[2] = XCDs,
 * info[3] = ghost blocks (the farthest pull, in 64-node blocks),
 * info[4] = ring slots per XCD (64-node blocks), info[5] = workgroups of
 * 256 lanes per launch.  No counterpart in the reference, whose loop has no launches. */
int zf_pair_info(zf_ctx *ctx, int32_t info[6]);
/* D, U, V, W of every listed node from the current populations without the
 * force shift (ZebraflowWT.c:6606-6614); each array dommax (any may be NULL). */
int zf_download_macro(zf_ctx *ctx, double *d, double *u, double *v, double *w);
/* Mass sum_p sum_c f and momentum over the interior nodes in FP64, reduced
 * on the device; out[0..3] = mass, jx, jy, jz. */
int zf_conservation_sums(zf_ctx *ctx, double out[4]);

/* Per-kernel timing pass: runs nsteps iterations iter0+1.. like zf_lbm_run,
 * bracketing every launch with HIP events on the context stream.
 * out_ms[0] = total ms in the stream-collide kernel, out_ms[1] = 0 (the
 * boundary update runs inside that kernel), out_ms[2] = launches timed. */
int zf_lbm_time_kernels(zf_ctx *ctx, int32_t iter0, int32_t nsteps, double out_ms[3]);

/* ---- parent (feeder) lattices ----------------------------------------------
 * The reference's three parent lattices (ZebraflowWT.c:91-112; lists
 * LBM_PARENTDOMAIN_SETUP :515-1013 and LBM_PARENTBC_SETUP :1015-1350; step
 * :6148-6286; boundary pass :6436-6594) are contexts of their own.
 * zf_set_parent, before zf_upload_lattice, makes ctx parent p (1..3; 0
 * undoes it): its BC list takes B == 0 or B >= 10, wall and upstream
 * (20+p) nodes take the parent thresholds for their neighbour, and the
 * copy code 10+p and the upstream velocity source name the main-domain node
 * (i + offset[0], j + offset[1], k_over + offset[2]) over the node's column
 * at the overlap plane k_over (the reference's (int)((X - X0)/DX + 1.01)
 * with its lattice origins; parent 1: k_over = KM_1). */
int zf_set_parent(zf_ctx *ctx, int32_t p, int32_t k_over, const int32_t offset[3]);
/* vBCNLIST of a parent lattice, bcmax*3 (zero except for upstream nodes). */
int zf_download_parent_lists(zf_ctx *ctx, int32_t *vbcn);
/* Couples parent lattice `parent` to `main` (both uploaded, same device):
 * every zf_lbm_step of main then exports the moments of the nodes the
 * parent reads, and every zf_lbm_step of the parent ends with its boundary
 * pass (:6436-6594) for codes 10+p and 20+p; walls stay with its stream
 * kernel.  Step main first, then its parents, in each iteration; while
 * linked, a parent issues its work on the main context's stream.  Fails with ZF_ERR_GEOMETRY if a coupled
 * node maps outside the main box or onto an unlisted node.  Destroying or
 * re-uploading either context unlinks them.  The adjoint entry points do
 * not cover the coupling: linked contexts refuse zf_adjoint_setup and a
 * context with an adjoint set up refuses the link.  A main lattice and its
 * linked parents are one unit: drive them from one host thread, main first
 * in each iteration; synchronising either context waits for both. */
int zf_parent_link(zf_ctx *parent, zf_ctx *main);

/* ---- RBC membranes ---------------------------------------------------------
 * Constants of the membrane model (ZebraflowWT.c:34-36, :118-134, :5654).
 * len_scale converts a coordinate difference to the force formulas' length
 * (1e-6 m per um in the reference; 1 in lattice units). */
typedef struct zf_membrane_config {
  double len_scale, kT, x0, es;     /* es: shear modulus of the WLC-POW fit */
  double kbend, kd, ka, kv;
  double area0, vol0;               /* global targets (RBC_SURF, MCV) */
  double x_max, x_floor, cos_cut, area_lo, area_hi;
  double cos_blend;                 /* 0: bending off above cos_cut (reference); w > 0:
                                       faded out by smoothstep over [cos_cut - w, cos_cut] */
  double force_limit, vol_limit;    /* per-component guards (1e-10) */
  double sliver_len, sliver_len_min, sliver_f1, sliver_f2;
  double relax, relax2;
  int32_t passes;                   /* fixed-point passes (51, :2338-2346) */
  int32_t max_cells;
} zf_membrane_config;

/* Reference SI constants (the zebrafish cell).  es = 3e-6 N/m. */
void zf_membrane_config_reference(zf_membrane_config *mc);

/* Builds the mesh topology (CGSM_ZFISHRBC_DATA_GEN, ZebraflowWT.c:1632-1901),
 * the equilibrium parameters from the rest shape (CGSM_FILAMENT_PARAMETERS /
 * CGSM_EQUILIBRIUM_METRICS, :1930-2223, :2338-2346) and rest triangle areas
 * (:2387-2460) on the host, and uploads them.  tri: 3*nt 1-based node ids
 * wound outward; rest_xyz: 3*nn. */
int zf_membrane_setup(zf_ctx *ctx, const zf_membrane_config *mc, int32_t nn, int32_t nt,
                      const int32_t *tri, const double *rest_xyz);
/* edge: 6*ne (n1 n2 n3 n4 t1 t2, 1-based); nbr_node / nbr_tri: nn rows of
 * 10 ints (count, up to 9 entries) as NEIGHNODE_LIST / NEIGHTRI_LIST. */
int zf_download_mesh_topology(zf_ctx *ctx, int32_t *ne, int32_t *edge, int32_t *nbr_node,
                              int32_t *nbr_tri);
/* Per edge: LMAX, Plength, kp, Theta0, sin(Theta0), cos(Theta0). */
int zf_download_mesh_equilibrium(zf_ctx *ctx, double *lmax, double *plength, double *kp,
                                 double *theta0, double *sin0, double *cos0);
/* Cell node coordinates, ncell*nn*3 (cell, node, axis); resets velocities. */
int zf_upload_cells(zf_ctx *ctx, int32_t ncell, const double *xyz);
int zf_download_cells(zf_ctx *ctx, double *xyz, double *vel, int32_t *status);

/* Membrane forces of every live cell: triangle geometry and node normals
 * (CGSM_AREA, :5757-5851, :6089-6093), anti-sliver terms (:6042-6086), cell
 * area/centroid (:6717-6796), spring, bending and area terms (:6800-7384) and
 * the volume constraint (:7512-7535).  Cell area, volume and centroid are
 * FP64 reductions on the device. */
int zf_membrane_forces(zf_ctx *ctx);
/* force, node_normal: ncell*nn*3; cell_area, cell_vol: ncell; centroid:
 * ncell*3; guard_fail: ncell (terms dropped by the guards).  Any may be NULL. */
int zf_download_membrane(zf_ctx *ctx, double *force, double *node_normal, double *cell_area,
                         double *cell_vol, double *cell_centroid, int32_t *guard_fail);

/* ---- nucleus network (nucleated RBCs) --------------------------------------
 * Each cell of the context may carry a nucleus: a network of nnuc nodes and
 * ne WLC-POW filaments (the reference's 121-node, 544-edge zebrafish
 * nucleus, NucleusNodes.dat / NucleusEdges.dat read at ZebraflowWT.c:
 * 1903-1922).  Its nodes are immersed-boundary points of their cell:
 * zf_step interpolates their velocities, spreads their forces and
 * advances them with the membrane's nodes (:8693-8996, :10875-10990,
 * :11975-11989; the speed check stays with the membrane nodes).  Per RBC
 * iteration:
 *   - spring forces F_Total_N (:7388-7466): the membrane's WLC + POW
 *     spring on every edge, with equilibrium parameters from the rest
 *     network (the 51-pass fixed point of :2003-2050 with es_factor * es,
 *     lmax of :2203-2221);
 *   - contact (:9519-9730): every list_interval iterations (and when cells
 *     or nuclei are uploaded) each nucleus node lists the triangles of its
 *     own membrane whose centroid lies within list_radius; a listed
 *     triangle whose centroid is within near_radius is probed at 19 points
 *     (centroid, sub-triangle centroids and midpoints), and every point
 *     within `contact` of the nucleus node pushes the triangle's three
 *     nodes along its normal: v_rep += k_contact (contact - d)/contact n/3.
 *     The contact runs after the cell-cell and wall repulsion.
 * The adjoint does not cover the nucleus: a context with one refuses
 * zf_adjoint_setup with cells. */
typedef struct zf_nucleus_config {
  double es_factor;     /* 0.1: network shear modulus / membrane's (:2040) */
  double list_radius;   /* 0.5: triangle list radius (:9582) */
  double near_radius;   /* 1.0: listed triangles probed within (:9629) */
  double contact;       /* 0.2: contact range of the probe points (:9705) */
  double k_contact;     /* 4.1666667e-07: contact speed scale (:9707) */
  int32_t list_interval; /* 1000: lists rebuilt when iter % list_interval == 0 (:9527) */
  int32_t list_cap;      /* list entries per nucleus node (the reference's NucMembMAX 1004) */
} zf_nucleus_config;

void zf_nucleus_config_reference(zf_nucleus_config *nc);
/* After zf_membrane_setup (and before or after zf_ibm_setup): edge 2*ne
 * 1-based node ids, rest_xyz 3*nnuc in the frame of the membrane rest
 * shape. */
int zf_nucleus_setup(zf_ctx *ctx, const zf_nucleus_config *nc, int32_t nnuc, int32_t ne,
                     const int32_t *edge, const double *rest_xyz);
/* Per edge: LMAX_N, Plength_N, kp_N. */
int zf_download_nucleus_equilibrium(zf_ctx *ctx, double *lmax, double *plength, double *kp);
/* Nucleus node coordinates of every uploaded cell, ncell*nnuc*3 (cell,
 * node, axis), after zf_upload_cells; resets their velocities. */
int zf_upload_nuclei(zf_ctx *ctx, const double *xyz);
/* xyz, vel, force (F_Total_N): ncell*nnuc*3 each (any may be NULL). */
int zf_download_nuclei(zf_ctx *ctx, double *xyz, double *vel, double *force);
/* The two nucleus stages alone (zf_step runs them in its order):
 * spring forces of the current positions, and the contact of iteration
 * iter added to the repulsion velocity. *list_overflow (nullable): list
 * entries dropped for want of room since the set-up. */
int zf_nucleus_forces(zf_ctx *ctx);
int zf_nucleus_contact(zf_ctx *ctx, int32_t iter, int32_t *list_overflow);
/* *nnuc and *list_cap of the set-up nucleus, 0 and 0 without one. */
int zf_get_nucleus_info(zf_ctx *ctx, int32_t *nnuc, int32_t *list_cap);
/* The contact lists as they stand (NUCLNEIGH between rebuilds,
 * ZebraflowWT.c:9527-9590), for exact restart files: the reference
 * rebuilds them at its first RBC iteration after a restart (NucMembUpdate
 * starts at 0), so its restarted run leaves the continuous one until the
 * next multiple of 1000.  list ncell*nnuc*list_cap local triangle ids
 * (0-based, the first count[q] of row q used), count ncell*nnuc, *valid 0
 * when the next zf_nucleus_contact rebuilds them whatever its iteration.
 * The upload checks every entry. */
int zf_download_nucleus_lists(zf_ctx *ctx, int32_t *list, int32_t *count, int32_t *valid);
int zf_upload_nucleus_lists(zf_ctx *ctx, const int32_t *list, const int32_t *count, int32_t valid);

/* ---- immersed-boundary coupling ------------------------------------------- */
typedef struct zf_ibm_config {
  double x0, y0, z0, dx;  /* node (i,j,k) at (x0+(i-1)dx, ...) (X0, Y0, Z0, DX) */
  double vel_scale;       /* VEL_SCALE */
  double force_scale;     /* 1/UNITMASS/LBM_BFORCE_SCALE */
  double gamma_p;         /* collision dissipation: F_fluid = gamma_p v_rep + F */
  double dt_advance;      /* DT*SOLVERSTAGGER */
  double vmax;            /* a cell is removed above this node speed (5e-3) */
  int32_t rbc_interval;   /* SOLVERSTAGGER: membrane/IBM every n-th iteration */
  int32_t flags;          /* ZF_IBM_* */
} zf_ibm_config;

/* Kernel support of interpolation and spreading: the reference's is the
 * 4x4x4 cosine kernel cut to the sphere |d| <= 2 DX, where the weight is not
 * zero, so the coupling jumps when a lattice node crosses that sphere.
 * ZF_IBM_CUBE_SUPPORT keeps the whole cube |d_a| < 2 DX instead: continuous
 * (C1) in the positions, for differentiating the model at scale. */
#define ZF_IBM_CUBE_SUPPORT 1

/* Reference values (ZebraflowWT.c:40-74, :95-97, :124, :151). */
void zf_ibm_config_reference(zf_ibm_config *ic);
int zf_ibm_setup(zf_ctx *ctx, const zf_ibm_config *ic);

/* Node velocities from the fluid: the 4-point cosine kernel with the
 * reference's spherical cut-off over fluid nodes 0 < B < 4998, using D,U of
 * the current populations (ZebraflowWT.c:8702-8996, macro of :6606-6614).
 * Overwrites the node velocities. */
int zf_ibm_interpolate(zf_ctx *ctx);
/* Node velocities += v_rep, then F_fluid = gamma_p v_rep + F_Total spread
 * onto every B >= 0 node (ZebraflowWT.c:10746-10867).  The force field is
 * overwritten: the reference zeroes it on the same iteration (:6622-6624).
 * Deterministic: nodes are binned into 4^3 lattice tiles and each tile
 * gathers from its 27 neighbours in a fixed order; no atomics. */
int zf_ibm_spread(zf_ctx *ctx);
/* x += v*dt_advance; a cell with a node faster than vmax (or NaN) is
 * removed (ZebraflowWT.c:11975-11989).  *n_removed may be NULL. */
int zf_advance_cells(zf_ctx *ctx, int32_t *n_removed);
/* Overrides for tests and external coupling: F_Total and v_rep, ncell*nn*3. */
int zf_upload_membrane_force(zf_ctx *ctx, const double *force);
int zf_upload_repulsion(zf_ctx *ctx, const double *vrep);
int zf_download_repulsion(zf_ctx *ctx, double *vrep);

/* ---- cell list and repulsion ---------------------------------------------- */
typedef struct zf_repulsion_config {
  double bin_cutoff;   /* a triangle enters its nearest node's bin below this (1 um) */
  double k_cell;       /* cell-cell repulsion speed scale (8.333e-7) */
  double cutoff;       /* cell-cell range (1 um) */
  double search;       /* triangle centroid search radius (2 um) */
  double k_wall;       /* wall repulsion speed scale (1e-4) */
  double wall_d0, wall_dmax, wall_search;  /* 0.25, 1.0, 1.5 */
} zf_repulsion_config;

void zf_repulsion_config_reference(zf_repulsion_config *rc);
int zf_repulsion_setup(zf_ctx *ctx, const zf_repulsion_config *rc);
/* Lumen wall mesh (the reference's WallMesh*.dat, ZebraflowWT.c:3715-4643):
 * node coordinates (3*nwn), triangles (3*nwt, 1-based), unit normals into
 * the fluid (3*nwt).  Builds the node-to-triangle lists (:3900-3935) and the
 * W_index map of each lattice node to its nearest wall node within the 9^3
 * window (:3715-3826) on the device. */
int zf_wall_setup(zf_ctx *ctx, int32_t nwn, const double *wall_xyz, int32_t nwt,
                  const int32_t *wall_tri, const double *wall_normal);
/* W_index over the box (1-based wall node or 0). */
int zf_download_windex(zf_ctx *ctx, int32_t *windex);

/* Wall-layer velocity Uvoid (built at ZebraflowWT.c:4788-4851, read by the
 * interpolation at :8980-8996): every box node with W_index > 0, B <= 0 and
 * k_lo <= k <= KM - k_hi carries Uvoid = w * speed * n, where w ramps
 * linearly from 0 at Wdistance (the FP32 distance to its wall node) d0 to 1
 * at d1, and n is the unit normal of the wall triangle, among those of its
 * wall node in order, with the vertex or centroid nearest to the selection
 * point (FP32 distances, strict <: vertices then centroid of each triangle).
 * zf_ibm_interpolate then adds vel_scale * w_q * Uvoid_q over the window
 * nodes that are not fluid (B <= 0 or B >= 4998) and have W_index > 0, with
 * the same kernel as the fluid nodes (WallStencilWidth = StencilWidth = 2).
 * The reference's selection point is not the node: it is whatever its
 * variables xp, yp, zp held when the loop ran, the coordinates of the last
 * row of the last start file read (PCV2parentdomainStart.csv, :3579);
 * select = 1 with select_point reproduces that, select = 0 uses the node's
 * own position.  Call after zf_wall_setup and zf_ibm_setup; NULL turns the
 * term off.  The node transpose of zf_ibm_vjp and the adjoint include it. */
typedef struct zf_uvoid_config {
  double speed;            /* 1e-3 / VEL_SCALE (ZebraflowWT.c:4845), units of U */
  double d0, d1;           /* 0.7, 3.0: ramp of Wdistance (coordinate units) */
  int32_t k_lo, k_hi;      /* nodes k_lo <= k <= KM - k_hi: 3, 2 */
  int32_t select;          /* 0: the node's own position, 1: select_point */
  int32_t reserved;
  double select_point[3];  /* the reference's: 78.5, 214.5, 392.5 */
} zf_uvoid_config;
void zf_uvoid_config_reference(zf_uvoid_config *uc);
int zf_uvoid_setup(zf_ctx *ctx, const zf_uvoid_config *uc);
/* Uvoid over the box, SoA uvw[a*box + q] (zeros where none). */
int zf_download_uvoid(zf_ctx *ctx, double *uvw);
/* Triangle cell list (EulNeigh, ZebraflowWT.c:5851-5900) from the triangle
 * centroids of the last membrane evaluation: bin[c*nt + t] = DOMID of the
 * triangle's bin (1-based) or 0.  Built as (bin, triangle) sorted keys, so
 * each bin lists its triangles in (cell, triangle) order. */
int zf_cell_list(zf_ctx *ctx);
int zf_download_cell_list(zf_ctx *ctx, int32_t *bin);
/* v_rep = cell-wall (ZebraflowWT.c:8998-9142, when a wall is set up) +
 * cell-cell (:9913-10052) repulsion velocities of every live node. */
int zf_repulsion(zf_ctx *ctx);
/* The wall triangle chosen for each node (NODE_WALLNEIGH), ncell*nn,
 * 1-based or 0. */
int zf_download_wall_neighbour(zf_ctx *ctx, int32_t *nearest_wall);

/* One iteration in the reference's order (ZebraflowWT.c:5646-12412): on an
 * iteration divisible by rbc_interval, the LBM step (clearing the force
 * field), membrane forces, interpolation, repulsion, spreading and the
 * position update; otherwise the LBM step alone. */
int zf_step(zf_ctx *ctx, int32_t iter);
/* zf_step for iterations iter0+1 .. iter0+nsteps (the time loop's body,
 * ZebraflowWT.c:5646-12412, without output). */
int zf_run(zf_ctx *ctx, int32_t iter0, int32_t nsteps);

/* ---- cells on the parent lattices ------------------------------------------
 * The reference's parent lattices carry cells of their own that feed the
 * main lattice (the region of interest, ROI).  A parent's cells live in a
 * periodic window along k, [Z0_p, ZM_p] with ZM_p = Z0_p + (KM_p - 1) DX:
 *   - every node is kept inside the window, wrapping by L = ZM_p - Z0_p
 *     when it leaves (:12196-12340), and the parent's stages address the
 *     lattice with k modulo KM_p - 1 (interpolation and wall repulsion
 *     :7541-8688, cell list :5898-6000, cell-cell repulsion :10199-10733,
 *     spreading :11179-11917); a cell across the wrap (status > 3, nodes
 *     spanning more than split_span) is unwrapped for the membrane and
 *     nucleus stages (:5769-5797, :6738-6756, :7410-7420);
 *   - status 2 is a cell inside the window; when one of its nodes leaves
 *     through the exit (ZM_p for direction +1, Z0_p for -1) and its
 *     REGRESSFLAG is 0, the cell is cloned into the first free slot (status
 *     -5000) of the main lattice, unwrapped so that it straddles the exit
 *     plane, with status -(id_base + slot + 1); the parent cell becomes
 *     status 4 with REGRESSFLAG 1 (:12345-12403), and returns to status 2
 *     inside the window's first three quarters (:12323-12336).  REGRESSFLAG
 *     returns to 0 when the cell's centroid is in the window's middle half
 *     and it is not across the wrap (:6762-6778);
 *   - in the main lattice a clone is a live cell whose nodes within
 *     clone_band*DX past the exit plane, or behind it, move with the same
 *     node of the parent cell and do not spread (:9295-9316, :11002-11015,
 *     :11927-11958); once all its membrane nodes lie in (roi_lo, roi_hi) it
 *     is released into status 1 (:11960-11962).
 * A main lattice's statuses: > 0 live, -4998 < s < 0 a clone, -5000 free.
 * Give the main lattice free slots for the clones (zf_upload_cells with
 * room, zf_upload_cell_status with -5000).  Not covered: the main lattice's
 * exits back into the parents (statuses 3001/3002, :12017-12160). */
typedef struct zf_cell_wrap_config {
  int32_t direction;    /* +1: cells leave at ZM_p (parent 1); -1: at Z0_p (parents 2, 3) */
  int32_t id_base;      /* the reference's number of the parent's first cell slot, minus 1 */
  double split_span;    /* 30: a cell with status > 3 spanning more lies across the wrap */
  double clone_band;    /* 5: band past the exit plane, in DX, where clone nodes follow */
  double roi_lo, roi_hi; /* release planes of the clones in the main lattice (ZM_1, Z0_2) */
} zf_cell_wrap_config;

/* The reference's values for its parent p = 1..3 (ZebraflowWT.c:91-112,
 * slots 1-188, 189-329, 330-344). */
void zf_cell_wrap_config_reference(zf_cell_wrap_config *wc, int32_t p);
/* Makes parent lattice `parent` (linked to `main` with zf_parent_link, both
 * with the same membrane mesh, immersed-boundary set-up and RBC interval,
 * and both with or without a nucleus) carry cells in its periodic window,
 * cloning into `main`.  Uploaded parent cells start with status 2 and
 * REGRESSFLAG 0 (zf_upload_cell_status may set 4). */
int zf_cell_wrap_setup(zf_ctx *parent, zf_ctx *main, const zf_cell_wrap_config *wc);
/* One iteration of a main lattice and its linked parents in the
 * reference's order: the LBM steps of the main lattice and of each parent;
 * then, on an RBC iteration, the cell stages of every parent and of the
 * main lattice (up to spreading), the main lattice's position update
 * (clones with their parent cells' velocities), the parents' wrapped
 * updates and the cloning, parent by parent. */
int zf_step_linked(zf_ctx *main, int32_t iter);
/* REGRESSFLAG and CLONEFLAG (of the last update) of a parent's cells,
 * ncell each (either may be NULL); the upload sets REGRESSFLAG. */
int zf_download_cell_flags(zf_ctx *ctx, int32_t *regress, int32_t *cloneflag);
int zf_upload_cell_flags(zf_ctx *ctx, const int32_t *regress);

/* ---- wall shear stress and wall pressure ----------------------------------
 * At n wall-mesh points (xyz: n*3, unit normals into the fluid: n*3), from
 * the current populations (PostProWSSandPressure.c:267-369):
 *   pressure[m] = (D - 1) * pressure_scale - pgauge at the nearest lattice
 *                 node with B >= 0 closer than 1.42 DX (valid[m] = 1), else
 *                 0 with valid[m] = 0;
 *   wss[3m..]   = -mu_over_dx * (u - (u.n) n), u the unnormalised cosine-
 *                 kernel velocity at x + DX n over nodes 0.11 < B < 4997.
 * Positions use the IBM grid (X0, DX) when set up, lattice units otherwise.
 * The reference's mu/DX is 0.0012 / (DX * 1e-6). */
int zf_wss(zf_ctx *ctx, int64_t n, const double *xyz, const double *normal, double mu_over_dx,
           double pgauge, double *wss, double *pressure, int32_t *valid);

/* Allocation-free zf_wss: same arguments, kernel and results, with the
 * points staged in a fixed device scratch of ZF_WSS_NOALLOC_MAX_POINTS
 * points (84 bytes each, about 16 MB per device, held until the process
 * ends) instead of buffers allocated per call; larger n fails with
 * ZF_ERR_ARG.  It calls no hipMalloc/hipFree.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (output
 * loops under HIP graph capture, serving loops that forbid allocation).
 * Calls on one device share the scratch: overlapping calls on one device
 * (other host threads, other contexts) race. */
#define ZF_WSS_NOALLOC_MAX_POINTS 196608
int zf_wss_noalloc(zf_ctx *ctx, int64_t n, const double *xyz, const double *normal,
                   double mu_over_dx, double pgauge, double *wss, double *pressure,
                   int32_t *valid);

/* ---- adjoint --------------------------------------------------------------
 * Reverse-mode gradients of a wall-shear-stress objective through the LBM
 * time step (the stream-collide of ZebraflowWT.c:6101-6145 with the
 * boundary rules of :6292-6433), written by hand: the adjoint of the pull
 * stream is a push along the same links, the collision's transpose is
 * applied per node, boundary values feed back to their neighbour's moments.
 * Populations are checkpointed every checkpoint_interval iterations in the
 * forward pass; the backward pass recomputes each segment from its
 * checkpoint and runs the adjoint steps through it in reverse.
 *
 * Objective: J = 1/((t_end - t_begin + 1) npts) * sum over iterations
 * t_begin..t_end and over the npts wall points of |tau|, with tau the wall
 * shear stress of zf_wss (mu_over_dx) on the populations after that
 * iteration.  Gradients: inlet amplitude and u0 of the velocity inlet
 * (ZF_B_VELOCITY_INLET), the uniform body force, the initial populations,
 * and the per-node force field when one is uploaded.  FP64 and FP32
 * storage (the adjoint populations use the storage's arithmetic type);
 * not FP16.  This entry point has no counterpart in the reference, which
 * has no adjoint. */
typedef struct zf_adjoint_config {
  int32_t checkpoint_interval;  /* K */
  int32_t max_steps;            /* longest forward pass to record */
  int32_t t_begin, t_end;       /* objective window (iteration numbers, inclusive) */
  double mu_over_dx;            /* WSS scale, as in zf_wss */
  int32_t cells;                /* 1: the forward pass is zf_step with its RBC
                                   iterations, and the gradient runs through them */
  int32_t reserved;
} zf_adjoint_config;

int zf_adjoint_setup(zf_ctx *ctx, const zf_adjoint_config *ac, int64_t npts, const double *xyz,
                     const double *normal);
/* Runs iterations iter0+1 .. iter0+nsteps from the current state (which
 * it advances, as zf_lbm_run does), records checkpoints and returns J. */
int zf_adjoint_forward(zf_ctx *ctx, int32_t iter0, int32_t nsteps, double *objective);
/* Backward pass of the last forward pass: grad[0] = dJ/d inlet_amp,
 * grad[1] = dJ/d inlet_u0, grad[2..4] = dJ/d body_force.  The state is
 * left as the forward pass left it. */
int zf_adjoint_backward(zf_ctx *ctx, double grad[5]);
/* dJ/d(initial populations), 19*dommax in the layout of
 * zf_upload_populations; dJ/d(force field) 3*dommax (zeros without one). */
int zf_adjoint_download(zf_ctx *ctx, double *dJ_df0, double *dJ_dforce);
/* With cells: dparam[0] = dJ/d ln es (every constant proportional to the
 * shear modulus scaled together, as the lattice-unit set-up makes them),
 * dparam[1] = dJ/d ln kbend, dJ_dx0 = dJ/d(cell positions at the start of
 * the forward pass), ncell*nn*3 (either may be NULL).  Each RBC iteration
 * of the forward pass is recorded (positions, spread node forces, force
 * field); the backward pass runs the transposes of the position update,
 * interpolation, spreading and membrane forces through each.  The
 * repulsion velocities enter as recorded values (their dependence on the
 * positions is not differentiated: exact while no pair is within range),
 * and no cell may be removed by the speed check during the pass. */
int zf_adjoint_cell_gradients(zf_ctx *ctx, double dparam[2], double *dJ_dx0);
/* Milliseconds of the last forward and backward passes (HIP events). */
int zf_adjoint_timing(zf_ctx *ctx, double ms[2]);
/* Inlet and body-force parameters of later iterations (the gradients'
 * variables), without rebuilding the context. */
int zf_set_forcing(zf_ctx *ctx, double inlet_u0, double inlet_amp, const double body_force[3]);

/* ---- transposes of the immersed-boundary coupling --------------------------
 * No counterpart in the reference (it has no adjoint); for callers that
 * differentiate through zf_ibm_interpolate / zf_ibm_spread.  With the
 * current positions X and populations f, for the interpolation
 * V = vs sum_q w(x_q - X) u_q and the spreading F(s) = fs sum_n w(x_s - X_n) G_n:
 *   lam_node_force = (dF/dG)^T lam_force               ncell*nn*3 (optional)
 *   lam_pos        = (dF/dX)^T lam_force + (dV/dX)^T lam_vel    ncell*nn*3
 *   lam_f          = (dV/df)^T lam_vel                  19*dommax (optional)
 * lam_force: 3*dommax (list order, NULL = 0); lam_vel, node_force:
 * ncell*nn*3 (NULL = 0; node_force is required with lam_force).  FP64
 * arithmetic on the stored populations. */
int zf_ibm_vjp(zf_ctx *ctx, const double *lam_force, const double *lam_vel, const double *node_force,
               double *lam_node_force, double *lam_pos, double *lam_f);

/* ---- transpose of the membrane forces ---------------------------------------
 * No counterpart in the reference; for callers that differentiate through
 * zf_membrane_forces.  At the current positions, for J = <lam_force, F>:
 *   lam_pos  = dJ/dX (ncell*nn*3, as lam_force)
 *   dparam[0] = dJ/d ln es: every constant proportional to the shear modulus
 *               (kT, kp, kd, ka, kv, the anti-sliver forces, as the lattice-
 *               unit set-up makes them) scaled together
 *   dparam[1] = dJ/d ln kbend
 * Exact local derivatives (dual-number evaluation of the force terms) plus
 * the cell centroid, area and volume terms; FP64. */
int zf_membrane_vjp(zf_ctx *ctx, const double *lam_force, double *lam_pos, double dparam[2]);

/* ---- state checkpoints ------------------------------------------------------
 * The reference restarts from its macroscopic CSV output and re-equilibrates
 * the populations (ZebraflowWT.c:5253-5262); these keep the exact state.
 * zf_checkpoint_save copies the current populations (as stored), the force
 * field, and with a membrane mesh the cells (positions, velocities, membrane
 * force, repulsion velocity, status) into device slot `slot` (0..1023,
 * allocated on first use); zf_checkpoint_restore copies them back, after
 * which the context continues bit-identically.  A slot restores only into
 * the configuration it was saved from. */
int zf_checkpoint_save(zf_ctx *ctx, int32_t slot);
int zf_checkpoint_restore(zf_ctx *ctx, int32_t slot);
/* Cell status (ncell; > -4998 live, -5000 removed or free; see the cells
 * on the parent lattices for 2, 4 and the clones' negative statuses), for
 * restarting from files (zf_upload_cells marks every uploaded cell live). */
int zf_upload_cell_status(zf_ctx *ctx, const int32_t *status);

/* ---- timing and synchronisation ------------------------------------------ */
/* Device-copy ceiling: a plain 16-byte-per-lane copy of `bytes` between two
 * fresh buffers, `reps` times; *gbps = 2*bytes*reps / time (read + write). */
int zf_copy_bandwidth(zf_ctx *ctx, int64_t bytes, int32_t reps, double *gbps);
int zf_sync(zf_ctx *ctx);
/* Record HIP event `slot` (0..15) on the context stream. */
int zf_event_record(zf_ctx *ctx, int32_t slot);
int zf_event_elapsed(zf_ctx *ctx, int32_t slot_a, int32_t slot_b, float *ms);

#ifdef __cplusplus
}
#endif
#endif
