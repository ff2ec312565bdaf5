// zf_internal.h -- shared state and helpers of libzfgpu.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/zf.h"

namespace zf {

// D3Q19 in the reference's order, ZebraflowWT.c:2471-2504: c and c^1 are
// opposites for c < 18, c = 18 is the rest direction.
constexpr int Q = 19;
constexpr int kE[Q][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
    {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0},
    {1, 0, 1},  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {0, 1, 1},
    {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 0, 0}};
// Wcoeff, ZebraflowWT.c:2329-2333.
constexpr double kW[Q] = {
    1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 3};

// Node word of the stream-collide kernel (one per interior node):
//   bits  0..17  halfway bounce-back mask (link c pulls f_{c^1} of the node)
//   bit   18     the node writes the boundary nodes that take it as neighbour
//   bits 20..31  chord index relative to the first chord of its 64-node block
constexpr uint32_t kWordMacro = 1u << 18;
constexpr int kWordChordShift = 20;
constexpr int kBlockNodes = 64;
// Chord record: 18 pull deltas (storage index of the source minus the
// node's own), the tau class B, and padding to 80 bytes.
constexpr int kChordInts = 20;

// One active boundary node for the boundary kernel.
struct BcEntry {
  int32_t s;      // storage index of the BC node
  int32_t sm;     // storage index of its neighbour (BCNLIST)
  int32_t code;   // B value of the BC node
  int32_t m_int;  // neighbour is updated by stream-collide (0<B<10): its rho,u
                  // come from this step's macro output, else from moments
};

// One coupled boundary node of a parent lattice: its populations are set to
// feq of moments taken from the main lattice (and, upstream, the parent's
// own density at its neighbour).
struct ParentBc {
  int32_t s;     // storage index of the BC node (parent)
  int32_t sm;    // storage index of its parent neighbour (upstream), else -1
  int32_t slot;  // slot of the main-lattice moments in the main's export buffer
  int32_t kind;  // 0 upstream, neighbour interior; 1 upstream, neighbour not
                 // interior; 2 copied from the main lattice
};

// One saved state (zf_checkpoint_save): regions packed at 256-byte offsets.
struct StateSlot {
  void *buf = nullptr;
  size_t bytes = 0;
  bool force = false;
  int32_t ncell = -1;  // -1: no cells in the slot
  bool nuc = false, nuc_lists_valid = false;  // nucleus regions held; their lists' validity
};

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
};

// Workgroup b of nblk renumbered so that each of the 8 XCDs (the hardware
// deals workgroups to them round-robin) runs one contiguous range of the
// grid: neighbouring workgroups (the nodes of one cell, cells adjacent in
// index order) then share one XCD's L2.  A permutation of [0, nblk).
__device__ __forceinline__ int xcd_range_block(int b, int nblk) {
  const int q = nblk / 8, r = nblk % 8, x = b % 8, i = b / 8;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Entries [beg, end) of a list of n that the workgroups on XCD blockIdx.x % 8
// share: 8 contiguous ranges, one per XCD, so that entries next to each
// other in the list (neighbouring nodes or tiles) meet in one L2.  The grid
// must be a multiple of 8; workgroup b takes every (gridDim.x / 8)-th part
// of its range from part b / 8 on.
__device__ __forceinline__ void xcd_list_span(int n, int &beg, int &end) {
  const int x = blockIdx.x % 8;
  beg = (int)((long long)n * x / 8);
  end = (int)((long long)n * (x + 1) / 8);
}

struct PipeState;  // two-step launches (zf_pipe.hip)

}  // namespace zf

struct zf_ctx {
  zf_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t events[16] = {};
  std::string err;

  // lattice
  bool have_lists = false;    // index lists built (downloadable)
  bool have_lattice = false;  // stream tables built and state allocated
  int32_t dommax = 0, bcmax = 0, n_int = 0, n_chords = 0, n_bc = 0;
  int32_t n_bb_links = 0, n_macro = 0;
  size_t stride = 0;          // elements per direction in the population arrays
  size_t box = 0;
  int32_t *d_b = nullptr;     // box B_index
  int32_t *d_domid = nullptr; // box DOMID (1-based or -1)
  int32_t *d_domlist = nullptr, *d_bclist = nullptr, *d_bcnlist = nullptr;
  int32_t *d_nbr = nullptr;   // 18 x dommax
  int32_t *d_s_of_p = nullptr;  // list position -> storage index
  int32_t *d_p_of_s = nullptr;  // storage index -> list position
  uint32_t *d_word = nullptr;   // n_int
  int32_t *d_block_chord0 = nullptr;
  int32_t *d_chords = nullptr;  // n_chords x kChordInts
  zf::BcEntry *d_bc = nullptr;  // n_bc: fused entries (by neighbour) then late ones
  int32_t n_bc_fused = 0, n_fused_nodes = 0;
  int32_t *d_fused_s = nullptr;   // sorted storage indices of nodes with fused entries
  int32_t *d_att_start = nullptr; // n_fused_nodes+1: their ranges in d_bc
  std::vector<int32_t> h_bnode;  // B value per list position

  // state
  void *d_f[2] = {nullptr, nullptr};  // population buffers (storage-typed)
  int cur = 0;                        // buffer holding the current state
  void *d_force = nullptr;            // 3 x stride, storage-indexed, field_elem bytes each; null = none
  double *d_sums = nullptr;           // reduction scratch
  int64_t device_bytes = 0;
  zf::PipeState *pipe = nullptr;      // tables of the two-step launch, built on first use
  bool pipe_used = false;             // a two-step launch since the last error check

  // membranes (shared mesh, per-cell state)
  bool have_mesh = false;
  zf_membrane_config mcfg{};
  int32_t nn = 0, nt = 0, ne = 0, ncell = 0;
  int32_t inc_max = 0;            // largest (edge, role) incidence count of a node
  std::vector<int32_t> h_edge, h_nbr_node, h_nbr_tri;  // reference row formats
  std::vector<double> h_eq;       // 6 x ne: lmax plength kp theta0 sin0 cos0
  int32_t *d_tri = nullptr;       // 3 x nt, 0-based
  int32_t *d_edge = nullptr;      // 8 x ne: n1 n2 n3 n4 t1 t2 own1 own2 (0-based)
  double *d_edge_eq = nullptr;    // 5 x ne: lmax plength kp sin0 cos0
  double *d_tri_area_eq = nullptr;
  int32_t *d_node_tri = nullptr;  // nn x 10 (count, tris 0-based)
  int32_t *d_inc_start = nullptr, *d_inc = nullptr;  // node -> (edge<<2|role), edge order
  double *d_x = nullptr;          // 3 x (ncap*nn) coordinates
  double *d_v = nullptr;          // 3 x (ncap*nn) node velocities
  double *d_mforce = nullptr;     // 3 x (ncap*nn) membrane force F_Total
  double *d_nnormal = nullptr;    // 3 x (ncap*nn) node normals
  double *d_tgeo = nullptr;       // 7 x (ncap*nt): area, normal(3), centroid(3)
  double *d_cell = nullptr;       // 8 x ncap: area, vol, cx, cy, cz, F_area, F_vol, spare
  int32_t *d_status = nullptr;    // ncap: >=0 live, <0 removed
  int32_t *d_fail = nullptr;      // ncap: guard failures this evaluation

  // nucleus network of each cell (nucleated RBCs)
  bool have_nuc = false;
  zf_nucleus_config ncfg{};
  int32_t nnuc = 0, ne_n = 0;
  std::vector<int32_t> h_nuc_edge;  // 2 x ne_n, 1-based (input order)
  std::vector<double> h_nuc_eq;     // 3 x ne_n: lmax, plength, kp
  int32_t *d_nuc_edge = nullptr;    // 2 x ne_n, 0-based
  double *d_nuc_eq = nullptr;       // 3 x ne_n
  int32_t *d_nuc_inc_start = nullptr, *d_nuc_inc = nullptr;  // node -> edge << 1 | role, edge order
  double *d_xn = nullptr, *d_vn = nullptr;    // 3 x (ncap*nnuc): positions, velocities
  double *d_fn = nullptr, *d_vrepn = nullptr; // 3 x (ncap*nnuc): F_Total_N, NODE_WALLREPUL_N
  void *d_ffln = nullptr;           // 3 x (ncap*nnuc): F_fluid of the nucleus nodes, field_elem
  int32_t *d_nlist = nullptr;       // (ncap*nnuc) x list_cap: listed triangles (0-based)
  int32_t *d_nlist_cnt = nullptr;   // ncap*nnuc, then one overflow counter
  int32_t *d_nuc_stale = nullptr;   // ncap: the cell's lists are rebuilt at the next contact
  bool nuc_lists_valid = false;     // cleared when cells or nuclei are uploaded
  size_t keys_cap = 0;              // entries of the spreading sort keys

  // immersed boundary
  bool have_ibm = false;
  zf_ibm_config icfg{};
  int32_t *d_sbox = nullptr;      // box: storage index of the listed node or -1
  void *d_umac = nullptr;         // 3 x stride: U,V,W without force shift (s order), field_elem
  double *d_vrep = nullptr;       // 3 x (ncap*nn) repulsion velocity v_rep
  void *d_ffl = nullptr;          // 3 x (ncap*nn): F_fluid of each node, field_elem
  uint64_t *d_keys[2] = {nullptr, nullptr};  // (tile << 32 | node) sort keys
  void *d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int32_t *d_tile_range = nullptr;  // 2 x ntiles: [begin, end) in the sorted keys
  int32_t *d_sbin_range = nullptr;  // 2 x 8 ntiles: the same per spreading bin (tile octant)
  int32_t *d_tile_occ = nullptr;    // ntiles: 1 where a tile holds a spreading node
  int32_t *d_spread_tiles = nullptr;  // 1 + ntiles: count, then the tiles the gather visits
  int32_t *d_tile_act = nullptr;    // ntiles: 1 where the gather visits a tile
  void *d_sel_tmp = nullptr;        // temporary storage of the ordered list selections
  size_t sel_tmp_bytes = 0;
  int32_t *d_flag = nullptr;        // ncap: speed check
  int ntx = 0, nty = 0, ntz = 0;

  // cell list and repulsion
  bool have_rep = false, have_wall = false;
  zf_repulsion_config rcfg{};
  uint64_t *d_tkeys[2] = {nullptr, nullptr};  // (bin << 32 | triangle) sort keys
  float4 *d_crec = nullptr;  // per sorted key: FP32 centroid and the packed normal
  void *d_tsort_tmp = nullptr;
  size_t tsort_tmp_bytes = 0;
  int32_t *d_tbin = nullptr;      // ncap*nt: 1-based DOMID of the triangle's bin or 0
  int32_t *d_bin_range = nullptr; // 2 x dommax: [begin, end) in the sorted keys
  int32_t *d_bin_cnt = nullptr;   // dommax + 1: keys per bin (scan input)
  int32_t *d_bin_off = nullptr;   // dommax + 1: first key of each bin, then the key count
  uint32_t *d_row_bd = nullptr;   // im*jm: largest centroid-to-bin distance per (i, j) row, FP32 bits
  void *d_bin_scan_tmp = nullptr;
  size_t bin_scan_tmp_bytes = 0;
  int32_t *d_near = nullptr;      // ncap*nn: chosen wall triangle (1-based) or 0
  int32_t *d_npair = nullptr;     // ncap*nn: passing triangles of each node (-1: not a live node)
  int32_t *d_pairs = nullptr;     // ncap*nn x 16: the first of them, in walk order
  int32_t *d_ochunk = nullptr;    // ncap*nn: first pool chunk of the rest
  int32_t *d_rep_pool = nullptr;  // chunk counter, then chunks of 31 triangles + next chunk
  int32_t *d_over = nullptr;      // 1 + ncap*nn: count, then the nodes whose list overflowed
  int32_t *d_rep_lists = nullptr; // nodes with listed triangles, by count class
  void *d_rsel_tmp = nullptr;     // temporary storage of the near-wall node selection
  size_t rsel_tmp_bytes = 0;
  // bounds recorded with the cell list: largest centroid-to-bin and
  // centroid-to-vertex distance (FP64 bits); valid while the triangle
  // geometry is the one the list was built from (geom_gen == bins_gen)
  unsigned long long *d_rep_bounds = nullptr;
  uint64_t geom_gen = 0, bins_gen = ~0ull;
  int32_t nwn = 0, nwt = 0;
  double wall_lmax = 0;  // largest vertex-to-vertex or vertex-to-centroid distance of a wall triangle
  double *d_wxyz = nullptr, *d_wnormal = nullptr;  // 3 x nwn, 3 x nwt (AoS)
  int32_t *d_wtri = nullptr, *d_wnode_tri = nullptr;  // 3 x nwt (1-based), 13 x nwn
  double *d_wtgeo = nullptr;  // 12 x nwt: vertices and centroid of each wall triangle
  float4 *d_wsph = nullptr;   // nwt: FP32 bounding sphere of each wall triangle
  float4 *d_wn_sph = nullptr; // 12 x nwn: the spheres of each wall node's triangle slots
  float4 *d_wpos = nullptr;   // nwn: FP32 position and wrad rounded up
  float4 *d_wnrm = nullptr;   // nwn: FP32 unit normal and the largest offset of its points along it
  double *d_wn_geo = nullptr; // 144 x nwn: the 12-double records of the same slots
  double *d_wrad = nullptr;   // nwn: largest distance from a wall node to its triangles' vertices/centroids
  int32_t *d_windex = nullptr;    // box
  float *d_wdist = nullptr;       // box: FP32 distance to that wall node (or +inf)
  // wall-layer velocity Uvoid: slot per box node (-1: none), 3 doubles per slot
  bool have_void = false;
  int32_t n_void = 0;
  int32_t *d_vslot = nullptr;
  double *d_uvoid = nullptr;

  // adjoint of the LBM step with a wall-shear-stress objective
  bool have_adj = false, adj_ready = false;  // set up / forward recorded
  zf_adjoint_config acfg{};
  int32_t adj_npts = 0, adj_iter0 = 0, adj_nsteps = 0;
  int32_t obj_batch = 1;           // states per batch of the objective's adjoint passes
  size_t adj_state_bytes = 0;
  double *d_adj_pts = nullptr;     // 6 x npts: xyz then unit normals
  void *d_ckpt = nullptr;          // checkpoints, one state every K iterations
  void *d_seg = nullptr;           // K - 1 recomputed states of one segment
  void *d_lam[2] = {nullptr, nullptr};  // adjoint populations (ping-pong)
  int adj_cur = 0;
  double *d_adj_a = nullptr;       // obj_batch x 3 x npts: -(mu/DX) dJ/dtau of each probe
  int32_t *d_support = nullptr;    // storage indices in any probe stencil
  int32_t *d_sup_off = nullptr;    // n_support + 1: CSR over (probe, weight) pairs
  int32_t *d_sup_probe = nullptr;
  double *d_sup_w = nullptr;
  double *d_sup_u = nullptr;       // obj_batch x 4 x n_support: u and 1 + drho of the support nodes
  double *d_sup_g = nullptr;       // obj_batch x 3 x n_support: dJ/du of the support nodes
  int32_t *d_prb_off = nullptr;    // npts + 1: CSR over (support index, weight) per probe
  int32_t *d_prb_sup = nullptr;
  double *d_prb_w = nullptr;
  int32_t n_support = 0;
  double *d_gpart = nullptr;       // 5 per workgroup (+1 row): gradient partial sums
  int64_t *d_unpulled = nullptr;   // (c*stride + s) read by the adjoint but pulled by no node
  int64_t n_unpulled = 0;
  double *d_grads = nullptr;       // dJ/d(inlet_amp, inlet_u0, g_x, g_y, g_z)
  double *d_gF = nullptr;          // 3 x stride: dJ/d(force field), when a field exists
  double *d_objbuf = nullptr;      // |tau| per objective iteration and probe
  float adj_ms[2] = {0, 0};        // forward, backward
  // cells in the adjoint: records of every RBC iteration of the forward pass
  bool adj_cells = false;
  int32_t adj_rec_cap = 0, adj_nrec = 0;
  double *d_rec_x = nullptr;       // per RBC iteration: positions before the update, 3 x ncap*nn
  double *d_rec_g = nullptr;       // per RBC iteration: spread node forces gamma_p v_rep + F_Total
  void *d_rec_field = nullptr;     // force field at the start, then after each RBC iteration, 3 x stride
  double *d_lamX = nullptr;        // dJ/d positions, 3 x ncap*nn
  double *d_lamV = nullptr, *d_lamG = nullptr;  // scratch of one RBC iteration, 3 x ncap*nn
  double *d_mscratch = nullptr;    // membrane transpose scratch
  double *d_cgrad = nullptr;       // dJ/d ln es, dJ/d ln kbend
  std::vector<int32_t> h_status0;  // cell status at the start of the forward pass

  std::vector<zf::StateSlot> slots;  // zf_checkpoint_save

  // parent (feeder) lattice p = 1..3 (0: a main lattice), its overlap plane
  // and the offset of its index space in the main lattice's
  int32_t parent_p = 0, parent_kover = 0, parent_off[3] = {0, 0, 0};
  int32_t *d_vbcn = nullptr;       // 3 x bcmax: upstream velocity source (main indices)
  zf_ctx *link_main = nullptr;     // main lattice this parent takes its boundary from
  zf::ParentBc *d_pbc = nullptr;   // coupled boundary nodes
  int32_t n_pbc = 0;
  hipStream_t own_stream = nullptr;  // while linked: its own stream (it uses the main's)
  // main lattice: moments of the nodes its linked parents read, written
  // by every step after the stream kernel
  std::vector<zf_ctx *> linked;
  std::vector<int32_t> h_export;   // per slot: storage index << 1 | interior
  int32_t *d_export_s = nullptr;
  double *d_export = nullptr;      // 4 x n_export: rho, u, v, w

  // cells on a parent lattice: the periodic cell window along k
  // [wz0, wzm] (wzm = wz0 + (km-1) dx), node coordinates wrapped into it
  bool have_wrap = false;
  zf_cell_wrap_config wcfg{};
  double wz0 = 0, wzm = 0;
  zf_ctx *wrap_main = nullptr;     // main lattice its clones go to
  int32_t *d_regress = nullptr;    // ncap: REGRESSFLAG
  int32_t *d_cloneflag = nullptr;  // ncap: CLONEFLAG of the last update
  double *d_span = nullptr;        // 2 x ncap: smallest and largest z (membrane and nucleus)
  int32_t *d_split = nullptr;      // ncap: the cell lies across the wrap (status > 3, span > split)
  double *d_xu = nullptr;          // 3 x (ncap*nn): the positions the membrane stages use
  double *d_xnu = nullptr;         // 3 x (ncap*nnuc): same for the nucleus
  int32_t *d_clone_pairs = nullptr;  // 1 + 2 x ncap: count, (parent slot, main slot) pairs
  // main lattice: parents whose cells clone into it, and per cell the
  // clone band of a clone (plane, direction; direction 0 otherwise)
  std::vector<zf_ctx *> cell_parents;
  double *d_band = nullptr;        // 2 x ncap
};

namespace zf {

int fail(zf_ctx *ctx, int code, const char *fmt, ...);
// HIP error check that records the failing call in the context.
#define ZF_HIP(ctx, call)                                                     \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::zf::fail((ctx), ZF_ERR_HIP, "%s failed: %s (%s:%d)", #call,    \
                        hipGetErrorString(e_), __FILE__, __LINE__);           \
  } while (0)

int dev_alloc(zf_ctx *ctx, void **p, size_t bytes);
void dev_free(zf_ctx *ctx, void **p);
int upload_lattice_impl(zf_ctx *ctx, const int32_t *b);
void free_lattice(zf_ctx *ctx);
void free_state(zf_ctx *ctx);
void free_membrane(zf_ctx *ctx);
void free_ibm(zf_ctx *ctx);
void free_cells(zf_ctx *ctx);
void free_adjoint(zf_ctx *ctx);
void free_state_slots(zf_ctx *ctx);
int membrane_forces_impl(zf_ctx *ctx);
int cell_list_impl(zf_ctx *ctx);
int repulsion_impl(zf_ctx *ctx);
// Nucleus network (zf_nucleus.hip): spring forces F_Total_N of every live
// cell; the nucleus-membrane contact of iteration iter (lists rebuilt on
// the rebuild rule), added to the membrane's repulsion velocity.
int nucleus_forces_impl(zf_ctx *ctx);
int nucleus_contact_impl(zf_ctx *ctx, int iter);
void free_nucleus(zf_ctx *ctx);
// Spreading sort keys for n entries (allocates only when they do not fit).
int ensure_sort_keys(zf_ctx *ctx, size_t n);
// Physical position of lattice node (i,j,k): (x0 + (i-1) dx, ...); the IBM
// configuration's when set up, lattice units otherwise.
struct GridGeom {
  double x0, y0, z0, dx;
};
GridGeom grid_geom(const zf_ctx *ctx);
// Looks up the allocation-free WSS scratch of ctx's device (once per device).
int wss_scratch_init(zf_ctx *ctx);
int alloc_state(zf_ctx *ctx);
size_t store_bytes(int storage);
// Two LBM steps per launch (zf_pipe.hip): whether ctx's lattice takes them
// (builds the tables on first use), the launch pair iter1, iter1 + 1 (the
// force field cleared after it when iter1 + 1 is an RBC iteration), the
// error check after a synchronisation, and the tables' release.
bool pair_eligible(zf_ctx *ctx);
int step_pair(zf_ctx *ctx, int iter1, int rbc2);
int pipe_check(zf_ctx *ctx);
void free_pipe(zf_ctx *ctx);
// zf_lbm.hip: whether the two-step launch is taken without
// ZF_FLAG_PAIRED_STEPS (never: measured slower) and its cached directions.
bool pair_default(const zf_ctx *ctx);
int pair_cached_dirs(const zf_ctx *ctx);
// Bytes of one element of the force field and of the IBM velocity and
// node-force buffers: FP64 with FP64 storage (the FP64 form is FP64
// throughout), FP32 otherwise.
inline size_t field_elem(const zf_ctx *ctx) { return ctx->cfg.storage == ZF_STORE_F64 ? 8 : 4; }
// The cell stages of an RBC iteration after its LBM step (membrane forces,
// cell list, interpolation, repulsion, spreading) and, with `advance`, the
// position update: the order of zf_step.
// iter: the iteration (the nucleus contact's list rebuild rule).
int rbc_update(zf_ctx *ctx, bool advance, int iter = 0);
// Parent coupling (zf_parent.hip): the main lattice's export of the moments
// its parents read (in_step: after its stream kernel, before the buffer
// swap and the force field's clearing; else of the current state), and a
// parent's coupled boundary pass (after its own step and swap).
int parent_export(zf_ctx *ctx, bool in_step);
int parent_couple(zf_ctx *ctx);
void free_parent(zf_ctx *ctx);
int advance_impl(zf_ctx *ctx, int32_t *n_removed);
// Cells on parent lattices (zf_wrap.hip).  A cell is alive while its
// status is above -4998 (-5000: free slot or removed; -4998 < s < 0: a
// clone of parent cell -s still following it).
constexpr int32_t kStatusDead = -4998;
// The positions the membrane and nucleus stages read: a parent's cells
// across the wrap unwrapped (wrap_prepare), the stored ones otherwise.
inline const double *cell_pos(const zf_ctx *ctx) { return ctx->have_wrap ? ctx->d_xu : ctx->d_x; }
inline const double *nuc_pos(const zf_ctx *ctx) { return ctx->have_wrap ? ctx->d_xnu : ctx->d_xn; }
// Parent: span, split flags and unwrapped positions (start of an RBC
// iteration); REGRESSFLAG reset from the centroids (after the membrane
// stage); the wrapped update with clone flags; the cloning into the main
// lattice.  Main: the update of its clones (band nodes with the parent
// cell's velocity, release into status 1).
int wrap_prepare(zf_ctx *ctx);
int wrap_regress(zf_ctx *ctx);
int wrap_advance(zf_ctx *ctx);
int wrap_clone(zf_ctx *ctx);
int clone_advance(zf_ctx *main);
void free_wrap(zf_ctx *ctx);
// The parent's stages with the periodic window (zf_ibm.hip, zf_cells.hip).
int wrap_interpolate_impl(zf_ctx *ctx);
int wrap_spread_impl(zf_ctx *ctx);
int wrap_cell_list_impl(zf_ctx *ctx);
int wrap_wall_impl(zf_ctx *ctx);
int wrap_repulsion_impl(zf_ctx *ctx);
int ensure_force_field(zf_ctx *ctx);
// Transposes of the IBM coupling at positions X and populations f (device
// arrays, SoA over the cell capacity; lamF and lam storage-indexed; lam of
// the adjoint arithmetic type, double when lam_double).  lamX and lam are
// accumulated, lamG written; null inputs count as zero.
int ibm_vjp_device(zf_ctx *ctx, const double *X, const void *f, const double *G,
                   const double *lamV, const double *lamF, double *lamG, double *lamX, void *lam,
                   bool lam_double);
// Transpose of the membrane forces at positions X: lamX += (dF/dX)^T lam,
// gparam[0..1] += dJ/d ln es, dJ/d ln kbend; scratch of
// membrane_vjp_scratch(ctx) bytes.
size_t membrane_vjp_scratch(const zf_ctx *ctx);
// Cells in the adjoint (zf_adjoint_cells.hip): buffers, the record of RBC
// iteration j of the forward pass (after spreading, before the position
// update), and its backward step: lamX holds dJ/d(positions after the
// update) on entry and dJ/d(positions before it) on return; gF (dJ/d the
// field spread at this iteration) is consumed and cleared; lam (dJ/d the
// populations the interpolation read) gains the interpolation's transpose.
int adjoint_cells_alloc(zf_ctx *ctx);
void free_adjoint_cells(zf_ctx *ctx);
int adjoint_cells_record(zf_ctx *ctx, int j);
int adjoint_cells_backward(zf_ctx *ctx, int j, const void *f, void *lam);
int membrane_vjp_device(zf_ctx *ctx, const double *X, const double *lam, double *lamX, double *gparam,
                        double *scratch);

}  // namespace zf
