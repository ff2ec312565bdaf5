// zf_wrap.hip -- cells on the parent (feeder) lattices.
//
// Replaces, for a parent lattice's cells and their clones in the main
// lattice, ZebraflowWT.c:
//   CAPS_MAX/MIN and the unwrapping of a cell across the wrap
//                                    :5709-5741, :5769-5797, :6738-6756
//   REGRESSFLAG reset                :6762-6778
//   wrapped position update, flags   :12196-12340
//   serial cloning into the ROI      :12345-12403
//   the clones' update in the ROI    :11927-11962
// The parent's lattice stages with the periodic window are in zf_ibm.hip
// (interpolation, spreading) and zf_cells.hip (cell list, repulsion).
// Parent cells are few (the reference's parents hold 188, 141 and 15
// slots), so these kernels take one workgroup per cell and reduce in LDS.
#include "zf_internal.h"

#include <algorithm>
#include <vector>

namespace zf {
namespace {

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct WrapWin {
  int dir;
  double z0, zm, L, mid;
};

WrapWin win_of(const zf_ctx *ctx) {
  return WrapWin{ctx->wcfg.direction, ctx->wz0, ctx->wzm, ctx->wzm - ctx->wz0,
                 0.5 * (ctx->wz0 + ctx->wzm)};
}

// Block reductions over 256 lanes.
__device__ double block_min(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ double block_max(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// CAPS_MIN/MAX[2] over the membrane and nucleus nodes of each cell
// (:5709-5741) and the split flag: status > 3 and a span above `split`
// (:5769-5770).
__global__ __launch_bounds__(256) void k_wrap_span(const double *__restrict__ X, size_t NT, int nn,
                                                   const double *__restrict__ Xn, size_t NTn, int nnuc,
                                                   const int32_t *__restrict__ status, double split,
                                                   double *__restrict__ span, int32_t *__restrict__ sflag) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double mn = 5000.0, mx = -5000.0;
  if (status[c] > kStatusDead) {
    for (int n = threadIdx.x; n < nn; n += 256) {
      const double z = X[2 * NT + (size_t)c * nn + n];
      mn = fmin(mn, z);
      mx = fmax(mx, z);
    }
    for (int n = threadIdx.x; n < nnuc; n += 256) {
      const double z = Xn[2 * NTn + (size_t)c * nnuc + n];
      mn = fmin(mn, z);
      mx = fmax(mx, z);
    }
  }
  mn = block_min(mn, sh);
  mx = block_max(mx, sh);
  if (threadIdx.x == 0) {
    span[2 * c] = mn;
    span[2 * c + 1] = mx;
    sflag[c] = status[c] > 3 && mx - mn > split ? 1 : 0;
  }
}

// The positions of the membrane and nucleus stages: a split cell's nodes
// behind the window's middle moved by L past its exit.
__global__ void k_wrap_unwrap(const double *__restrict__ X, size_t NT, int nn, int ncell,
                              const int32_t *__restrict__ sflag, WrapWin w, double *__restrict__ XU) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (size_t)ncell * nn) return;
  const int c = (int)(g / nn);
  double z = X[2 * NT + g];
  if (sflag[c] && (w.dir > 0 ? z < w.mid : z > w.mid)) z += w.dir > 0 ? w.L : -w.L;
  XU[g] = X[g];
  XU[NT + g] = X[NT + g];
  XU[2 * NT + g] = z;
}

// REGRESSFLAG back to 0 (:6762-6778): not across the wrap, centroid in the
// window's middle half.
__global__ void k_wrap_regress(const int32_t *__restrict__ status, const double *__restrict__ span,
                               const double *__restrict__ cell, WrapWin w, double split, int ncell,
                               int32_t *__restrict__ regress) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell || status[c] <= kStatusDead) return;
  if (span[2 * c + 1] - span[2 * c] < split) {
    const double cz = cell[8 * c + 4];
    if (cz > w.z0 + 0.25 * w.L && cz < w.z0 + 0.75 * w.L) regress[c] = 0;
  }
}

// Moves node q by v dt and wraps it into the window; true when it left
// through the exit.
__device__ __forceinline__ bool move_wrap(double *X, const double *V, size_t NT, size_t q, double dt,
                                          const WrapWin &w) {
  for (int a = 0; a < 3; a++) X[a * NT + q] += V[a * NT + q] * dt;
  double z = X[2 * NT + q];
  bool out = false;
  if (w.dir > 0) {
    if (z >= w.zm) { out = true; z -= w.L; }
    else if (z < w.z0) z += w.L;
  } else {
    if (z <= w.z0) { out = true; z += w.L; }
    else if (z > w.zm) z -= w.L;
  }
  X[2 * NT + q] = z;
  return out;
}

// The parent's position update (:12196-12340), one workgroup per cell:
// status 2 raises CLONEFLAG when a node leaves through the exit with
// REGRESSFLAG 0; status 4 sets REGRESSFLAG and returns to 2 inside the
// window's first three quarters, 5 from the entry.
__global__ __launch_bounds__(256) void k_wrap_advance(double *__restrict__ X, const double *__restrict__ V,
                                                      size_t NT, int nn, double *__restrict__ Xn,
                                                      const double *__restrict__ Vn, size_t NTn, int nnuc,
                                                      int32_t *__restrict__ status,
                                                      int32_t *__restrict__ regress,
                                                      int32_t *__restrict__ cloneflag, WrapWin w,
                                                      double dt) {
  __shared__ double sh[256];
  __shared__ int anyout;
  const int c = blockIdx.x;
  const int st = status[c];
  if (st != 2 && st != 4) {
    if (threadIdx.x == 0) cloneflag[c] = 0;
    return;
  }
  if (threadIdx.x == 0) anyout = 0;
  const int reg = regress[c];
  __syncthreads();
  double mn = 5000.0, mx = -5000.0;
  bool out = false;
  for (int n = threadIdx.x; n < nn; n += 256) {
    const size_t q = (size_t)c * nn + n;
    out |= move_wrap(X, V, NT, q, dt, w);
    const double z = X[2 * NT + q];
    mn = fmin(mn, z);
    mx = fmax(mx, z);
  }
  for (int n = threadIdx.x; n < nnuc; n += 256) out |= move_wrap(Xn, Vn, NTn, (size_t)c * nnuc + n, dt, w);
  if (out) atomicOr(&anyout, 1);
  mn = block_min(mn, sh);
  mx = block_max(mx, sh);  // also orders the atomicOr before the read below
  if (threadIdx.x != 0) return;
  cloneflag[c] = (st == 2 && reg == 0 && anyout) ? 1 : 0;
  if (st == 4) {
    regress[c] = 1;
    if (w.dir > 0 && mn > w.z0 + 5.0 && mx < w.z0 + 0.75 * w.L) status[c] = 2;
    if (w.dir < 0 && mx < w.zm - 5.0 && mn > w.zm - 0.75 * w.L) status[c] = 2;
  }
}

// The serial cloning of one parent (:12345-12403) in one lane: each
// flagged cell in slot order takes the first free main slot; the parent
// cell becomes status 4 with REGRESSFLAG 1 either way.  pairs[0] = count.
__global__ void k_clone_plan(int32_t *__restrict__ pstatus, int32_t *__restrict__ pregress,
                             const int32_t *__restrict__ cloneflag, int pncell, int id_base,
                             int32_t *__restrict__ mstatus, int mncell, int32_t *__restrict__ pairs) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int cnt = 0, m = 0;
  for (int s = 0; s < pncell; s++) {
    if (cloneflag[s] <= 0) continue;
    while (m < mncell && mstatus[m] != -5000) m++;
    if (m < mncell) {
      mstatus[m] = -(id_base + s + 1);
      pairs[1 + 2 * cnt] = s;
      pairs[2 + 2 * cnt] = m;
      cnt++;
      m++;
    }
    pstatus[s] = 4;
    pregress[s] = 1;
  }
  pairs[0] = cnt;
}

// Copies the planned clones' nodes (positions unwrapped to straddle the
// exit plane, velocities); with band, records the clone band of the new
// main cell, and with stale, marks its nucleus lists for rebuilding.
__global__ void k_clone_copy(const int32_t *__restrict__ pairs, int maxpairs, int nn,
                             const double *__restrict__ PX, const double *__restrict__ PV, size_t PNT,
                             double *__restrict__ MX, double *__restrict__ MV, size_t MNT, WrapWin w,
                             double *__restrict__ band, double plane, int32_t *__restrict__ stale) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (size_t)maxpairs * nn) return;
  const int p = (int)(g / nn), n = (int)(g % nn);
  if (p >= pairs[0]) return;
  const int s = pairs[1 + 2 * p], m = pairs[2 + 2 * p];
  const size_t src = (size_t)s * nn + n, dst = (size_t)m * nn + n;
  for (int a = 0; a < 3; a++) {
    double x = PX[a * PNT + src];
    if (a == 2 && (w.dir > 0 ? x < w.mid : x > w.mid)) x += w.dir > 0 ? w.L : -w.L;
    MX[a * MNT + dst] = x;
    MV[a * MNT + dst] = PV[a * PNT + src];
  }
  if (n == 0 && band) {
    band[2 * m] = plane;
    band[2 * m + 1] = (double)w.dir;
  }
  if (n == 0 && stale) stale[m] = 1;
}

struct CloneSrc {
  const double *v, *vn;
  size_t NT, NTn;
  int id_base, cap;
};
struct CloneSrcs {
  CloneSrc s[8];
  int n;
};

// A clone node moves with the same node of its parent cell while it lies
// within the band (plane = exit plane + direction * band): not past the
// plane along the direction of travel (:11933-11941).
__device__ __forceinline__ bool in_band(double z, double plane, double dir) {
  return dir > 0 ? !(z > plane) : !(z < plane);
}

// The clones' update in the main lattice (:11927-11962), one workgroup per
// cell: band nodes take the parent node's velocity, every node moves, and
// a clone whose membrane lies inside (roi_lo, roi_hi) is released.
__global__ __launch_bounds__(256) void k_clone_advance(double *__restrict__ X, double *__restrict__ V,
                                                       size_t NT, int nn, double *__restrict__ Xn,
                                                       double *__restrict__ Vn, size_t NTn, int nnuc,
                                                       int32_t *__restrict__ status,
                                                       const double *__restrict__ band, CloneSrcs src,
                                                       double dt, double roi_lo, double roi_hi) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  const int st = status[c];
  if (!(st < 0 && st > kStatusDead)) return;
  const int id = -st;
  int k = 0;
  while (k < src.n && !(id > src.s[k].id_base && id <= src.s[k].id_base + src.s[k].cap)) k++;
  if (k == src.n) return;  // not a clone of a linked parent: left in place
  const CloneSrc ps = src.s[k];
  const size_t pslot = (size_t)(id - ps.id_base - 1);
  const double plane = band[2 * c], dir = band[2 * c + 1];
  double mn = 5000.0, mx = -5000.0;
  for (int n = threadIdx.x; n < nn; n += 256) {
    const size_t q = (size_t)c * nn + n;
    if (in_band(X[2 * NT + q], plane, dir))
      for (int a = 0; a < 3; a++) V[a * NT + q] = ps.v[a * ps.NT + pslot * nn + n];
    for (int a = 0; a < 3; a++) X[a * NT + q] += V[a * NT + q] * dt;
    mn = fmin(mn, X[2 * NT + q]);
    mx = fmax(mx, X[2 * NT + q]);
  }
  for (int n = threadIdx.x; n < nnuc; n += 256) {
    const size_t q = (size_t)c * nnuc + n;
    if (in_band(Xn[2 * NTn + q], plane, dir))
      for (int a = 0; a < 3; a++) Vn[a * NTn + q] = ps.vn[a * ps.NTn + pslot * nnuc + n];
    for (int a = 0; a < 3; a++) Xn[a * NTn + q] += Vn[a * NTn + q] * dt;
  }
  mn = block_min(mn, sh);
  mx = block_max(mx, sh);
  if (threadIdx.x == 0 && mn > roi_lo && mx < roi_hi) status[c] = 1;
}

__global__ void k_wrap_fill(int32_t *a, size_t n, int32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

int require_wrap_pair(zf_ctx *parent, zf_ctx *main) {
  if (!parent->have_ibm || !main->have_ibm)
    return fail(parent, ZF_ERR_STATE, "zf_ibm_setup on the parent and the main lattice first");
  if (parent->nn != main->nn || parent->nt != main->nt)
    return fail(parent, ZF_ERR_ARG, "the parent and the main lattice need the same membrane mesh");
  if (parent->icfg.rbc_interval != main->icfg.rbc_interval)
    return fail(parent, ZF_ERR_ARG, "the parent and the main lattice need the same RBC interval");
  if (parent->have_nuc != main->have_nuc || (parent->have_nuc && parent->nnuc != main->nnuc))
    return fail(parent, ZF_ERR_ARG, "the parent and the main lattice need the same nucleus set-up");
  return ZF_OK;
}

}  // namespace

void free_wrap(zf_ctx *ctx) {
  // a main lattice: its parents stop cloning into it
  if (!ctx->cell_parents.empty()) (void)hipStreamSynchronize(ctx->stream);
  for (zf_ctx *p : std::vector<zf_ctx *>(ctx->cell_parents)) free_wrap(p);
  ctx->cell_parents.clear();
  dev_free(ctx, (void **)&ctx->d_band);
  void **bufs[] = {(void **)&ctx->d_regress, (void **)&ctx->d_cloneflag, (void **)&ctx->d_span,
                   (void **)&ctx->d_split,   (void **)&ctx->d_xu,        (void **)&ctx->d_xnu,
                   (void **)&ctx->d_clone_pairs};
  for (auto b : bufs) dev_free(ctx, b);
  if (ctx->wrap_main) {
    (void)hipStreamSynchronize(ctx->stream);
    auto &v = ctx->wrap_main->cell_parents;
    v.erase(std::remove(v.begin(), v.end(), ctx), v.end());
  }
  ctx->wrap_main = nullptr;
  ctx->have_wrap = false;
}

int wrap_prepare(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const size_t NTn = ctx->have_nuc ? ncap * ctx->nnuc : 0;
  const int nc = ctx->ncell, nnuc = ctx->have_nuc ? ctx->nnuc : 0;
  if (nc == 0) return ZF_OK;
  const WrapWin w = win_of(ctx);
  hipLaunchKernelGGL(k_wrap_span, dim3(nc), dim3(256), 0, ctx->stream, ctx->d_x, NT, ctx->nn,
                     ctx->d_xn, NTn, nnuc, ctx->d_status, ctx->wcfg.split_span, ctx->d_span,
                     ctx->d_split);
  hipLaunchKernelGGL(k_wrap_unwrap, dim3(nb((size_t)nc * ctx->nn, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_x, NT, ctx->nn, nc, ctx->d_split, w, ctx->d_xu);
  if (nnuc > 0)
    hipLaunchKernelGGL(k_wrap_unwrap, dim3(nb((size_t)nc * nnuc, 256)), dim3(256), 0, ctx->stream,
                       ctx->d_xn, NTn, nnuc, nc, ctx->d_split, w, ctx->d_xnu);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int wrap_regress(zf_ctx *ctx) {
  if (ctx->ncell == 0) return ZF_OK;
  hipLaunchKernelGGL(k_wrap_regress, dim3(nb(ctx->ncell, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_status, ctx->d_span, ctx->d_cell, win_of(ctx), ctx->wcfg.split_span,
                     ctx->ncell, ctx->d_regress);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int wrap_advance(zf_ctx *ctx) {
  ctx->geom_gen++;
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const size_t NTn = ctx->have_nuc ? ncap * ctx->nnuc : 0;
  if (ctx->ncell == 0) return ZF_OK;
  hipLaunchKernelGGL(k_wrap_advance, dim3(ctx->ncell), dim3(256), 0, ctx->stream, ctx->d_x, ctx->d_v,
                     NT, ctx->nn, ctx->d_xn, ctx->d_vn, NTn, ctx->have_nuc ? ctx->nnuc : 0,
                     ctx->d_status, ctx->d_regress, ctx->d_cloneflag, win_of(ctx),
                     ctx->icfg.dt_advance);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int wrap_clone(zf_ctx *ctx) {
  zf_ctx *mc = ctx->wrap_main;
  if (ctx->ncell == 0 || !mc || mc->ncell == 0) return ZF_OK;
  const WrapWin w = win_of(ctx);
  mc->geom_gen++;
  hipLaunchKernelGGL(k_clone_plan, dim3(1), dim3(1), 0, ctx->stream, ctx->d_status, ctx->d_regress,
                     ctx->d_cloneflag, ctx->ncell, ctx->wcfg.id_base, mc->d_status, mc->ncell,
                     ctx->d_clone_pairs);
  const double plane = w.dir > 0 ? w.zm + ctx->wcfg.clone_band * ctx->icfg.dx
                                 : w.z0 - ctx->wcfg.clone_band * ctx->icfg.dx;
  const size_t PNT = (size_t)ctx->mcfg.max_cells * ctx->nn, MNT = (size_t)mc->mcfg.max_cells * mc->nn;
  hipLaunchKernelGGL(k_clone_copy, dim3(nb((size_t)ctx->ncell * ctx->nn, 256)), dim3(256), 0,
                     ctx->stream, ctx->d_clone_pairs, ctx->ncell, ctx->nn, ctx->d_x, ctx->d_v, PNT,
                     mc->d_x, mc->d_v, MNT, w, mc->d_band, plane, mc->have_nuc ? mc->d_nuc_stale : nullptr);
  if (ctx->have_nuc && ctx->nnuc > 0) {
    const size_t PNTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc, MNTn = (size_t)mc->mcfg.max_cells * mc->nnuc;
    hipLaunchKernelGGL(k_clone_copy, dim3(nb((size_t)ctx->ncell * ctx->nnuc, 256)), dim3(256), 0,
                       ctx->stream, ctx->d_clone_pairs, ctx->ncell, ctx->nnuc, ctx->d_xn, ctx->d_vn,
                       PNTn, mc->d_xn, mc->d_vn, MNTn, w, nullptr, plane, nullptr);
  }
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int clone_advance(zf_ctx *mc) {
  if (mc->cell_parents.empty() || mc->ncell == 0) return ZF_OK;
  CloneSrcs src{};
  for (zf_ctx *p : mc->cell_parents) {
    if (src.n == 8) break;
    const size_t ncap = p->mcfg.max_cells;
    src.s[src.n++] = CloneSrc{p->d_v, p->d_vn, ncap * p->nn, p->have_nuc ? ncap * p->nnuc : 0,
                              p->wcfg.id_base, p->mcfg.max_cells};
  }
  const zf_cell_wrap_config &wc = mc->cell_parents.front()->wcfg;
  const size_t ncap = mc->mcfg.max_cells;
  hipLaunchKernelGGL(k_clone_advance, dim3(mc->ncell), dim3(256), 0, mc->stream, mc->d_x, mc->d_v,
                     ncap * mc->nn, mc->nn, mc->d_xn, mc->d_vn, mc->have_nuc ? ncap * mc->nnuc : 0,
                     mc->have_nuc ? mc->nnuc : 0, mc->d_status, mc->d_band, src,
                     mc->icfg.dt_advance, wc.roi_lo, wc.roi_hi);
  ZF_HIP(mc, hipGetLastError());
  return ZF_OK;
}

// The cell stages of a parent lattice up to spreading, in the reference's
// order for its parents: CAPS_MAX/MIN, membrane (on the unwrapped
// positions) and REGRESSFLAG, nucleus forces, cell list, interpolation
// with the wall repulsion, nucleus contact, cell-cell repulsion, spreading.
int wrap_rbc_update(zf_ctx *ctx, int iter) {
  int rc;
  if ((rc = wrap_prepare(ctx)) || (rc = membrane_forces_impl(ctx)) || (rc = wrap_regress(ctx)))
    return rc;
  if (ctx->have_nuc && (rc = nucleus_forces_impl(ctx))) return rc;
  if (ctx->have_rep && (rc = wrap_cell_list_impl(ctx))) return rc;
  if ((rc = wrap_interpolate_impl(ctx))) return rc;
  if (ctx->have_rep && (rc = wrap_wall_impl(ctx))) return rc;
  if (ctx->have_nuc && (rc = nucleus_contact_impl(ctx, iter))) return rc;
  if (ctx->have_rep && (rc = wrap_repulsion_impl(ctx))) return rc;
  return wrap_spread_impl(ctx);
}

}  // namespace zf

using namespace zf;

extern "C" {

void zf_cell_wrap_config_reference(zf_cell_wrap_config *wc, int32_t p) {
  *wc = zf_cell_wrap_config{};
  wc->direction = p == 1 ? 1 : -1;
  wc->id_base = p == 1 ? 0 : p == 2 ? 188 : 329;
  wc->split_span = 30.0;
  wc->clone_band = 5.0;
  wc->roi_lo = -86.0 + 0.5 * (121 - 1);  // ZM_1 = Z0_1 + (KM_1 - 1) DX
  wc->roi_hi = 332.5;                    // Z0_2
}

int zf_cell_wrap_setup(zf_ctx *parent, zf_ctx *main, const zf_cell_wrap_config *wc) {
  if (!parent || !main || !wc || parent == main || (wc->direction != 1 && wc->direction != -1) ||
      wc->id_base < 0 || !(wc->split_span > 0) || wc->clone_band < 0)
    return ZF_ERR_ARG;
  if (parent->link_main != main)
    return fail(parent, ZF_ERR_STATE, "link the parent to the main lattice first (zf_parent_link)");
  int rc = require_wrap_pair(parent, main);
  if (rc) return rc;
  if (main->cell_parents.size() >= 8 && !parent->have_wrap)
    return fail(parent, ZF_ERR_ARG, "at most 8 parents with cells per main lattice");
  if (parent->have_adj || main->have_adj)
    return fail(parent, ZF_ERR_STATE, "the adjoint does not cover cells on parent lattices");
  for (zf_ctx *q : main->cell_parents)
    if (q != parent && !(wc->id_base + (int64_t)parent->mcfg.max_cells <= q->wcfg.id_base ||
                         q->wcfg.id_base + (int64_t)q->mcfg.max_cells <= wc->id_base))
      return fail(parent, ZF_ERR_ARG, "the clone numbers of two parents overlap");
  ZF_HIP(parent, hipSetDevice(parent->device));
  ZF_HIP(parent, hipStreamSynchronize(parent->stream));
  free_wrap(parent);
  const size_t ncap = parent->mcfg.max_cells, NT = ncap * parent->nn;
  const size_t NTn = parent->have_nuc ? ncap * parent->nnuc : 0;
  if ((rc = dev_alloc(parent, (void **)&parent->d_regress, 4 * ncap)) ||
      (rc = dev_alloc(parent, (void **)&parent->d_cloneflag, 4 * ncap)) ||
      (rc = dev_alloc(parent, (void **)&parent->d_span, 16 * ncap)) ||
      (rc = dev_alloc(parent, (void **)&parent->d_split, 4 * ncap)) ||
      (rc = dev_alloc(parent, (void **)&parent->d_xu, 24 * NT)) ||
      (NTn && (rc = dev_alloc(parent, (void **)&parent->d_xnu, 24 * NTn))) ||
      (rc = dev_alloc(parent, (void **)&parent->d_clone_pairs, 4 * (2 * ncap + 1))))
    return rc;
  if (!main->d_band) {
    const size_t mcap = main->mcfg.max_cells;
    if ((rc = dev_alloc(main, (void **)&main->d_band, 16 * mcap))) return rc;
    ZF_HIP(main, hipMemset(main->d_band, 0, 16 * mcap));
  }
  // spreading keys: every point may carry one periodic image
  if ((rc = ensure_sort_keys(parent, 2 * (NT + NTn)))) return rc;
  hipLaunchKernelGGL(k_wrap_fill, dim3(nb(ncap, 256)), dim3(256), 0, parent->stream,
                     parent->d_regress, ncap, 0);
  hipLaunchKernelGGL(k_wrap_fill, dim3(nb(ncap, 256)), dim3(256), 0, parent->stream,
                     parent->d_cloneflag, ncap, 0);
  ZF_HIP(parent, hipGetLastError());
  ZF_HIP(parent, hipStreamSynchronize(parent->stream));
  if (parent->ncell > 0) {  // cells uploaded before: live ones start in status 2
    std::vector<int32_t> st(parent->ncell);
    ZF_HIP(parent, hipMemcpy(st.data(), parent->d_status, 4 * st.size(), hipMemcpyDeviceToHost));
    for (auto &s : st)
      if (s > kStatusDead) s = 2;
    ZF_HIP(parent, hipMemcpy(parent->d_status, st.data(), 4 * st.size(), hipMemcpyHostToDevice));
  }
  parent->wcfg = *wc;
  parent->wz0 = parent->icfg.z0;
  parent->wzm = parent->icfg.z0 + (parent->cfg.km - 1) * parent->icfg.dx;
  parent->wrap_main = main;
  parent->have_wrap = true;
  // parents clone in the order of their cell numbers
  auto &v = main->cell_parents;
  v.push_back(parent);
  std::sort(v.begin(), v.end(),
            [](const zf_ctx *a, const zf_ctx *b) { return a->wcfg.id_base < b->wcfg.id_base; });
  return ZF_OK;
}

int zf_step_linked(zf_ctx *mc, int32_t iter) {
  if (!mc || mc->parent_p) return ZF_ERR_ARG;
  const bool rbc_iter = mc->have_ibm && iter % mc->icfg.rbc_interval == 0;
  auto has_cells = [](const zf_ctx *c) { return c->have_ibm && c->ncell > 0; };
  int rc = zf_lbm_step(mc, iter, rbc_iter && has_cells(mc) ? 1 : 0);
  if (rc) return rc;
  for (zf_ctx *p : mc->linked)
    if ((rc = zf_lbm_step(p, iter, rbc_iter && p->have_wrap && has_cells(p) ? 1 : 0))) return rc;
  if (!rbc_iter) return ZF_OK;
  for (zf_ctx *p : mc->cell_parents)
    if (has_cells(p) && (rc = wrap_rbc_update(p, iter))) return rc;
  if (has_cells(mc)) {
    if ((rc = rbc_update(mc, false, iter)) || (rc = advance_impl(mc, nullptr)) ||
        (rc = clone_advance(mc)))
      return rc;
  }
  for (zf_ctx *p : mc->cell_parents)
    if (has_cells(p) && (rc = wrap_advance(p))) return rc;
  for (zf_ctx *p : mc->cell_parents)
    if (has_cells(p) && (rc = wrap_clone(p))) return rc;
  return ZF_OK;
}

int zf_download_cell_flags(zf_ctx *ctx, int32_t *regress, int32_t *cloneflag) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_wrap) return fail(ctx, ZF_ERR_STATE, "zf_cell_wrap_setup not called");
  if (ctx->ncell == 0) return ZF_OK;
  if (regress)
    ZF_HIP(ctx, hipMemcpyAsync(regress, ctx->d_regress, 4 * (size_t)ctx->ncell, hipMemcpyDeviceToHost,
                               ctx->stream));
  if (cloneflag)
    ZF_HIP(ctx, hipMemcpyAsync(cloneflag, ctx->d_cloneflag, 4 * (size_t)ctx->ncell,
                               hipMemcpyDeviceToHost, ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ZF_OK;
}

int zf_upload_cell_flags(zf_ctx *ctx, const int32_t *regress) {
  if (!ctx || !regress) return ZF_ERR_ARG;
  if (!ctx->have_wrap) return fail(ctx, ZF_ERR_STATE, "zf_cell_wrap_setup not called");
  if (ctx->ncell == 0) return ZF_OK;
  ZF_HIP(ctx, hipMemcpyAsync(ctx->d_regress, regress, 4 * (size_t)ctx->ncell, hipMemcpyHostToDevice,
                             ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ZF_OK;
}

}  // extern "C"
