// zf_parent.hip -- parent (feeder) lattices and their boundary coupling to
// the main lattice.
//
// The reference runs three parent lattices beside the main one
// (ZebraflowWT.c:91-112, set up by LBM_PARENTDOMAIN_SETUP :515 and
// LBM_PARENTBC_SETUP :1015): straight extrusions of a vessel cross-section
// whose overlap plane (k = KM for parent 1, k = 1 for parents 2 and 3) lies
// on the main lattice.  A parent's interior nodes take the ordinary
// stream + collide (:6148-6286); its boundary pass (:6436-6594) sets
//   B == 0      feq(D_p[m], 0)            wall
//   B == 20+p   feq(D_p[m], U[o])         upstream, velocity from the main
//                                          node o over the parent column
//   B == 10+p   feq(D[v], U[v])           copied from the main node v
// with D, U of the main lattice as its stream + collide of the same step
// left them (density and force-shifted velocity at interior nodes; the last
// macro recompute elsewhere).
//
// Here a parent is a context of its own (zf_set_parent before its lattice
// is uploaded: parent thresholds in the BC list, main-domain indices for the
// coupled codes), stepped like any other; its walls are the stream kernel's.
// zf_parent_link gives the main context an export list: after each of its
// steps a small kernel writes the moments of those nodes into a buffer, and
// the parent's step ends with a kernel that writes feq of them into its
// coupled boundary nodes.  The moments come from the stored populations:
// after a BGK-Guo collision sum f' = rho and sum f' c = rho u_shift + DT F/2,
// so an interior node's (rho, u_shift) is recovered from its post-collision
// populations and the force the step used; a node the main stream kernel
// does not update keeps its previous populations, whose moments are its
// last macro recompute.  The results equal the reference's to rounding.
// While linked, a parent issues its work on the main context's stream, so
// the one stream orders main step, export, parent step and coupling pass.
// Separate streams ordered by events were measured slower: each event wait
// cost more than the overlap it allowed (the parent's kernels are short
// next to the main step).  Call order per iteration: the main lattice's
// step, then its parents' steps.
#include "zf_d3q19.h"

#include <algorithm>
#include <vector>

namespace zf {
namespace {

using namespace d3q19;

// Moments of node s from storage buffer f: (rho, u) with u = (sum g c - shift
// * F) / rho, g the stored deviations f - w (their weights sum to one and
// have no first moment).
template <int S>
__device__ void node_moments(const typename Store<S>::T *f, size_t stride, int s, double shift,
                             double fx, double fy, double fz, double *out) {
  double dr = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int c = 0; c < Q; c++) {
    const double g = (double)Store<S>::ld(f, (size_t)c * stride + s, WQ[c]);
    dr += g;
    mx += g * EX[c];
    my += g * EY[c];
    mz += g * EZ[c];
  }
  const double rho = 1.0 + dr;
  out[0] = rho;
  out[1] = (mx - shift * fx) / rho;
  out[2] = (my - shift * fy) / rho;
  out[3] = (mz - shift * fz) / rho;
}

// Main lattice: one lane per exported node.  ent = storage index << 1 |
// interior; interior nodes are read from the step's output with the force
// shift, the others from its input without (every node, with `plain`).
template <int S>
__global__ void k_export(const typename Store<S>::T *__restrict__ fnew,
                         const typename Store<S>::T *__restrict__ fold, size_t stride,
                         const int32_t *__restrict__ ent, int n, int plain,
                         const typename Store<S>::F *__restrict__ field, double g0, double g1,
                         double g2, double half_dt, double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int s = ent[e] >> 1;
  const bool inner = (ent[e] & 1) && !plain;
  double fx = g0, fy = g1, fz = g2;
  if (field) {
    fx += (double)field[s];
    fy += (double)field[stride + s];
    fz += (double)field[2 * stride + s];
  }
  double m[4];
  if (inner) node_moments<S>(fnew, stride, s, half_dt, fx, fy, fz, m);
  else node_moments<S>(fold, stride, s, 0.0, 0, 0, 0, m);
  out[e] = m[0];
  out[n + e] = m[1];
  out[2 * (size_t)n + e] = m[2];
  out[3 * (size_t)n + e] = m[3];
}

// Parent lattice: one lane per coupled boundary node; writes feq of its
// moments into the step's output.
template <int S>
__global__ void k_parent_bc(typename Store<S>::T *__restrict__ fnew,
                            const typename Store<S>::T *__restrict__ fold, size_t stride,
                            const ParentBc *__restrict__ pbc, int n,
                            const double *__restrict__ mom, int nexp) {
  using R = typename Store<S>::R;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const ParentBc b = pbc[e];
  double rho = mom[b.slot];
  const double u = mom[nexp + b.slot], v = mom[2 * (size_t)nexp + b.slot],
               w = mom[3 * (size_t)nexp + b.slot];
  if (b.kind < 2) {
    // upstream: the parent's own density at its neighbour, from this step
    // (interior neighbour) or the last one
    const typename Store<S>::T *src = b.kind == 0 ? fnew : fold;
    double dr = 0;
#pragma unroll
    for (int c = 0; c < Q; c++) dr += (double)Store<S>::ld(src, (size_t)c * stride + b.sm, WQ[c]);
    rho = 1.0 + dr;
  }
  const double dr = rho - 1.0, usq = u * u + v * v + w * w;
#pragma unroll
  for (int c = 0; c < Q; c++)
    fnew[(size_t)c * stride + b.s] = Store<S>::st((R)geq<double>(c, dr, u, v, w, usq), WQ[c]);
}

// Storage index of the main-lattice node at each (i, j, k) triple (1-based,
// inside the box), or -1 where it is unlisted.
__global__ void k_main_index(const int32_t *__restrict__ ijk, int n, const int32_t *__restrict__ domid,
                             const int32_t *__restrict__ s_of_p, int jm, int km,
                             int32_t *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const size_t q = ((size_t)(ijk[3 * e] - 1) * jm + (ijk[3 * e + 1] - 1)) * km + (ijk[3 * e + 2] - 1);
  const int d = domid[q];
  out[e] = d > 0 ? s_of_p[d - 1] : -1;
}

template <class F>
int by_store(int storage, F &&f) {
  switch (storage) {
    case ZF_STORE_F64: return f(std::integral_constant<int, ZF_STORE_F64>());
    case ZF_STORE_F32DEV: return f(std::integral_constant<int, ZF_STORE_F32DEV>());
    default: return f(std::integral_constant<int, ZF_STORE_F16DEV>());
  }
}

int nblk(int n) { return (n + 255) / 256; }
// The export and coupling passes cover a few thousand nodes, each lane a
// gather of 19 scattered population words: one wave per workgroup spreads
// them over as many CUs as possible.
constexpr int kGatherLanes = 64;
int nwg(int n) { return (n + kGatherLanes - 1) / kGatherLanes; }

}  // namespace

int parent_export(zf_ctx *ctx, bool in_step) {
  const int n = (int)ctx->h_export.size();
  if (n == 0) return ZF_OK;
  const zf_config &c = ctx->cfg;
  by_store(c.storage, [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    using T = typename Store<S>::T;
    // inside a step (before the buffer swap) d_f[cur] is its input and
    // d_f[cur ^ 1] its output; outside one, d_f[cur] is the state
    hipLaunchKernelGGL(k_export<S>, dim3(nwg(n)), dim3(kGatherLanes), 0, ctx->stream,
                       static_cast<const T *>(ctx->d_f[ctx->cur ^ 1]),
                       static_cast<const T *>(ctx->d_f[ctx->cur]), ctx->stride, ctx->d_export_s, n,
                       in_step ? 0 : 1,
                       static_cast<const typename Store<S>::F *>(ctx->d_force), c.body_force[0],
                       c.body_force[1], c.body_force[2], 0.5 * c.force_dt, ctx->d_export);
    return 0;
  });
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int parent_couple(zf_ctx *ctx) {
  zf_ctx *mc = ctx->link_main;
  if (!mc || ctx->n_pbc == 0) return ZF_OK;
  const int nexp = (int)mc->h_export.size();
  by_store(ctx->cfg.storage, [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    using T = typename Store<S>::T;
    hipLaunchKernelGGL(k_parent_bc<S>, dim3(nwg(ctx->n_pbc)), dim3(kGatherLanes), 0, ctx->stream,
                       static_cast<T *>(ctx->d_f[ctx->cur]),
                       static_cast<const T *>(ctx->d_f[ctx->cur ^ 1]), ctx->stride, ctx->d_pbc,
                       ctx->n_pbc, mc->d_export, nexp);
    return 0;
  });
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

// Removes the main context's export (its slots are rebuilt from the
// parents still linked).
static int rebuild_export(zf_ctx *mc);

void free_parent(zf_ctx *ctx) {
  free_wrap(ctx);
  if (ctx->link_main) {
    zf_ctx *mc = ctx->link_main;
    (void)hipStreamSynchronize(mc->stream);
    ctx->stream = ctx->own_stream;  // back to its own stream
    ctx->own_stream = nullptr;
    mc->linked.erase(std::remove(mc->linked.begin(), mc->linked.end(), ctx), mc->linked.end());
    ctx->link_main = nullptr;
    (void)rebuild_export(mc);
  }
  dev_free(ctx, (void **)&ctx->d_pbc);
  ctx->n_pbc = 0;
  // a main lattice: unlink its parents (they keep their lattices)
  if (!ctx->linked.empty()) (void)hipStreamSynchronize(ctx->stream);
  for (zf_ctx *p : std::vector<zf_ctx *>(ctx->linked)) {
    p->stream = p->own_stream;
    p->own_stream = nullptr;
    p->link_main = nullptr;
    dev_free(p, (void **)&p->d_pbc);
    p->n_pbc = 0;
  }
  ctx->linked.clear();
  ctx->h_export.clear();
  dev_free(ctx, (void **)&ctx->d_export_s);
  dev_free(ctx, (void **)&ctx->d_export);
}

static int download_ints(zf_ctx *ctx, std::vector<int32_t> &h, const int32_t *d, size_t n) {
  h.resize(n);
  if (n) ZF_HIP(ctx, hipMemcpy(h.data(), d, 4 * n, hipMemcpyDeviceToHost));
  return ZF_OK;
}

// The parent's coupling table and the main's export slots.  Slots are shared
// by the parents of one main lattice; each parent's table names its slots.
static int link_tables(zf_ctx *pc, zf_ctx *mc, std::vector<int32_t> &slot_of_main_s) {
  const int nb = pc->bcmax, p = pc->parent_p;
  std::vector<int32_t> bcl, bcn, vb, pdomid, psop;
  int rc;
  if ((rc = download_ints(pc, bcl, pc->d_bclist, 4 * (size_t)nb))) return rc;
  if ((rc = download_ints(pc, bcn, pc->d_bcnlist, 3 * (size_t)nb))) return rc;
  if ((rc = download_ints(pc, vb, pc->d_vbcn, 3 * (size_t)nb))) return rc;
  if ((rc = download_ints(pc, pdomid, pc->d_domid, pc->box))) return rc;
  if ((rc = download_ints(pc, psop, pc->d_s_of_p, pc->dommax))) return rc;
  const zf_config &P = pc->cfg, &M = mc->cfg;
  auto pq = [&](int i, int j, int k) { return ((size_t)(i - 1) * P.jm + (j - 1)) * P.km + (k - 1); };
  // coupled entries and the main node each one reads
  std::vector<ParentBc> tab;
  std::vector<int32_t> ijk, row;
  for (int l = 0; l < nb; l++) {
    const int code = bcl[4 * l + 3];
    if (code != 10 + p && code != 20 + p) continue;
    if (bcn[3 * l] == 0) continue;  // no neighbour: left alone, as the reference's stale entry
    ParentBc e{};
    e.s = psop[pdomid[pq(bcl[4 * l], bcl[4 * l + 1], bcl[4 * l + 2])] - 1];
    const int32_t *m3 = code == 20 + p ? &vb[3 * l] : &bcn[3 * l];
    if (code == 20 + p) {
      e.sm = psop[pdomid[pq(bcn[3 * l], bcn[3 * l + 1], bcn[3 * l + 2])] - 1];
      e.kind = e.sm < pc->n_int ? 0 : 1;
    } else {
      e.sm = -1;
      e.kind = 2;
    }
    if (m3[0] < 1 || m3[0] > M.im || m3[1] < 1 || m3[1] > M.jm || m3[2] < 1 || m3[2] > M.km)
      return fail(pc, ZF_ERR_GEOMETRY, "parent BC entry %d maps to (%d,%d,%d), outside the main box",
                  l + 1, m3[0], m3[1], m3[2]);
    ijk.insert(ijk.end(), m3, m3 + 3);
    row.push_back(l);
    tab.push_back(e);
  }
  // their main storage indices, gathered on the device
  const int n = (int)tab.size();
  std::vector<int32_t> ms(n);
  if (n) {
    int32_t *d_ijk = nullptr, *d_out = nullptr;
    if ((rc = dev_alloc(pc, (void **)&d_ijk, 12 * (size_t)n))) return rc;
    if ((rc = dev_alloc(pc, (void **)&d_out, 4 * (size_t)n))) {
      dev_free(pc, (void **)&d_ijk);
      return rc;
    }
    hipError_t err = hipMemcpy(d_ijk, ijk.data(), 12 * (size_t)n, hipMemcpyHostToDevice);
    if (err == hipSuccess) {
      hipLaunchKernelGGL(k_main_index, dim3(nblk(n)), dim3(256), 0, mc->stream, d_ijk, n, mc->d_domid,
                         mc->d_s_of_p, M.jm, M.km, d_out);
      err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(mc->stream);
    if (err == hipSuccess) err = hipMemcpy(ms.data(), d_out, 4 * (size_t)n, hipMemcpyDeviceToHost);
    dev_free(pc, (void **)&d_ijk);
    dev_free(pc, (void **)&d_out);
    if (err != hipSuccess)
      return fail(pc, ZF_ERR_HIP, "parent link gather failed: %s", hipGetErrorString(err));
  }
  for (int e = 0; e < n; e++) {
    if (ms[e] < 0) {
      const int32_t *m3 = &ijk[3 * (size_t)e];
      return fail(pc, ZF_ERR_GEOMETRY, "parent BC entry %d maps to (%d,%d,%d), not a listed main node",
                  row[e] + 1, m3[0], m3[1], m3[2]);
    }
    if (slot_of_main_s[ms[e]] < 0) {
      slot_of_main_s[ms[e]] = (int32_t)mc->h_export.size();
      mc->h_export.push_back(ms[e] << 1 | (ms[e] < mc->n_int ? 1 : 0));
    }
    tab[e].slot = slot_of_main_s[ms[e]];
  }
  dev_free(pc, (void **)&pc->d_pbc);
  pc->n_pbc = (int32_t)tab.size();
  if (!tab.empty()) {
    if ((rc = dev_alloc(pc, (void **)&pc->d_pbc, sizeof(ParentBc) * tab.size()))) return rc;
    ZF_HIP(pc, hipMemcpy(pc->d_pbc, tab.data(), sizeof(ParentBc) * tab.size(), hipMemcpyHostToDevice));
  }
  return ZF_OK;
}

static int rebuild_export(zf_ctx *mc) {
  (void)hipStreamSynchronize(mc->stream);
  mc->h_export.clear();
  dev_free(mc, (void **)&mc->d_export_s);
  dev_free(mc, (void **)&mc->d_export);
  std::vector<int32_t> slot(mc->stride, -1);
  for (zf_ctx *p : mc->linked) {
    const int rc = link_tables(p, mc, slot);
    if (rc) {
      mc->err = p->err;
      return rc;
    }
  }
  const size_t n = mc->h_export.size();
  if (n == 0) return ZF_OK;
  int rc;
  if ((rc = dev_alloc(mc, (void **)&mc->d_export_s, sizeof(int32_t) * n))) return rc;
  if ((rc = dev_alloc(mc, (void **)&mc->d_export, sizeof(double) * 4 * n))) return rc;
  ZF_HIP(mc, hipMemcpy(mc->d_export_s, mc->h_export.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  // the moments of the current state, so that a parent stepped before its
  // main lattice's next step reads defined values
  return parent_export(mc, false);
}

}  // namespace zf

using namespace zf;

extern "C" {

int zf_set_parent(zf_ctx *ctx, int32_t p, int32_t k_over, const int32_t offset[3]) {
  if (!ctx || p < 0 || p > 3 || (p && !offset)) return ZF_ERR_ARG;
  if (ctx->have_lists) return fail(ctx, ZF_ERR_STATE, "zf_set_parent must precede zf_upload_lattice");
  if (p && (k_over < 1 || k_over > ctx->cfg.km))
    return fail(ctx, ZF_ERR_ARG, "overlap plane %d outside 1..%d", k_over, ctx->cfg.km);
  if (p && (ctx->cfg.periodic[0] || ctx->cfg.periodic[1] || ctx->cfg.periodic[2]))
    return fail(ctx, ZF_ERR_ARG, "a parent lattice is not periodic");
  ctx->parent_p = p;
  ctx->parent_kover = p ? k_over : 0;
  for (int a = 0; a < 3; a++) ctx->parent_off[a] = p ? offset[a] : 0;
  return ZF_OK;
}

int zf_download_parent_lists(zf_ctx *ctx, int32_t *vbcn) {
  if (!ctx || !vbcn) return ZF_ERR_ARG;
  if (!ctx->have_lists || !ctx->parent_p)
    return fail(ctx, ZF_ERR_STATE, "not a parent lattice with an uploaded lattice");
  if (ctx->bcmax)
    ZF_HIP(ctx, hipMemcpy(vbcn, ctx->d_vbcn, 12 * (size_t)ctx->bcmax, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_parent_link(zf_ctx *parent, zf_ctx *main) {
  if (!parent || !main || parent == main) return ZF_ERR_ARG;
  if (!parent->parent_p || !parent->have_lattice)
    return fail(parent, ZF_ERR_STATE, "zf_parent_link needs a parent lattice (zf_set_parent) uploaded");
  if (main->parent_p || !main->have_lattice)
    return fail(parent, ZF_ERR_STATE, "zf_parent_link needs a main lattice uploaded");
  if (main->device != parent->device) return fail(parent, ZF_ERR_ARG, "contexts on different devices");
  if (main->linked.size() >= 8) return fail(parent, ZF_ERR_ARG, "at most 8 parents per main lattice");
  if (main->have_adj || parent->have_adj)
    return fail(parent, ZF_ERR_STATE, "the adjoint does not cover parent-lattice coupling");
  free_parent(parent);
  ZF_HIP(parent, hipSetDevice(parent->device));
  ZF_HIP(parent, hipStreamSynchronize(parent->stream));
  parent->link_main = main;
  main->linked.push_back(parent);
  int rc = rebuild_export(main);
  if (rc) {
    main->linked.pop_back();
    parent->link_main = nullptr;
    (void)rebuild_export(main);
    if (parent->err.empty()) parent->err = main->err;
    return rc;
  }
  // the parent's work goes to the main lattice's stream while linked
  parent->own_stream = parent->stream;
  parent->stream = main->stream;
  return ZF_OK;
}

}  // extern "C"
