// zf_lists.hip -- index lists on the device and the stream kernel's run-length
// neighbour form.
//
// Lists built here (bit-exact with the reference):
//   DOMLIST / DOMID   LBM_DOM_SETUP     ZebraflowWT.c:1481-1502
//   BCLIST / BCNLIST  LBM_BC_SETUP      ZebraflowWT.c:1506-1589
//   pull neighbours   (implicit)        ZebraflowWT.c:6113-6125
// Both box lists are a stream compaction of the box in lexicographic order
// (i outer, k inner); it runs as tile count -> tile scan -> ordered scatter,
// with wave ballots giving each lane its rank inside a 64-node wave.
#include "zf_internal.h"

#include <algorithm>
#include <cstring>

namespace zf {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;  // per thread -> 4096 box nodes per tile
constexpr int kTile = kScanThreads * kScanItems;

struct Box {
  int im, jm, km;
  int per[3];
};

// 0: DOMLIST (B >= 0); 1: BCLIST of a main lattice (B == 0 or B >= 1000,
// ZebraflowWT.c:1515); 2: BCLIST of a parent lattice (B == 0 or B >= 10,
// :1045).
__device__ __forceinline__ bool pred(int mode, int b) {
  return mode == 0 ? (b >= 0) : (b == 0 || b >= (mode == 1 ? 1000 : 10));
}

// Parent lattice p (0: main lattice): overlap plane and index offset of the
// main-domain node over a parent column.
struct ParentMap {
  int p, kover, off[3];
};

__device__ __forceinline__ void unflatten(size_t q, int jm, int km, int &i, int &j, int &k) {
  k = (int)(q % (size_t)km) + 1;
  size_t r = q / (size_t)km;
  j = (int)(r % (size_t)jm) + 1;
  i = (int)(r / (size_t)jm) + 1;
}

__device__ __forceinline__ size_t flatten(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

// Step from (i,j,k) by sign*e_c with periodic wrap; false if outside a
// non-periodic face (the reference's `continue`, ZebraflowWT.c:1541).
__device__ __forceinline__ bool step_to(const Box &bx, int i, int j, int k, int c, int sign,
                                        int &oi, int &oj, int &ok) {
  const int di[18] = {1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0};
  const int dj[18] = {0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
  const int dk[18] = {0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
  int ii = i + sign * di[c], jj = j + sign * dj[c], kk = k + sign * dk[c];
  if (bx.per[0]) { if (ii < 1) ii += bx.im; else if (ii > bx.im) ii -= bx.im; }
  if (bx.per[1]) { if (jj < 1) jj += bx.jm; else if (jj > bx.jm) jj -= bx.jm; }
  if (bx.per[2]) { if (kk < 1) kk += bx.km; else if (kk > bx.km) kk -= bx.km; }
  if (ii < 1 || ii > bx.im || jj < 1 || jj > bx.jm || kk < 1 || kk > bx.km) return false;
  oi = ii; oj = jj; ok = kk;
  return true;
}

__global__ __launch_bounds__(kScanThreads) void k_tile_count(const int32_t *__restrict__ b,
                                                             size_t n, int mode,
                                                             int32_t *__restrict__ counts) {
  __shared__ int wsum[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kTile;
  int cnt = 0;
  for (int it = 0; it < kScanItems; it++) {
    size_t q = base + (size_t)it * kScanThreads + threadIdx.x;
    bool f = q < n && pred(mode, b[q]);
    cnt += f ? 1 : 0;
  }
  // wave reduction through ballots is not needed for a sum of small ints:
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kScanThreads / 64; w++) t += wsum[w];
    counts[blockIdx.x] = t;
  }
}

// Exclusive scan of the tile counts in one workgroup; total -> counts[ntiles].
__global__ __launch_bounds__(1024) void k_scan_counts(int32_t *__restrict__ counts, int ntiles) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (ntiles + 1023) / 1024;
  const int lo = min(ntiles, t * per), hi = min(ntiles, lo + per);
  int s = 0;
  for (int x = lo; x < hi; x++) s += counts[x];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int w = 0; w < 1024; w++) {
      int v = part[w];
      part[w] = run;
      run += v;
    }
    counts[ntiles] = run;
  }
  __syncthreads();
  int run = part[t];
  for (int x = lo; x < hi; x++) {
    int v = counts[x];
    counts[x] = run;
    run += v;
  }
}

// Ordered scatter: lane rank = popcount of the ballot below the lane, wave
// offsets within an iteration from LDS, iterations in index order.
__global__ __launch_bounds__(kScanThreads) void k_tile_scatter(
    const int32_t *__restrict__ b, size_t n, int mode, int jm, int km,
    const int32_t *__restrict__ offsets, int32_t *__restrict__ domid,
    int32_t *__restrict__ list) {
  __shared__ int wcnt[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kTile;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int run = offsets[blockIdx.x];
  for (int it = 0; it < kScanItems; it++) {
    size_t q = base + (size_t)it * kScanThreads + threadIdx.x;
    int bv = q < n ? b[q] : -1;
    bool f = q < n && pred(mode, bv);
    uint64_t m = __ballot(f);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < kScanThreads / 64; w++) {
      if (w < wave) woff += wcnt[w];
      tot += wcnt[w];
    }
    const uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
    const int pos = run + woff + __popcll(below);
    if (q < n) {
      int i, j, k;
      unflatten(q, jm, km, i, j, k);
      if (mode == 0) {
        domid[q] = f ? pos + 1 : -1;
        if (f) {
          list[3 * (size_t)pos + 0] = i;
          list[3 * (size_t)pos + 1] = j;
          list[3 * (size_t)pos + 2] = k;
        }
      } else if (f) {
        list[4 * (size_t)pos + 0] = i;
        list[4 * (size_t)pos + 1] = j;
        list[4 * (size_t)pos + 2] = k;
        list[4 * (size_t)pos + 3] = bv;
      }
    }
    run += tot;
    __syncthreads();
  }
}

// BCNLIST: three tiers, each scanning c = 1..18 in order
// (ZebraflowWT.c:1535-1574).  No neighbour -> (0,0,0).  On a parent lattice
// (LBM_PARENTBC_SETUP, :1053-1122) the tiers use its thresholds (0<B<10,
// B>10, B==0) and run for B == 0 and the upstream code 20+p only; the copy
// code 10+p takes the main-domain node over its column at the overlap plane,
// and an upstream node's velocity source (vbcn) is that same node.
__global__ void k_bc_neighbour(const int32_t *__restrict__ b, Box bx, int bcmax,
                               const int32_t *__restrict__ bclist, int32_t *__restrict__ bcnlist,
                               ParentMap pm, int32_t *__restrict__ vbcn) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= bcmax) return;
  const int i = bclist[4 * t], j = bclist[4 * t + 1], k = bclist[4 * t + 2], code = bclist[4 * t + 3];
  const int lim = pm.p ? 10 : 1000;
  int fi = 0, fj = 0, fk = 0;
  bool found = false;
  const bool search = !pm.p || code == 0 || code == 20 + pm.p;
  for (int tier = 0; search && tier < 3 && !found; tier++) {
    for (int c = 0; c < 18; c++) {
      int ii, jj, kk;
      if (!step_to(bx, i, j, k, c, +1, ii, jj, kk)) continue;
      const int nb = b[flatten(ii, jj, kk, bx.jm, bx.km)];
      const bool ok = tier == 0 ? (nb > 0 && nb < lim) : tier == 1 ? (nb > lim) : (nb == 0);
      if (ok) { fi = ii; fj = jj; fk = kk; found = true; break; }
    }
  }
  if (pm.p && code == 10 + pm.p) {
    fi = i + pm.off[0];
    fj = j + pm.off[1];
    fk = pm.kover + pm.off[2];
  }
  bcnlist[3 * t] = fi;
  bcnlist[3 * t + 1] = fj;
  bcnlist[3 * t + 2] = fk;
  if (pm.p) {
    const bool up = code == 20 + pm.p;
    vbcn[3 * t] = up ? i + pm.off[0] : 0;
    vbcn[3 * t + 1] = up ? j + pm.off[1] : 0;
    vbcn[3 * t + 2] = up ? pm.kover + pm.off[2] : 0;
  }
}

// Pull sources: DOMID of (i,j,k) - e_c (ZebraflowWT.c:6114-6119).
// Storage index of the listed node at each box position, or -1.
__global__ void k_sbox(const int32_t *__restrict__ domid, size_t box, const int32_t *__restrict__ s_of_p,
                       int32_t *__restrict__ sbox) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= box) return;
  const int d = domid[q];
  sbox[q] = d > 0 ? s_of_p[d - 1] : -1;
}

__global__ void k_pull_neighbours(const int32_t *__restrict__ domid, Box bx, int dommax,
                                  const int32_t *__restrict__ domlist, int32_t *__restrict__ nbr) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dommax) return;
  const int i = domlist[3 * p], j = domlist[3 * p + 1], k = domlist[3 * p + 2];
  for (int c = 0; c < 18; c++) {
    int ii, jj, kk, v = -1;
    if (step_to(bx, i, j, k, c, -1, ii, jj, kk)) v = domid[flatten(ii, jj, kk, bx.jm, bx.km)];
    nbr[(size_t)c * dommax + p] = v;
  }
}

int compact(zf_ctx *ctx, int mode, int32_t *d_counts, int ntiles, int32_t *list, int32_t *count) {
  const Box bx{ctx->cfg.im, ctx->cfg.jm, ctx->cfg.km, {0, 0, 0}};
  hipLaunchKernelGGL(k_tile_count, dim3(ntiles), dim3(kScanThreads), 0, ctx->stream, ctx->d_b,
                     ctx->box, mode, d_counts);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, ctx->stream, d_counts, ntiles);
  ZF_HIP(ctx, hipGetLastError());
  if (list) {
    hipLaunchKernelGGL(k_tile_scatter, dim3(ntiles), dim3(kScanThreads), 0, ctx->stream,
                       ctx->d_b, ctx->box, mode, bx.jm, bx.km, d_counts, ctx->d_domid, list);
    ZF_HIP(ctx, hipGetLastError());
  }
  ZF_HIP(ctx, hipMemcpyAsync(count, d_counts + ntiles, sizeof(int32_t), hipMemcpyDeviceToHost,
                             ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ZF_OK;
}

}  // namespace

void free_lattice(zf_ctx *ctx) {
  free_pipe(ctx);
  void **bufs[] = {(void **)&ctx->d_b,        (void **)&ctx->d_domid,  (void **)&ctx->d_domlist,
                   (void **)&ctx->d_bclist,   (void **)&ctx->d_bcnlist, (void **)&ctx->d_nbr,
                   (void **)&ctx->d_s_of_p,   (void **)&ctx->d_p_of_s, (void **)&ctx->d_word,
                   (void **)&ctx->d_block_chord0, (void **)&ctx->d_chords, (void **)&ctx->d_bc,
                   (void **)&ctx->d_fused_s, (void **)&ctx->d_att_start, (void **)&ctx->d_sbox,
                   (void **)&ctx->d_vbcn};
  for (auto b : bufs) dev_free(ctx, b);
  ctx->have_lists = false;
  ctx->have_lattice = false;
  ctx->dommax = ctx->bcmax = ctx->n_int = ctx->n_chords = ctx->n_bc = 0;
}

// Storage order, node words, chords and boundary entries (host side).
static int build_kernel_tables(zf_ctx *ctx, const int32_t *b, const std::vector<int32_t> &domlist,
                               const std::vector<int32_t> &nbr, const std::vector<int32_t> &bclist,
                               const std::vector<int32_t> &bcnlist) {
  const int N = ctx->dommax, jm = ctx->cfg.jm, km = ctx->cfg.km;
  const bool bb = ctx->cfg.wall_mode == ZF_WALL_HALFWAY_BB;
  auto boxq = [&](int i, int j, int k) {
    return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
  };
  std::vector<int32_t> &bnode = ctx->h_bnode;
  bnode.assign(N, 0);
  for (int p = 0; p < N; p++)
    bnode[p] = b[boxq(domlist[3 * p], domlist[3 * p + 1], domlist[3 * p + 2])];
  auto interior = [&](int p) { return bnode[p] > 0 && bnode[p] < 10; };

  // storage: interior nodes first (list order), then all others (list order)
  std::vector<int32_t> s_of_p(N), p_of_s(N);
  int nint = 0;
  for (int p = 0; p < N; p++)
    if (interior(p)) s_of_p[p] = nint++;
  int nt = nint;
  for (int p = 0; p < N; p++)
    if (!interior(p)) s_of_p[p] = nt++;
  for (int p = 0; p < N; p++) p_of_s[s_of_p[p]] = p;
  ctx->n_int = nint;

  // node words and chords
  std::vector<uint32_t> word(nint, 0);
  std::vector<int32_t> chords;
  std::vector<int32_t> block0((nint + kBlockNodes - 1) / kBlockNodes + 1, 0);
  std::vector<int32_t> node_chord(nint, 0);
  int nbb = 0;
  int cur_def = 0;  // bitmask of defined deltas in the open chord
  int32_t cur[kChordInts];
  int nch = 0;
  for (int s = 0; s < nint; s++) {
    const int p = p_of_s[s];
    uint32_t mask = 0;
    int32_t d[18];
    for (int c = 0; c < 18; c++) {
      const int src = nbr[(size_t)c * N + p];
      if (src <= 0) {
        const int *x = &domlist[3 * p];
        return fail(ctx, ZF_ERR_GEOMETRY,
                    "interior node (%d,%d,%d) pulls direction %d from a node outside the list "
                    "(B<0 or outside the box)", x[0], x[1], x[2], c + 1);
      }
      if (bb && bnode[src - 1] == 0) {
        mask |= 1u << c;
        d[c] = 0;
        nbb++;
      } else {
        d[c] = s_of_p[src - 1] - s;
      }
    }
    bool join = nch > 0 && cur[18] == bnode[p];
    if (join) {
      for (int c = 0; c < 18 && join; c++)
        if (!(mask >> c & 1) && (cur_def >> c & 1) && cur[c] != d[c]) join = false;
    }
    if (!join) {
      if (nch > 0) chords.insert(chords.end(), cur, cur + kChordInts);
      std::memset(cur, 0, sizeof cur);
      cur[18] = bnode[p];
      cur_def = 0;
      nch++;
    }
    for (int c = 0; c < 18; c++)
      if (!(mask >> c & 1) && !(cur_def >> c & 1)) {
        cur[c] = d[c];
        cur_def |= 1 << c;
      }
    node_chord[s] = nch - 1;
    word[s] = mask;
  }
  if (nch > 0) chords.insert(chords.end(), cur, cur + kChordInts);
  for (int blk = 0; blk * kBlockNodes < nint; blk++) block0[blk] = node_chord[blk * kBlockNodes];
  for (int s = 0; s < nint; s++) {
    const uint32_t rel = (uint32_t)(node_chord[s] - block0[s / kBlockNodes]);
    word[s] |= rel << kWordChordShift;
  }
  ctx->n_chords = nch;
  ctx->n_bb_links = nbb;

  // boundary entries
  std::vector<BcEntry> bce;
  int nmacro = 0;
  for (int t = 0; t < ctx->bcmax; t++) {
    const int code = bclist[4 * t + 3];
    const bool active = (code == 0 && !bb) || code == 4998 || (code >= 1001 && code <= 1009) ||
                        code == ZF_B_VELOCITY_INLET || code == ZF_B_PRESSURE_OUTLET;
    if (!active || bcnlist[3 * t] == 0) continue;
    const size_t qn = boxq(bclist[4 * t], bclist[4 * t + 1], bclist[4 * t + 2]);
    const size_t qm = boxq(bcnlist[3 * t], bcnlist[3 * t + 1], bcnlist[3 * t + 2]);
    (void)qn;
    // DOMID of BC node and its neighbour via the list order
    BcEntry e{};
    int pn = -1, pm = -1;
    // binary search in DOMLIST (lexicographic) for both nodes
    auto find = [&](int i, int j, int k) {
      int lo = 0, hi = N - 1;
      while (lo <= hi) {
        int mid = (lo + hi) / 2;
        const int *x = &domlist[3 * mid];
        int cmp = x[0] != i ? (x[0] < i ? -1 : 1) : x[1] != j ? (x[1] < j ? -1 : 1)
                                                              : x[2] != k ? (x[2] < k ? -1 : 1) : 0;
        if (cmp == 0) return mid;
        if (cmp < 0) lo = mid + 1; else hi = mid - 1;
      }
      return -1;
    };
    (void)qm;
    pn = find(bclist[4 * t], bclist[4 * t + 1], bclist[4 * t + 2]);
    pm = find(bcnlist[3 * t], bcnlist[3 * t + 1], bcnlist[3 * t + 2]);
    if (pn < 0 || pm < 0) return fail(ctx, ZF_ERR_STATE, "BC list entry %d not in DOMLIST", t + 1);
    e.s = s_of_p[pn];
    e.sm = s_of_p[pm];
    e.code = code;
    e.m_int = interior(pm) ? 1 : 0;
    if (e.m_int && !(word[e.sm] & kWordMacro)) {
      word[e.sm] |= kWordMacro;
      nmacro++;
    }
    bce.push_back(e);
  }
  // Boundary entries whose neighbour is an interior node are written by that
  // node's lane in the stream kernel, right after it computes rho and u
  // (fused); they come first, grouped by neighbour.  The rest read their
  // neighbour's moments from the previous state and run in extra workgroups
  // of the same launch ("late" entries), in list order.
  std::stable_sort(bce.begin(), bce.end(), [](const BcEntry &a, const BcEntry &b) {
    if (a.m_int != b.m_int) return a.m_int > b.m_int;
    return a.m_int ? a.sm < b.sm : false;
  });
  std::vector<int32_t> fused_s, att_start;
  int n_t1 = 0;
  for (const BcEntry &e : bce) {
    if (!e.m_int) break;
    if (fused_s.empty() || fused_s.back() != e.sm) {
      fused_s.push_back(e.sm);
      att_start.push_back(n_t1);
    }
    n_t1++;
  }
  att_start.push_back(n_t1);
  ctx->n_bc = (int)bce.size();
  ctx->n_bc_fused = n_t1;
  ctx->n_fused_nodes = (int)fused_s.size();
  ctx->n_macro = nmacro;

  // upload
  int rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_s_of_p, sizeof(int32_t) * (size_t)N))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_p_of_s, sizeof(int32_t) * (size_t)N))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_word, sizeof(uint32_t) * (size_t)std::max(nint, 1))))
    return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_block_chord0, sizeof(int32_t) * block0.size())))
    return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_chords, sizeof(int32_t) * std::max<size_t>(chords.size(), 1))))
    return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_bc, sizeof(BcEntry) * std::max<size_t>(bce.size(), 1))))
    return rc;
  ZF_HIP(ctx, hipMemcpy(ctx->d_s_of_p, s_of_p.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_p_of_s, p_of_s.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
  if (nint)
    ZF_HIP(ctx, hipMemcpy(ctx->d_word, word.data(), sizeof(uint32_t) * nint, hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_block_chord0, block0.data(), sizeof(int32_t) * block0.size(),
                        hipMemcpyHostToDevice));
  if (!chords.empty())
    ZF_HIP(ctx, hipMemcpy(ctx->d_chords, chords.data(), sizeof(int32_t) * chords.size(),
                          hipMemcpyHostToDevice));
  if (!bce.empty())
    ZF_HIP(ctx, hipMemcpy(ctx->d_bc, bce.data(), sizeof(BcEntry) * bce.size(), hipMemcpyHostToDevice));
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_fused_s, sizeof(int32_t) * std::max<size_t>(fused_s.size(), 1))))
    return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_att_start, sizeof(int32_t) * att_start.size()))) return rc;
  if (!fused_s.empty())
    ZF_HIP(ctx, hipMemcpy(ctx->d_fused_s, fused_s.data(), sizeof(int32_t) * fused_s.size(),
                          hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_att_start, att_start.data(), sizeof(int32_t) * att_start.size(),
                        hipMemcpyHostToDevice));
  return ZF_OK;
}

int upload_lattice_impl(zf_ctx *ctx, const int32_t *b) {
  free_parent(ctx);  // coupling tables index the old lists
  free_adjoint(ctx);
  free_cells(ctx);  // cell-list and IBM box tables belong to the old lattice
  free_ibm(ctx);
  free_state(ctx);
  free_lattice(ctx);
  const zf_config &c = ctx->cfg;
  ctx->box = (size_t)c.im * c.jm * c.km;
  const int ntiles = (int)((ctx->box + kTile - 1) / kTile);
  int rc;
  int32_t *d_counts = nullptr;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_b, sizeof(int32_t) * ctx->box))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_domid, sizeof(int32_t) * ctx->box))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&d_counts, sizeof(int32_t) * (ntiles + 1)))) return rc;
  ZF_HIP(ctx, hipMemcpyAsync(ctx->d_b, b, sizeof(int32_t) * ctx->box, hipMemcpyHostToDevice,
                             ctx->stream));
  // count both lists first so the list buffers are sized exactly
  int32_t nd = 0, nb = 0;
  if ((rc = compact(ctx, 0, d_counts, ntiles, nullptr, &nd))) goto out;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_domlist, sizeof(int32_t) * 3 * (size_t)std::max(nd, 1))))
    goto out;
  if ((rc = compact(ctx, 0, d_counts, ntiles, ctx->d_domlist, &nd))) goto out;
  if ((rc = compact(ctx, ctx->parent_p ? 2 : 1, d_counts, ntiles, nullptr, &nb))) goto out;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_bclist, sizeof(int32_t) * 4 * (size_t)std::max(nb, 1))))
    goto out;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_bcnlist, sizeof(int32_t) * 3 * (size_t)std::max(nb, 1))))
    goto out;
  if (ctx->parent_p &&
      (rc = dev_alloc(ctx, (void **)&ctx->d_vbcn, sizeof(int32_t) * 3 * (size_t)std::max(nb, 1))))
    goto out;
  if ((rc = compact(ctx, ctx->parent_p ? 2 : 1, d_counts, ntiles, ctx->d_bclist, &nb))) goto out;
  ctx->dommax = nd;
  ctx->bcmax = nb;
  {
    const Box bx{c.im, c.jm, c.km, {c.periodic[0], c.periodic[1], c.periodic[2]}};
    const ParentMap pm{ctx->parent_p, ctx->parent_kover,
                       {ctx->parent_off[0], ctx->parent_off[1], ctx->parent_off[2]}};
    if (nb > 0)
      hipLaunchKernelGGL(k_bc_neighbour, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream,
                         ctx->d_b, bx, nb, ctx->d_bclist, ctx->d_bcnlist, pm, ctx->d_vbcn);
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_nbr, sizeof(int32_t) * 18 * (size_t)std::max(nd, 1))))
      goto out;
    if (nd > 0)
      hipLaunchKernelGGL(k_pull_neighbours, dim3((nd + 255) / 256), dim3(256), 0, ctx->stream,
                         ctx->d_domid, bx, nd, ctx->d_domlist, ctx->d_nbr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      rc = fail(ctx, ZF_ERR_HIP, "list kernels failed: %s", hipGetErrorString(e));
      goto out;
    }
  }
  {
    std::vector<int32_t> domlist(3 * (size_t)nd), nbr(18 * (size_t)nd), bclist(4 * (size_t)nb),
        bcnlist(3 * (size_t)nb);
    hipError_t e = hipSuccess;
    if (nd) e = hipMemcpy(domlist.data(), ctx->d_domlist, 4 * domlist.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess && nd) e = hipMemcpy(nbr.data(), ctx->d_nbr, 4 * nbr.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess && nb) e = hipMemcpy(bclist.data(), ctx->d_bclist, 4 * bclist.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess && nb) e = hipMemcpy(bcnlist.data(), ctx->d_bcnlist, 4 * bcnlist.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(ctx, ZF_ERR_HIP, "list download failed: %s", hipGetErrorString(e));
      goto out;
    }
    ctx->have_lists = true;
    if ((rc = build_kernel_tables(ctx, b, domlist, nbr, bclist, bcnlist))) goto out;
    // box -> storage index map, for the operators that look up lattice
    // nodes by position (IBM, wall shear stress)
    if ((rc = dev_alloc(ctx, (void **)&ctx->d_sbox, 4 * ctx->box))) goto out;
    hipLaunchKernelGGL(k_sbox, dim3((unsigned)((ctx->box + 255) / 256)), dim3(256), 0, ctx->stream, ctx->d_domid,
                       ctx->box, ctx->d_s_of_p, ctx->d_sbox);
    if ((e = hipGetLastError()) == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      rc = fail(ctx, ZF_ERR_HIP, "box map: %s", hipGetErrorString(e));
      goto out;
    }
  }
  ctx->stride = ((size_t)std::max(ctx->dommax, 1) + 255) / 256 * 256;
  ctx->have_lattice = true;
out:
  dev_free(ctx, (void **)&d_counts);
  // a rejected geometry keeps its (valid) lists for inspection; any other
  // failure releases everything
  if (rc && !(rc == ZF_ERR_GEOMETRY && ctx->have_lists)) free_lattice(ctx);
  return rc;
}

}  // namespace zf

using namespace zf;

extern "C" {

int zf_upload_lattice(zf_ctx *ctx, const int32_t *b_index, int32_t *dommax, int32_t *bcmax,
                      int32_t *n_interior) {
  if (!ctx || !b_index) return ZF_ERR_ARG;
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  int rc = upload_lattice_impl(ctx, b_index);
  if (rc) return rc;
  if ((rc = alloc_state(ctx))) return rc;
  if (dommax) *dommax = ctx->dommax;
  if (bcmax) *bcmax = ctx->bcmax;
  if (n_interior) *n_interior = ctx->n_int;
  return ZF_OK;
}

int zf_download_lists(zf_ctx *ctx, int32_t *domlist, int32_t *bclist, int32_t *bcnlist) {
  if (!ctx) return ZF_ERR_ARG;
  if (!ctx->have_lists) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  if (domlist && ctx->dommax)
    ZF_HIP(ctx, hipMemcpy(domlist, ctx->d_domlist, 12 * (size_t)ctx->dommax, hipMemcpyDeviceToHost));
  if (bclist && ctx->bcmax)
    ZF_HIP(ctx, hipMemcpy(bclist, ctx->d_bclist, 16 * (size_t)ctx->bcmax, hipMemcpyDeviceToHost));
  if (bcnlist && ctx->bcmax)
    ZF_HIP(ctx, hipMemcpy(bcnlist, ctx->d_bcnlist, 12 * (size_t)ctx->bcmax, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_download_domid(zf_ctx *ctx, int32_t *domid) {
  if (!ctx || !domid) return ZF_ERR_ARG;
  if (!ctx->have_lists) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  ZF_HIP(ctx, hipMemcpy(domid, ctx->d_domid, 4 * ctx->box, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_download_neighbours(zf_ctx *ctx, int32_t *nbr) {
  if (!ctx || !nbr) return ZF_ERR_ARG;
  if (!ctx->have_lists) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  if (ctx->dommax)
    ZF_HIP(ctx, hipMemcpy(nbr, ctx->d_nbr, 72 * (size_t)ctx->dommax, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_get_lattice_info(zf_ctx *ctx, zf_lattice_info *info) {
  if (!ctx || !info) return ZF_ERR_ARG;
  if (!ctx->have_lattice) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  info->dommax = ctx->dommax;
  info->bcmax = ctx->bcmax;
  info->n_interior = ctx->n_int;
  info->n_chords = ctx->n_chords;
  info->n_bc_active = ctx->n_bc;
  info->n_bb_links = ctx->n_bb_links;
  info->n_macro_nodes = ctx->n_macro;
  info->stride = (int32_t)ctx->stride;
  info->device_bytes = ctx->device_bytes;
  return ZF_OK;
}

}  // extern "C"
