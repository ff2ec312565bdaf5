// zf_cells.hip -- triangle cell list, cell-cell and cell-wall repulsion.
//
// Replaces, for the lattice of this context, ZebraflowWT.c:
//   EulNeigh cell list       :5851-5900   (built by atomic capture there;
//                                           here a radix sort of
//                                           (bin << 32 | triangle) keys, so
//                                           each bin's triangles come in
//                                           (cell, triangle) order)
//   cell-wall repulsion      :8998-9142   (nearest wall triangle through the
//                                           W_index map, plane distance)
//   cell-cell repulsion      :9913-10052  (16-point sub-triangle probe)
//   W_index / wall lists     :3715-3826, :3900-3935
// One lane per membrane node (repulsion) or per triangle (binning); all
// sums in FP64 in a fixed order.
#include "zf_internal.h"

#include <hipcub/hipcub.hpp>

#include <vector>

namespace zf {
namespace {

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct Grid {
  double x0, y0, z0, dx;
  int im, jm, km;
};

Grid grid_of(const zf_ctx *ctx) {
  return Grid{ctx->icfg.x0, ctx->icfg.y0, ctx->icfg.z0, ctx->icfg.dx, ctx->cfg.im, ctx->cfg.jm,
              ctx->cfg.km};
}

// The reference's 9^3 search window, (int)((x-X0)/DX + 1.01) +- 4, clipped.
__device__ __forceinline__ void window(const Grid &g, double x, double y, double z, int lo[3],
                                       int hi[3]) {
  const int i = (int)((x - g.x0) / g.dx + 1.01), j = (int)((y - g.y0) / g.dx + 1.01),
            k = (int)((z - g.z0) / g.dx + 1.01);
  lo[0] = max(1, i - 4); hi[0] = min(g.im, i + 4);
  lo[1] = max(1, j - 4); hi[1] = min(g.jm, j + 4);
  lo[2] = max(1, k - 4); hi[2] = min(g.km, k + 4);
}

__device__ __forceinline__ size_t bq(const Grid &g, int i, int j, int k) {
  return ((size_t)(i - 1) * g.jm + (j - 1)) * g.km + (k - 1);
}

// Cell list: nearest listed node to each triangle centroid, strict < in
// (i,j,k) order, kept below the bin cut-off (ZebraflowWT.c:5857-5887).
// The nearest lattice point of a centroid is a corner of the unit cell
// holding it, and every other lattice point is strictly farther than one
// of those corners; so when the nearest corner is listed (B >= 0), the
// minimum over the 8 corners, visited in (i,j,k) order, is the window's
// minimum with the same tie rule.  Otherwise the full window is scanned.
// Also records, for the repulsion pass, the largest centroid-to-bin
// distance and the largest centroid-to-vertex distance of any triangle
// (atomicMax on the bits of non-negative doubles, which order like the
// values).
__global__ void k_tri_bin(const double *__restrict__ tg, size_t TT, const double *__restrict__ X,
                          size_t NT, int nn, const int32_t *__restrict__ tri, int nt, int ncell,
                          const int32_t *__restrict__ status, const int32_t *__restrict__ b,
                          const int32_t *__restrict__ domid, Grid g, double cutoff, int dommax,
                          int32_t *__restrict__ tbin, uint64_t *__restrict__ keys,
                          unsigned long long *__restrict__ bounds, uint32_t *__restrict__ row_bd) {
  const int q = xcd_range_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool live = q < ncell * nt;
  int bin = 0;
  double bdist = 0, ext = 0;
  if (live && status[q / nt] > kStatusDead) {
    const double x = tg[4 * TT + q], y = tg[5 * TT + q], z = tg[6 * TT + q];
    double mind = 5000.0;
    int bi = 0, bj = 0, bk = 0;
    // corners of the unit cell around the centroid (1-based lattice indices)
    const int ci = (int)floor((x - g.x0) / g.dx) + 1, cj = (int)floor((y - g.y0) / g.dx) + 1,
              ck = (int)floor((z - g.z0) / g.dx) + 1;
    // fast path only where all 8 corners lie in the box
    const bool inside = ci >= 1 && ci + 1 <= g.im && cj >= 1 && cj + 1 <= g.jm && ck >= 1 &&
                        ck + 1 <= g.km;
    double cmin = 5000.0;  // nearest corner, listed or not
    for (int i = ci; i <= ci + 1 && inside; i++)
      for (int j = cj; j <= cj + 1; j++)
        for (int k = ck; k <= ck + 1; k++) {
          const double dx = g.x0 + (i - 1) * g.dx - x, dy = g.y0 + (j - 1) * g.dx - y,
                       dz = g.z0 + (k - 1) * g.dx - z;
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          cmin = fmin(cmin, d);
          if (b[bq(g, i, j, k)] < 0) continue;
          if (d < mind) { mind = d; bi = i; bj = j; bk = k; }
        }
    if (!inside || !(mind <= cmin)) {
      // Every lattice point outside the 4^3 block ci-1..ci+2 (per axis) is
      // at least 2 dx from the centroid, so a listed block point nearer
      // than that (by a margin far above rounding) is the window's nearest,
      // found first in the same (i,j,k) order; the block's b words are read
      // a plane at a time (in-bounds addresses, out-of-box points skipped).
      mind = 5000.0;
#pragma unroll 1
      for (int a = 0; a < 4; a++) {
        const int i = ci - 1 + a;
        int bw[16];
#pragma unroll
        for (int u = 0; u < 16; u++)
          bw[u] = b[bq(g, min(max(i, 1), g.im), min(max(cj - 1 + u / 4, 1), g.jm),
                       min(max(ck - 1 + u % 4, 1), g.km))];
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const int j = cj - 1 + u / 4, k = ck - 1 + u % 4;
          if (i < 1 || i > g.im || j < 1 || j > g.jm || k < 1 || k > g.km || bw[u] < 0) continue;
          const double dx = g.x0 + (i - 1) * g.dx - x, dy = g.y0 + (j - 1) * g.dx - y,
                       dz = g.z0 + (k - 1) * g.dx - z;
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d < mind) { mind = d; bi = i; bj = j; bk = k; }
        }
      }
      if (!(mind < 2.0 * g.dx * (1.0 - 1e-9))) {
        int lo[3], hi[3];
        window(g, x, y, z, lo, hi);
        mind = 5000.0;
        for (int i = lo[0]; i <= hi[0]; i++)
          for (int j = lo[1]; j <= hi[1]; j++)
            for (int k = lo[2]; k <= hi[2]; k++) {
              if (b[bq(g, i, j, k)] < 0) continue;
              const double dx = g.x0 + (i - 1) * g.dx - x, dy = g.y0 + (j - 1) * g.dx - y,
                           dz = g.z0 + (k - 1) * g.dx - z;
              const double d = sqrt(dx * dx + dy * dy + dz * dz);
              if (d < mind) { mind = d; bi = i; bj = j; bk = k; }
            }
      }
    }
    if (mind < cutoff) {
      bin = domid[bq(g, bi, bj, bk)];
      bdist = mind;
      // the largest centroid-to-bin distance of the bins of row (bi, bj),
      // FP32 rounded up (the bits of non-negative floats order like them);
      // read first, the atomic only when it raises the bound
      float bf = (float)mind;
      if ((double)bf < mind) bf = nextafterf(bf, INFINITY);
      uint32_t *rb = &row_bd[(size_t)(bi - 1) * g.jm + (bj - 1)];
      if (__float_as_uint(bf) > __hip_atomic_load(rb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(rb, __float_as_uint(bf));
    }
    const size_t b0 = (size_t)(q / nt) * nn;
    const int t = q % nt;
    for (int v = 0; v < 3; v++) {
      const size_t n = b0 + tri[3 * t + v];
      const double dx = X[n] - x, dy = X[NT + n] - y, dz = X[2 * NT + n] - z;
      ext = fmax(ext, sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  if (live) {
    tbin[q] = bin;
    const uint64_t key_bin = bin > 0 ? (uint64_t)(bin - 1) : (uint64_t)dommax;
    keys[q] = (key_bin << 32) | (uint64_t)q;
  }
  // one atomic per wave for each bound (every lane of the wave takes part)
  for (int off = 32; off > 0; off >>= 1) {
    bdist = fmax(bdist, __shfl_xor(bdist, off));
    ext = fmax(ext, __shfl_xor(ext, off));
  }
  // The bounds are read first and the atomic made only by a wave that
  // raises one (the read may be older than the bound, never above it), so
  // the many waves of one launch do not queue on the same two words.
  if ((threadIdx.x & 63) == 0) {
    const unsigned long long bb = (unsigned long long)__double_as_longlong(bdist),
                             be = (unsigned long long)__double_as_longlong(ext);
    if (bb > __hip_atomic_load(&bounds[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(&bounds[0], bb);
    if (be > __hip_atomic_load(&bounds[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(&bounds[1], be);
  }
}

__global__ void k_bin_ranges(const uint64_t *__restrict__ keys, int n, uint64_t nbins,
                             int32_t *__restrict__ range) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t t = keys[i] >> 32;
  if (t >= nbins) return;
  if (i == 0 || (keys[i - 1] >> 32) != t) range[2 * t] = i;
  if (i == n - 1 || (keys[i + 1] >> 32) != t) range[2 * t + 1] = i + 1;
}

// Keys per bin (from the bin ranges), the input of the exclusive scan that
// gives each bin the offset of its first key (bins 0..nbins, the last 0).
__global__ void k_bin_counts(const int32_t *__restrict__ range, int nbins, int32_t *__restrict__ cnt) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d > nbins) return;
  cnt[d] = d < nbins ? range[2 * d + 1] - range[2 * d] : 0;
}

__global__ void k_fill(int32_t *a, size_t n, int32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

struct RepParams {
  double k_cell, cutoff, search, k_wall, wall_d0, wall_dmax, wall_search;
  double wall_reach2;  // (wall_search + largest wall-triangle extent)^2, see k_wall_repulsion
  int use_bounds;      // the cell list's distance bounds match the current geometry
};

// Slack on the distance prefilters below, far above FP64 rounding of the
// distances and far below any geometric scale.
constexpr double kReachSlack = 1e-9;

// entries of a lane's queue of passing triangles in the repulsion kernel
constexpr int kRepQueue = 32;

// Sub-triangle centroids of a triangle refined four times
// (ZebraflowWT.c:9990-10025): the reference forms them through chained edge
// midpoints; centroid s lies at (A a + B b + C c)/12 with integer weights
// (10,1,1) (8,2,2) (7,4,1) (5,5,2) (4,7,1) (2,8,2) (1,10,1) (7,1,4) (5,2,5)
// (4,4,4) (2,5,5) (1,7,4) (4,1,7) (2,2,8) (1,4,7) (1,1,10), so the values
// agree to rounding.  The weights are packed 4 bits per centroid (s in bits
// 4s..4s+3), which lets a lane form one centroid at a time from scalar
// constants instead of holding all 16.
constexpr uint64_t kSubWA = 0x112412457124578aull, kSubWB = 0x142175421a875421ull,
                   kSubWC = 0xa787454541212121ull;

// Margins of the FP32 prefilters of the repulsion kernel.  A coordinate a
// rounded to FP32 is off by at most |a| 2^-24, so a node-to-centroid
// distance formed from FP32 values is off by less than
// 2^-22 (|x| + |y| + |z| + 3 reach) for candidates within reach; the margin
// 2^-20 (|x| + |y| + |z| + 4 reach) is four times that.  Unit normals are off
// by 6e-8 per component and their FP32 dot product by less than 1e-6.  A
// candidate is dropped in FP32 only when it is beyond the FP64 tests' reach
// by more than these margins, so the set of candidates reaching the FP64
// tests, and the sums, are unchanged.  Likewise a candidate is taken as
// within reach, or facing away, on the FP32 values alone only when it is
// so by more than the margins.  The normals are stored packed (each
// component off by at most 1/1022), so their dot product with the node's
// normal is off by less than 3/1022 + 1e-6; the margin is 4e-3.
constexpr float kRepDotMargin = 4e-3f;

// FP32 records of the sorted cell list, one float4 per key in sorted order:
// the centroid, and the normal packed as three 10-bit signed fractions of
// 511 (each component off by at most 1/1022), so that the repulsion kernel
// reads a row's candidates contiguously, one 16-byte load each.
__device__ __forceinline__ uint32_t pack_normal(double x, double y, double z) {
  auto q = [](double v) { return (uint32_t)((int)rint(fmin(fmax(v, -1.0), 1.0) * 511.0) & 1023); };
  return q(x) | (q(y) << 10) | (q(z) << 20);
}
__device__ __forceinline__ float unpack_comp(uint32_t w, int shift) {
  return (float)((int)(w << (22 - shift)) >> 22) * (1.0f / 511.0f);
}

__global__ void k_cand_records(const uint64_t *__restrict__ keys, int n, const double *__restrict__ tg,
                               size_t TT, float4 *__restrict__ rec) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int q = (int)(keys[e] & 0xffffffffu);
  rec[e] = make_float4((float)tg[4 * TT + q], (float)tg[5 * TT + q], (float)tg[6 * TT + q],
                       __uint_as_float(pack_normal(tg[TT + q], tg[2 * TT + q], tg[3 * TT + q])));
}

// The 16-point contribution of triangle q (cell m = q / nt) to the
// repulsion velocity of a node at (xc, yc, zc), ZebraflowWT.c:9990-10052,
// added to (v0, v1, v2) sub-centroid by sub-centroid.
__device__ __forceinline__ void rep_triangle(int q, int nt, int nn, size_t NT, const int32_t *__restrict__ tri,
                                             const double *__restrict__ X, double xc, double yc, double zc,
                                             const RepParams &r, double &v0, double &v1, double &v2) {
  const int m = q / nt, t = q % nt;
  const int n1 = tri[3 * t], n2 = tri[3 * t + 1], n3 = tri[3 * t + 2];
  const size_t b0 = (size_t)m * nn;
  const double ax = X[b0 + n1], ay = X[NT + b0 + n1], az = X[2 * NT + b0 + n1];
  const double bx = X[b0 + n2], by = X[NT + b0 + n2], bz = X[2 * NT + b0 + n2];
  const double cx = X[b0 + n3], cy = X[NT + b0 + n3], cz = X[2 * NT + b0 + n3];
#pragma unroll 2
  for (int sc = 0; sc < 16; sc++) {
    const double wa = (double)((kSubWA >> (4 * sc)) & 15u);
    const double wb = (double)((kSubWB >> (4 * sc)) & 15u);
    const double wc = (double)((kSubWC >> (4 * sc)) & 15u);
    const double px = (wa * ax + wb * bx + wc * cx) * (1.0 / 12.0) - xc;
    const double py = (wa * ay + wb * by + wc * cy) * (1.0 / 12.0) - yc;
    const double pz = (wa * az + wb * bz + wc * cz) * (1.0 / 12.0) - zc;
    const double d = sqrt(px * px + py * py + pz * pz);
    if (d < r.cutoff) {
      const double f = -r.k_cell * (r.cutoff - d) / r.cutoff;
      v0 += f * px / d;
      v1 += f * py / d;
      v2 += f * pz / d;
    }
  }
}

// The terms of rep_triangle, unsummed: term sc of each component at
// t[sc * STRIDE], +0 where sub-centroid sc is beyond the cut-off.
template <int STRIDE>
__device__ __forceinline__ void rep_terms(int q, int nt, int nn, size_t NT, const int32_t *__restrict__ tri,
                                              const double *__restrict__ X, double xc, double yc, double zc,
                                              const RepParams &r, double *t0, double *t1, double *t2) {
  const int m = q / nt, t = q % nt;
  const int n1 = tri[3 * t], n2 = tri[3 * t + 1], n3 = tri[3 * t + 2];
  const size_t b0 = (size_t)m * nn;
  const double ax = X[b0 + n1], ay = X[NT + b0 + n1], az = X[2 * NT + b0 + n1];
  const double bx = X[b0 + n2], by = X[NT + b0 + n2], bz = X[2 * NT + b0 + n2];
  const double cx = X[b0 + n3], cy = X[NT + b0 + n3], cz = X[2 * NT + b0 + n3];
#pragma unroll 2
  for (int sc = 0; sc < 16; sc++) {
    const double wa = (double)((kSubWA >> (4 * sc)) & 15u);
    const double wb = (double)((kSubWB >> (4 * sc)) & 15u);
    const double wc = (double)((kSubWC >> (4 * sc)) & 15u);
    const double px = (wa * ax + wb * bx + wc * cx) * (1.0 / 12.0) - xc;
    const double py = (wa * ay + wb * by + wc * cy) * (1.0 / 12.0) - yc;
    const double pz = (wa * az + wb * bz + wc * cz) * (1.0 / 12.0) - zc;
    const double d = sqrt(px * px + py * py + pz * pz);
    double a = 0, b = 0, c = 0;
    if (d < r.cutoff) {
      const double f = -r.k_cell * (r.cutoff - d) / r.cutoff;
      a = f * px / d;
      b = f * py / d;
      c = f * pz / d;
    }
    t0[sc * STRIDE] = a;
    t1[sc * STRIDE] = b;
    t2[sc * STRIDE] = c;
  }
}

// Triangles listed per node by the walk (kernel k_cell_repulsion<1>), in
// walk order: the first kRepSlots in the node's own slots, the rest in
// chunks of kRepChunk words taken from a pool (one atomic per chunk), each
// holding kRepChunk - 1 triangles and, last, the next chunk's number.  A
// node whose chunk the pool cannot supply goes through the whole walk
// again with the evaluation (k_cell_repulsion<0>).  At 1200 cells of 642
// nodes in the 15-segment vessel pair (haematocrit 0.094) 92 % of the nodes
// list none and 1.4 % more than 16.
constexpr int kRepSlots = 16;
constexpr int kRepChunk = 32;
constexpr int kRepPoolHead = 64;  // pool words: the chunk counter, padded
// Nodes with listed triangles are appended to one of kRepClasses lists by
// their count (1-4, 5-8, 9-16, 17-32, more), so an evaluation wave holds
// nodes of similar cost.  Layout of the list buffer: kRepClasses counts,
// padded to 64 words, then one region of NT entries per class.
constexpr int kRepClasses = 5;
constexpr int kRepListHead = 64;
__device__ __forceinline__ int rep_class(int np) {
  return np <= 4 ? 0 : np <= 8 ? 1 : np <= 16 ? 2 : np <= 32 ? 3 : 4;
}
// Chunks in the pool: one per 8 nodes (room for about 4 extra triangles per
// node on average; 1.4 % of the nodes need one or two chunks at 1200 cells).
inline int rep_pool_chunks(size_t NT) { return (int)std::max<size_t>(4096, NT / 8); }

// Cell-cell repulsion, one lane per node of a released cell (status > 0).
// The window is walked row by row (fixed i, j): DOMID is k-fastest and the
// cell list is sorted by bin, so the triangles of a row's bins in reach are
// one contiguous run of the sorted keys, visited in the same (bin, key)
// order as a bin-by-bin walk.  A row's 9 DOMID and bin-range loads are
// issued together (in-bounds addresses, results selected afterwards), and
// candidates are fetched 4 at a time, so a lane waits on a few batched
// loads per row instead of a chain of dependent loads per bin.
template <int MODE>
__device__ __forceinline__ void rep_walk(
    int gq, const double *__restrict__ X, size_t NT, int nn, int nt,
    const int32_t *__restrict__ tri, const int32_t *__restrict__ status,
    const double *__restrict__ nnormal, const double *__restrict__ tg, size_t TT,
    const int32_t *__restrict__ domid, const float4 *__restrict__ crec, const uint64_t *__restrict__ keys,
    const int32_t *__restrict__ bin_off, const unsigned long long *__restrict__ bounds,
    const uint32_t *__restrict__ row_bd, Grid g,
    RepParams r, int32_t *__restrict__ npair, int32_t *__restrict__ pairs, int32_t *__restrict__ ochunk,
    int32_t *__restrict__ pool, int pool_cap, int32_t *__restrict__ lists, int32_t *__restrict__ over,
    double *__restrict__ vrep) {
  const int c = gq / nn, node = gq % nn;
  // released cells and clones (the reference's :9920-9921, :10055)
  if (status[c] == 0 || status[c] <= kStatusDead) {
    if (MODE == 1) npair[gq] = -1;
    return;
  }
  const double xc = X[gq], yc = X[NT + gq], zc = X[2 * NT + gq];
  const double e1x = nnormal[gq], e1y = nnormal[NT + gq], e1z = nnormal[2 * NT + gq];
  int lo[3], hi[3];
  window(g, xc, yc, zc, lo, hi);
  double v0 = 0, v1 = 0, v2 = 0;
  // Exact prefilters (they skip only work that adds nothing to the sum):
  //  - a triangle contributes only through sub-centroids closer than
  //    `cutoff`; they lie within `ext` (largest centroid-to-vertex distance)
  //    of its centroid, so a centroid farther than min(search, cutoff + ext)
  //    contributes nothing;
  //  - a triangle's bin is within `bd` (largest centroid-to-bin distance)
  //    of its centroid, so a bin farther than that reach plus bd holds no
  //    contributing triangle.
  // Both bounds were recorded when the cell list was built; the three tests
  // on a triangle are filters of one sum, so their order is free.
  double treach = r.search + kReachSlack, breach = 1e300;
  if (r.use_bounds) {
    const double bd = __longlong_as_double((long long)bounds[0]);
    const double ext = __longlong_as_double((long long)bounds[1]);
    treach = fmin(r.search, r.cutoff + ext) + kReachSlack;
    breach = treach + bd + kReachSlack;
  }
  const double treach2 = treach * treach, breach2 = breach * breach;
  const double rmargin = std::ldexp(fabs(xc) + fabs(yc) + fabs(zc) + 4.0 * treach, -20);
  const float treach_f2 = (float)((treach + rmargin) * (treach + rmargin));
  // within reach for certain: below treach - margin (and so below search,
  // treach being at most search plus a slack far below the margin)
  const float tinner_f2 = (float)(fmax(treach - rmargin, 0.0) * fmax(treach - rmargin, 0.0));
  const float xf = (float)xc, yf = (float)yc, zf = (float)zc;
  const float e1xf = (float)e1x, e1yf = (float)e1y, e1zf = (float)e1z;
  // Triangles that pass the three filters go to the lane's queue in LDS,
  // and the 16-point sub-triangle evaluation runs when any active lane's
  // queue is nearly full (and once at the end), every active lane draining
  // its own queue in order: the lanes then evaluate their candidates
  // together instead of one or two at a time, and each lane's sum keeps its
  // candidate order.
  __shared__ int queue[MODE == 0 ? kRepQueue : 1][128];
  int qlen = 0;
  auto drain = [&]() {
    if constexpr (MODE == 0) {
      for (int k = 0; k < qlen; k++)
        rep_triangle(queue[k][threadIdx.x], nt, nn, NT, tri, X, xc, yc, zc, r, v0, v1, v2);
    }
    qlen = 0;
  };
  int np = 0;       // mode 1: triangles listed
  int chunk = -1;   // mode 1: the pool chunk being filled
  bool lost = false;  // mode 1: the pool ran out
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++) {
      const double qx = g.x0 + (i - 1) * g.dx - xc, qy = g.y0 + (j - 1) * g.dx - yc;
      const double rij = qx * qx + qy * qy;
      if (rij >= breach2) continue;
      const size_t row = bq(g, i, j, 1);
      int ov[9];
#pragma unroll
      for (int t = 0; t < 9; t++) ov[t] = domid[row + (size_t)(min(lo[2] + t, hi[2]) - 1)];
      // the bins of this row hold triangles within rbd of them (the row's
      // own bound, at most the list's): the reach of this row's bins
      double rbreach2 = breach2;
      if (r.use_bounds) {
        const double rb = treach + (double)__uint_as_float(row_bd[(size_t)(i - 1) * g.jm + (j - 1)]) + kReachSlack;
        rbreach2 = fmin(breach2, rb * rb);
      }
      // the bins in reach are a k-interval of the row; its listed nodes
      // have consecutive DOMIDs (k is the fastest index of the list), so
      // their triangles are the sorted keys from the first one's offset to
      // past the last one's
      int dfirst = 0, dlast = -1;
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int k = lo[2] + t;
        const double qz = g.z0 + (k - 1) * g.dx - zc;
        const bool ok = k <= hi[2] && rij + qz * qz < rbreach2 && ov[t] > 0;
        dfirst = (ok && dlast < 0) ? ov[t] : dfirst;
        dlast = ok ? ov[t] : dlast;
      }
      if (dlast < 0) continue;
      const int beg = bin_off[dfirst - 1], end = bin_off[dlast];
      for (int e0 = beg; e0 < end; e0 += 4) {
        // room for this batch's 4 candidates in every active lane's queue
        if (MODE == 0 && __ballot(qlen > kRepQueue - 4)) drain();
        float4 ca[4];
#pragma unroll
        for (int u = 0; u < 4; u++) ca[u] = crec[min(e0 + u, end - 1)];
#pragma unroll 1
        for (int u = 0; u < 4 && e0 + u < end; u++) {
          // FP32 tests first: a candidate clearly beyond reach or facing
          // the same way is dropped, one clearly within reach and facing
          // away passes; only one within a margin of either bound takes
          // the FP64 test (same decisions as the FP64 tests alone)
          const float fx = xf - ca[u].x, fy = yf - ca[u].y, fz = zf - ca[u].z;
          const float d2f = fx * fx + fy * fy + fz * fz;
          if (d2f > treach_f2) continue;
          const uint32_t pn = __float_as_uint(ca[u].w);
          const float dotf = e1xf * unpack_comp(pn, 0) + e1yf * unpack_comp(pn, 10) + e1zf * unpack_comp(pn, 20);
          if (dotf > kRepDotMargin) continue;
          const int q = (int)(keys[e0 + u] & 0xffffffffu);
          if (d2f >= tinner_f2) {
            const double dx = xc - tg[4 * TT + q], dy = yc - tg[5 * TT + q], dz = zc - tg[6 * TT + q];
            const double dc2 = dx * dx + dy * dy + dz * dz;
            if (dc2 >= treach2 || sqrt(dc2) >= r.search) continue;
          }
          const int m = q / nt, t = q % nt;
          const int n1 = tri[3 * t], n2 = tri[3 * t + 1], n3 = tri[3 * t + 2];
          if (m == c && (n1 == node || n2 == node || n3 == node)) continue;
          if (dotf >= -kRepDotMargin && e1x * tg[TT + q] + e1y * tg[2 * TT + q] + e1z * tg[3 * TT + q] >= 0)
            continue;
          if constexpr (MODE == 0) {
            queue[qlen][threadIdx.x] = q;
            qlen++;
          } else {
            if (np < kRepSlots) {
              pairs[(size_t)gq * kRepSlots + np] = q;
            } else {
              const int at = (np - kRepSlots) % (kRepChunk - 1);
              if (at == 0 && !lost) {
                const int cn = atomicAdd(pool, 1);
                if (cn >= pool_cap) {
                  lost = true;
                } else {
                  if (chunk < 0) ochunk[gq] = cn;
                  else pool[kRepPoolHead + (size_t)chunk * kRepChunk + kRepChunk - 1] = cn;
                  chunk = cn;
                }
              }
              if (!lost) pool[kRepPoolHead + (size_t)chunk * kRepChunk + at] = q;
            }
            np++;
          }
        }
      }
    }
  if constexpr (MODE == 1) {
    npair[gq] = np;
    if (lost) over[1 + atomicAdd(over, 1)] = gq;
    const int cls = np >= 1 && !lost ? rep_class(np) : -1;
    // one class per pass, so each atomic has one address across the wave
#pragma unroll
    for (int k = 0; k < kRepClasses; k++)
      if (cls == k) lists[kRepListHead + (size_t)k * NT + atomicAdd(lists + k, 1)] = gq;
  } else {
    drain();
    vrep[gq] += v0;
    vrep[NT + gq] += v1;
    vrep[2 * NT + gq] += v2;
  }
}

template <int MODE>
__global__ __launch_bounds__(128) void k_cell_repulsion(
    const double *__restrict__ X, size_t NT, int nn, int nt, int ncell,
    const int32_t *__restrict__ tri, const int32_t *__restrict__ status,
    const double *__restrict__ nnormal, const double *__restrict__ tg, size_t TT,
    const int32_t *__restrict__ domid, const float4 *__restrict__ crec, const uint64_t *__restrict__ keys,
    const int32_t *__restrict__ bin_off, const unsigned long long *__restrict__ bounds,
    const uint32_t *__restrict__ row_bd, Grid g,
    RepParams r, int32_t *__restrict__ npair, int32_t *__restrict__ pairs, int32_t *__restrict__ ochunk,
    int32_t *__restrict__ pool, int pool_cap, int32_t *__restrict__ lists, int32_t *__restrict__ over,
    double *__restrict__ vrep) {
  if constexpr (MODE == 1) {
    const int gq = xcd_range_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (gq >= ncell * nn) return;
    rep_walk<MODE>(gq, X, NT, nn, nt, tri, status, nnormal, tg, TT, domid, crec, keys, bin_off, bounds, row_bd, g, r,
                   npair, pairs, ochunk, pool, pool_cap, lists, over, vrep);
  } else {
    // the nodes whose passing triangles did not fit their list (over[0]
    // of them, listed from over[1] on)
    const int nover = over[0];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nover; i += gridDim.x * blockDim.x)
      rep_walk<MODE>(over[1 + i], X, NT, nn, nt, tri, status, nnormal, tg, TT, domid, crec, keys, bin_off,
                     bounds, row_bd, g, r, npair, pairs, ochunk, pool, pool_cap, lists, over, vrep);
  }
}

// The listed triangle p of node gq (its slots, then its pool chunks).
__device__ __forceinline__ int rep_pair_at(int gq, int p, const int32_t *__restrict__ pairs,
                                           const int32_t *__restrict__ ochunk, const int32_t *__restrict__ pool) {
  if (p < kRepSlots) return pairs[(size_t)gq * kRepSlots + p];
  int chunk = ochunk[gq];
  for (int h = (p - kRepSlots) / (kRepChunk - 1); h > 0; h--)
    chunk = pool[kRepPoolHead + (size_t)chunk * kRepChunk + kRepChunk - 1];
  return pool[kRepPoolHead + (size_t)chunk * kRepChunk + (p - kRepSlots) % (kRepChunk - 1)];
}

// The nodes of the two heaviest classes (more than 16 listed triangles),
// one wave per node: lane p forms the 16 terms of triangle p (32 at a
// time) in LDS, and lanes 0, 1, 2 add those of one component each in walk
// order, term by term, as the per-node evaluation does, so that a node's
// triangles are gathered together instead of one after another.  A term
// beyond the cut-off is stored as +0 and added: the sums start at +0 and a
// sum of doubles is -0 only if both addends are, so adding +0 leaves every
// partial sum as it is and the sums are the per-node evaluation's, bit for
// bit.
constexpr int kRepWideClass = 3;
__global__ __launch_bounds__(64) void k_rep_eval_wide(const double *__restrict__ X, size_t NT, int nn, int nt,
                                                      const int32_t *__restrict__ tri,
                                                      const int32_t *__restrict__ npair,
                                                      const int32_t *__restrict__ pairs,
                                                      const int32_t *__restrict__ ochunk,
                                                      const int32_t *__restrict__ pool,
                                                      const int32_t *__restrict__ lists, RepParams r,
                                                      double *__restrict__ vrep) {
  constexpr int B = 32;  // triangles per batch (LDS 12 KB per wave)
  __shared__ double t0[16 * B], t1[16 * B], t2[16 * B];  // [term][lane]
  const int lane = threadIdx.x;
  int cnt[kRepClasses], total = 0;
#pragma unroll
  for (int k = kRepWideClass; k < kRepClasses; k++) total += (cnt[k] = lists[k]);
  for (int i = blockIdx.x; i < total; i += gridDim.x) {
    int k = kRepClasses - 1, at = i;
    while (at >= cnt[k]) at -= cnt[k--];
    const int gq = lists[kRepListHead + (size_t)k * NT + at];
    const int np = npair[gq];
    const double xc = X[gq], yc = X[NT + gq], zc = X[2 * NT + gq];
    double v0 = 0;
    for (int p0 = 0; p0 < np; p0 += B) {
      const int p = p0 + lane;
      if (lane < B && p < np)
        rep_terms<B>(rep_pair_at(gq, p, pairs, ochunk, pool), nt, nn, NT, tri, X, xc, yc, zc, r, t0 + lane,
                     t1 + lane, t2 + lane);
      __syncthreads();
      if (lane < 3) {
        const double *tc = lane == 0 ? t0 : lane == 1 ? t1 : t2;
        const int n = min(B, np - p0);
        for (int e = 0; e < n; e++) {
          double tv[16];
#pragma unroll
          for (int sc = 0; sc < 16; sc++) tv[sc] = tc[sc * B + e];
#pragma unroll
          for (int sc = 0; sc < 16; sc++) v0 += tv[sc];
        }
      }
      __syncthreads();
    }
    // lane c holds component c
    if (lane < 3) vrep[(size_t)lane * NT + gq] += v0;
  }
}

// The listed triangles of every node in the lighter class lists, in the
// walk's order (the same sums as the walk's own evaluation).  Entry i of
// the class lists taken heaviest class first, grid-stride.
__global__ __launch_bounds__(128) void k_rep_eval(const double *__restrict__ X, size_t NT, int nn, int nt,
                                                  const int32_t *__restrict__ tri,
                                                  const int32_t *__restrict__ npair,
                                                  const int32_t *__restrict__ pairs,
                                                  const int32_t *__restrict__ ochunk,
                                                  const int32_t *__restrict__ pool,
                                                  const int32_t *__restrict__ lists, RepParams r,
                                                  double *__restrict__ vrep) {
  int cnt[kRepWideClass], total = 0;
#pragma unroll
  for (int k = 0; k < kRepWideClass; k++) total += (cnt[k] = lists[k]);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    int k = kRepWideClass - 1, at = i;
    while (at >= cnt[k]) at -= cnt[k--];
    const int gq = lists[kRepListHead + (size_t)k * NT + at];
    const int np = npair[gq];
    const double xc = X[gq], yc = X[NT + gq], zc = X[2 * NT + gq];
    double v0 = 0, v1 = 0, v2 = 0;
    int chunk = 0;
    for (int p = 0; p < np; p++) {
      int q;
      if (p < kRepSlots) {
        q = pairs[(size_t)gq * kRepSlots + p];
      } else {
        const int at = (p - kRepSlots) % (kRepChunk - 1);
        if (at == 0)
          chunk = p == kRepSlots ? ochunk[gq] : pool[kRepPoolHead + (size_t)chunk * kRepChunk + kRepChunk - 1];
        q = pool[kRepPoolHead + (size_t)chunk * kRepChunk + at];
      }
      rep_triangle(q, nt, nn, NT, tri, X, xc, yc, zc, r, v0, v1, v2);
    }
    vrep[gq] += v0;
    vrep[NT + gq] += v1;
    vrep[2 * NT + gq] += v2;
  }
}

// Whole-node skip of the wall repulsion: with q0 the lattice node nearest x
// and d0 its distance to its nearest wall node (W_index), every wall node is
// at least d0 - |x - q0| from x.  A lattice node without a W_index entry lies
// outside the 9^3 window of every wall node, at least 3.9 dx from all of
// them.  If that bound reaches sqrt(wall_reach2), no triangle can be chosen
// (see the per-node test in k_wall_repulsion) and the node gets nothing.
__device__ __forceinline__ bool wall_clear(double x, double y, double z, const float *__restrict__ wdist,
                                           const Grid &g, const RepParams &r) {
  const int i0 = (int)floor((x - g.x0) / g.dx + 0.5) + 1, j0 = (int)floor((y - g.y0) / g.dx + 0.5) + 1,
            k0 = (int)floor((z - g.z0) / g.dx + 0.5) + 1;
  if (i0 < 1 || i0 > g.im || j0 < 1 || j0 > g.jm || k0 < 1 || k0 > g.km) return false;
  const float w0 = wdist[bq(g, i0, j0, k0)];
  const double d0 = isinf(w0) ? 3.9 * g.dx : (double)w0 * (1.0 - 1e-6);
  const double qx = g.x0 + (i0 - 1) * g.dx - x, qy = g.y0 + (j0 - 1) * g.dx - y,
               qz = g.z0 + (k0 - 1) * g.dx - z;
  const double lb = d0 - sqrt(qx * qx + qy * qy + qz * qz);
  return lb > 0 && lb * lb >= r.wall_reach2;
}

// flag[gq] = 1 for the nodes of released cells that wall_clear does not
// clear, 0 for the others.
__global__ __launch_bounds__(256) void k_wall_select(const double *__restrict__ X, size_t NT, int nn, int ncell,
                                                     const int32_t *__restrict__ status,
                                                     const float *__restrict__ wdist, Grid g, RepParams r,
                                                     int32_t *__restrict__ flag) {
  const int gq = blockIdx.x * blockDim.x + threadIdx.x;
  if (gq >= ncell * nn) return;
  flag[gq] = status[gq / nn] != 0 && status[gq / nn] > kStatusDead &&
                     !wall_clear(X[gq], X[NT + gq], X[2 * NT + gq], wdist, g, r)
                 ? 1
                 : 0;
}

// Cell-wall repulsion (ZebraflowWT.c:10030-10052), one wave per listed node
// (the list in node order, a contiguous part per XCD, so that the nodes of
// one cell, which search the same wall nodes, share an L2).  The reference
// walks the 9^3 window in (i, j, k) order, and for each W_index wall node
// each of its triangles' three vertices and centroid, keeping the triangle
// of the first point strictly nearer than all before it, starting from
// wall_search: the winner is the point of least distance, the earliest in
// that order among equal distances.  Here every point carries its order
// (its window position, then its triangle slot and point) and the
// lexicographically least (distance, order) wins, whatever the order of
// evaluation: the same triangle.  The wave's lanes take the window's rows
// (lane l rows l and l + 64), put the wall nodes that pass the bounds below
// in a list in LDS, examine the list's entries in parallel and reduce.
// A wall node repeated along the walk is listed again; its points come
// later with the same distances and never win.
constexpr int kWallList = 729;  // at most 9 entries per row of the 9 x 9 rows
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_wall_repulsion(const double *__restrict__ X, size_t NT,
                                 const int32_t *__restrict__ list, const int32_t *__restrict__ windex,
                                 const float *__restrict__ wdist,
                                 const int32_t *__restrict__ wnode_tri, const int32_t *__restrict__ wtri,
                                 const double *__restrict__ wxyz, const float4 *__restrict__ wpos,
                                 const float4 *__restrict__ wnrm,
                                 const float4 *__restrict__ wn_sph, const double *__restrict__ wn_geo,
                                 const double *__restrict__ wnormal,
                                 Grid g, RepParams r, double *__restrict__ vrep,
                                 int32_t *__restrict__ nearest) {
  __shared__ int ewn[kWallList], eord[kWallList];
  __shared__ float elb[kWallList];
  const int lane = threadIdx.x;
  // the reach in FP32, rounded down: the tests against it below skip only
  // wall nodes beyond the reach
  const float reach_f = (float)sqrt(r.wall_reach2) * (1.f - 1e-6f);
  // wall_search in FP32, rounded up
  float search_f = (float)r.wall_search;
  if ((double)search_f < r.wall_search) search_f = nextafterf(search_f, INFINITY);
  int lbeg, lend;
  xcd_list_span(list[0], lbeg, lend);
  for (int li = lbeg + (int)blockIdx.x / 8; li < lend; li += (int)gridDim.x / 8) {
    const int gq = list[1 + li];
    const double x = X[gq], y = X[NT + gq], z = X[2 * NT + gq];
    int lo[3], hi[3];
    window(g, x, y, z, lo, hi);
    const float xf = (float)x, yf = (float)y, zf = (float)z;
    // A coordinate rounded to FP32 is off by at most |a| 2^-24, so an FP32
    // distance from x to a point within reach is off by less than
    // 2^-22 (|x| + |y| + |z| + 3 reach); the margin is four times that
    const float smargin =
        (float)std::ldexp(fabs(x) + fabs(y) + fabs(z) + 4.0 * sqrt(r.wall_reach2), -20) + 1e-6f;
    const int nj = hi[1] - lo[1] + 1, nrow = (hi[0] - lo[0] + 1) * nj;
    // ub: an upper bound on the winning distance, the distance (rounded up)
    // to the nearest listed wall node, itself a vertex of its triangles
    float ub = INFINITY;
    int ne = 0;
    for (int r0 = 0; r0 < nrow; r0 += 64) {
      const int rw = r0 + lane;
      int wv[9];
      float lbt[9];
      unsigned keep = 0;  // the row's slots listed
      if (rw < nrow) {
        const int i = lo[0] + rw / nj, j = lo[1] + rw % nj;
        const double qx = g.x0 + (i - 1) * g.dx - x, qy = g.y0 + (j - 1) * g.dx - y;
        const double rij = qx * qx + qy * qy;
        // the row's 9 W_index entries and their distances in one batch
        // (in-bounds addresses; entries past hi[2] are not used)
        const size_t row = bq(g, i, j, 1);
        float wd[9];
#pragma unroll
        for (int t = 0; t < 9; t++) {
          const size_t at = row + (size_t)(min(lo[2] + t, hi[2]) - 1);
          wv[t] = windex[at];
          wd[t] = wdist[at];
        }
        bool cand[9];
        int last = 0;
#pragma unroll
        for (int t = 0; t < 9; t++) {
          const int wn = wv[t];
          // a wall node just seen in the row is not listed again
          cand[t] = lo[2] + t <= hi[2] && wn > 0 && wn != last;
          last = cand[t] ? wn : last;
          // wn lies wd of lattice node p = (i, j, lo[2] + t), so at least
          // |x - p| - wd from x: beyond reach (tested below exactly) for
          // certain if that exceeds the reach lowered by a margin
          const double qz = g.z0 + (lo[2] + t - 1) * g.dx - z;
          const float dp = sqrtf((float)(rij + qz * qz));
          cand[t] = cand[t] && dp - wd[t] - 1e-4f < reach_f;
        }
        float4 wp[9], wnr[9];
#pragma unroll
        for (int t = 0; t < 9; t++)
          if (cand[t]) {
            wp[t] = wpos[wv[t] - 1];
            wnr[t] = wnrm[wv[t] - 1];
          }
#pragma unroll
        for (int t = 0; t < 9; t++) {
          if (!cand[t]) continue;
          const float ex = xf - wp[t].x, ey = yf - wp[t].y, ez = zf - wp[t].z;
          const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
          // a lower bound on the distance from x to wn
          const float e = dist - smargin;
          // every vertex and the centroid of a triangle of wall node wn lie
          // within wall_lmax of it: if wn is farther than wall_search +
          // wall_lmax, none of them can beat wall_search
          if (e >= reach_f) continue;
          // they lie within wrad (rounded up) of wn, and within h of its
          // tangent plane along n: none is nearer than wall_search if the
          // larger of the two bounds reaches it
          const float dn = fabsf(wnr[t].x * ex + wnr[t].y * ey + wnr[t].z * ez) - smargin;
          const float lbf = fmaxf(e - wp[t].w, dn - wnr[t].w);
          if (lbf >= search_f) continue;
          // a node with triangles (h >= 0) is one of their vertices
          if (wnr[t].w >= 0.f) ub = fminf(ub, dist + smargin);
          lbt[t] = lbf;
          keep |= 1u << t;
        }
      }
      // the entries go to the list in row order
      const int cnt = __popc(keep);
      int incl = cnt;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        incl += lane >= d ? v : 0;
      }
      int at = ne + incl - cnt;
#pragma unroll
      for (int t = 0; t < 9; t++)
        if (keep >> t & 1u) {
          ewn[at] = wv[t];
          elb[at] = lbt[t];
          eord[at] = (rw * 9 + t) * 64;  // window position, then 4 (slot + 1) + point
          at++;
        }
      ne += __shfl(incl, 63);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ub = fminf(ub, __shfl_xor(ub, off));
    __syncthreads();
    // bound of the examination: every point nearer than the winner's
    // distance is at least as near as ub, and a winning point is nearer
    // than wall_search
    const float bnd = fminf(ub, search_f);
    double bd = r.wall_search;
    int bo = -1, bw = 0;
    for (int e = lane; e < ne; e += 64) {
      if (elb[e] > bnd) continue;
      const int wn = ewn[e];
      // the bounding spheres of the node's triangles in one batch (unused
      // slots hold a sphere at infinity)
      const float4 *sph = &wn_sph[12 * (size_t)(wn - 1)];
      float4 sp[12];
#pragma unroll
      for (int u = 0; u < 12; u++) sp[u] = sph[u];
      unsigned live = 0;
#pragma unroll
      for (int u = 0; u < 12; u++) {
        const float fx = xf - sp[u].x, fy = yf - sp[u].y, fz = zf - sp[u].z;
        live |= sqrtf(fx * fx + fy * fy + fz * fz) - smargin - sp[u].w <= bnd ? 1u << u : 0u;
      }
      const int order0 = eord[e];
#pragma unroll 1
      while (live) {
        const int u = __builtin_ctz(live);
        live &= live - 1;
        // vertices and centroid of the node's triangle u, one record
        const double *tgp = &wn_geo[12 * (12 * (size_t)(wn - 1) + u)];
        double gv[12];
#pragma unroll
        for (int v = 0; v < 12; v++) gv[v] = tgp[v];
#pragma unroll
        for (int sv = 0; sv < 4; sv++) {  // vertices 0..2, then the centroid
          const double wx = gv[3 * sv], wy = gv[3 * sv + 1], wz = gv[3 * sv + 2];
          const double d = sqrt((x - wx) * (x - wx) + (y - wy) * (y - wy) + (z - wz) * (z - wz));
          const int o = order0 + 4 * (u + 1) + sv;
          if (d < bd || (d == bd && bo >= 0 && o < bo)) {
            bd = d;
            bo = o;
            bw = wn;
          }
        }
      }
    }
    // the wave's least (distance, order)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double od = __shfl_xor(bd, off);
      const int oo = __shfl_xor(bo, off), ow = __shfl_xor(bw, off);
      if (oo >= 0 && (bo < 0 || od < bd || (od == bd && oo < bo))) {
        bd = od;
        bo = oo;
        bw = ow;
      }
    }
    if (lane == 0 && bo >= 0) {
      // the winner's triangle: slot (bo % 64) / 4 - 1 of wall node bw
      const int triwall = wnode_tri[13 * (size_t)(bw - 1) + (bo & 63) / 4];
      nearest[gq] = triwall;
      const double *w1 = &wxyz[3 * (size_t)(wtri[3 * (size_t)(triwall - 1)] - 1)];
      const double *n = &wnormal[3 * (size_t)(triwall - 1)];
      const double d = (x - w1[0]) * n[0] + (y - w1[1]) * n[1] + (z - w1[2]) * n[2];
      if (d < r.wall_d0 && fabs(d) < r.wall_dmax) {
        double f = (r.wall_d0 - d) / r.wall_dmax;
        if (d < r.wall_d0 - r.wall_dmax) f = 1.0;
        vrep[gq] += f * r.k_wall * n[0];
        vrep[NT + gq] += f * r.k_wall * n[1];
        vrep[2 * NT + gq] += f * r.k_wall * n[2];
      }
    }
    __syncthreads();
  }
}

// ---- a parent lattice's cells: the periodic window along k (zf_wrap.hip)
struct WrapCells {
  int dir;
  double z0, zm, L, mid;
};

__device__ __forceinline__ int kmap_ibm(int k, int km) {
  return k < 1 ? k + (km - 1) : k > km ? k - (km - 1) : k;
}

// Cell list of a parent (ZebraflowWT.c:5898-5950): a triangle of a cell
// not in status 2 whose centroid lies past the exit is binned from inside
// the window; window planes beyond the box wrap to k -+ (KM - 1) at their
// unwrapped place; the nearest listed node wins with no cut-off.
__global__ void k_wrap_tri_bin(const double *__restrict__ tg, size_t TT, int nt, int ncell,
                               const int32_t *__restrict__ status, const int32_t *__restrict__ b,
                               const int32_t *__restrict__ domid, Grid g, WrapCells w, int dommax,
                               int32_t *__restrict__ tbin, uint64_t *__restrict__ keys) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncell * nt) return;
  const int st = status[q / nt];
  int bin = 0;
  if (st > kStatusDead) {
    const double x = tg[4 * TT + q], y = tg[5 * TT + q];
    double z = tg[6 * TT + q];
    if (st != 2) {
      if (w.dir > 0 && z >= w.zm) z -= w.L;
      if (w.dir < 0 && z <= w.z0) z += w.L;
    }
    const int i = (int)((x - g.x0) / g.dx + 1.01), j = (int)((y - g.y0) / g.dx + 1.01),
              k = (int)((z - g.z0) / g.dx + 1.01);
    const int i1 = max(1, i - 4), i2 = min(g.im, i + 4), j1 = max(1, j - 4), j2 = min(g.jm, j + 4);
    double mind = 5000.0;
    for (int ii = i1; ii <= i2; ii++)
      for (int jj = j1; jj <= j2; jj++)
        for (int kk = k - 4; kk <= k + 4; kk++) {
          const int km = kk > g.km ? kk - g.km + 1 : kk < 1 ? kk + g.km - 1 : kk;
          const size_t bqq = bq(g, ii, jj, km);
          if (b[bqq] < 0) continue;
          const double dx = g.x0 + (ii - 1) * g.dx - x, dy = g.y0 + (jj - 1) * g.dx - y,
                       dz = g.z0 + (kk - 1) * g.dx - z;
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d < mind) {
            mind = d;
            bin = domid[bqq];
          }
        }
  }
  tbin[q] = bin;
  const uint64_t key_bin = bin > 0 ? (uint64_t)(bin - 1) : (uint64_t)dommax;
  keys[q] = (key_bin << 32) | (uint64_t)q;
}

// Cell-cell repulsion among a parent's cells (:10204-10383), one lane per
// node of a live cell.  Window planes k >= KM read plane k - KM + 1 (plane
// KM itself read as plane 1) and k < 1 read k + KM - 1 (checks 1, 2);
// otherwise a node in the window's last or first quarter is check 3 or 4.
// Across a wrapped plane a candidate's centroid moves by L (check 1: if
// behind the middle, 2: if past it); its vertices move by the quarter
// rules (checks 1, 3: first quarter +L; 2, 4: last quarter -L).
__global__ __launch_bounds__(128) void k_wrap_cell_repulsion(
    const double *__restrict__ X, size_t NT, int nn, int nt, int ncell,
    const int32_t *__restrict__ tri, const int32_t *__restrict__ status,
    const double *__restrict__ nnormal, const double *__restrict__ tg, size_t TT,
    const int32_t *__restrict__ domid, const uint64_t *__restrict__ keys,
    const int32_t *__restrict__ range, Grid g, WrapCells w, RepParams r, double *__restrict__ vrep) {
  const int gq = blockIdx.x * blockDim.x + threadIdx.x;
  if (gq >= ncell * nn) return;
  const int c = gq / nn, node = gq % nn;
  if (status[c] <= kStatusDead) return;
  const double xc = X[gq], yc = X[NT + gq], zc = X[2 * NT + gq];
  const double e1x = nnormal[gq], e1y = nnormal[NT + gq], e1z = nnormal[2 * NT + gq];
  const double q1 = w.z0 + 0.25 * w.L, q3 = w.z0 + 0.75 * w.L;
  const int ic = (int)((xc - g.x0) / g.dx + 1.01), jc = (int)((yc - g.y0) / g.dx + 1.01),
            kc = (int)((zc - g.z0) / g.dx + 1.01);
  const int i1 = max(1, ic - 4), i2 = min(g.im, ic + 4), j1 = max(1, jc - 4), j2 = min(g.jm, jc + 4);
  double v0 = 0, v1 = 0, v2 = 0;
  for (int i = i1; i <= i2; i++)
    for (int j = j1; j <= j2; j++)
      for (int k = kc - 4; k <= kc + 4; k++) {
        int check = 0, kk = k;
        if (k >= g.km) { kk = k - g.km + 1; check = 1; }
        else if (k < 1) { kk = k + g.km - 1; check = 2; }
        else if (zc >= q3) check = 3;
        else if (zc < q1) check = 4;
        const int o = domid[bq(g, i, j, kk)];
        if (o <= 0) continue;
        const int beg = range[2 * (o - 1)], end = range[2 * (o - 1) + 1];
        for (int e = beg; e < end; e++) {
          const int q = (int)(keys[e] & 0xffffffffu);
          const int m = q / nt, t = q % nt;
          if (status[m] <= kStatusDead) continue;
          const int n1 = tri[3 * t], n2 = tri[3 * t + 1], n3 = tri[3 * t + 2];
          if (m == c && (n1 == node || n2 == node || n3 == node)) continue;
          double zp = tg[6 * TT + q];
          if (check == 1 && zp < w.mid) zp += w.L;
          if (check == 2 && zp > w.mid) zp -= w.L;
          if (e1x * tg[TT + q] + e1y * tg[2 * TT + q] + e1z * tg[3 * TT + q] >= 0) continue;
          const double dx = xc - tg[4 * TT + q], dy = yc - tg[5 * TT + q], dz = zc - zp;
          if (sqrt(dx * dx + dy * dy + dz * dz) >= r.search) continue;
          const size_t b0 = (size_t)m * nn;
          double vz[3];
          const int nv[3] = {n1, n2, n3};
          for (int s = 0; s < 3; s++) {
            double z = X[2 * NT + b0 + nv[s]];
            if ((check == 1 || check == 3) && z < q1) z += w.L;
            else if ((check == 2 || check == 4) && z >= q3) z -= w.L;
            vz[s] = z;
          }
          const double ax = X[b0 + n1], ay = X[NT + b0 + n1], az = vz[0];
          const double bx = X[b0 + n2], by = X[NT + b0 + n2], bz = vz[1];
          const double cx = X[b0 + n3], cy = X[NT + b0 + n3], cz = vz[2];
          for (int sc = 0; sc < 16; sc++) {
            const double wa = (double)((kSubWA >> (4 * sc)) & 15u);
            const double wb = (double)((kSubWB >> (4 * sc)) & 15u);
            const double wc = (double)((kSubWC >> (4 * sc)) & 15u);
            const double px = (wa * ax + wb * bx + wc * cx) * (1.0 / 12.0) - xc;
            const double py = (wa * ay + wb * by + wc * cy) * (1.0 / 12.0) - yc;
            const double pz = (wa * az + wb * bz + wc * cz) * (1.0 / 12.0) - zc;
            const double d = sqrt(px * px + py * py + pz * pz);
            if (d < r.cutoff) {
              const double f = -r.k_cell * (r.cutoff - d) / r.cutoff;
              v0 += f * px / d;
              v1 += f * py / d;
              v2 += f * pz / d;
            }
          }
        }
      }
  vrep[gq] += v0;
  vrep[NT + gq] += v1;
  vrep[2 * NT + gq] += v2;
}

// Wall repulsion of a parent's cells (the interpolation loop of
// :7541-8688): the window as the interpolation's (centred at zpmap for a
// crossing cell, planes through kmap), the nearest wall triangle by vertex
// or centroid distance from the node's own position, then the plane rule.
__global__ void k_wrap_wall(const double *__restrict__ X, size_t NT, int nn, int ncell,
                            const int32_t *__restrict__ status, const int32_t *__restrict__ windex,
                            const int32_t *__restrict__ wnode_tri, const int32_t *__restrict__ wtri,
                            const double *__restrict__ wxyz, const double *__restrict__ wtgeo,
                            const double *__restrict__ wnormal, Grid g, WrapCells w, RepParams r,
                            double *__restrict__ vrep, int32_t *__restrict__ nearest) {
  const int gq = blockIdx.x * blockDim.x + threadIdx.x;
  if (gq >= ncell * nn) return;
  nearest[gq] = 0;
  const int st = status[gq / nn];
  if (st < 2) return;
  const double x = X[gq], y = X[NT + gq], z = X[2 * NT + gq];
  double ze = z;
  if (st > 2) {
    if (w.dir > 0 && z < w.mid) ze = z + w.L;
    if (w.dir < 0 && z > w.mid) ze = z - w.L;
  }
  const int i = (int)((x - g.x0) / g.dx + 1.01), j = (int)((y - g.y0) / g.dx + 1.01),
            k = (int)((ze - g.z0) / g.dx + 1.01);
  const int i1 = max(1, i - 4), i2 = min(g.im, i + 4), j1 = max(1, j - 4), j2 = min(g.jm, j + 4);
  double mind = r.wall_search;
  int triwall = 0;
  for (int ii = i1; ii <= i2; ii++)
    for (int jj = j1; jj <= j2; jj++)
      for (int kk = k - 4; kk <= k + 4; kk++) {
        const int wn = windex[bq(g, ii, jj, kmap_ibm(kk, g.km))];
        if (wn <= 0) continue;
        const int32_t *trow = &wnode_tri[13 * (size_t)(wn - 1)];
        const int cnt = min(trow[0], 12);
        for (int l = 1; l <= cnt; l++) {
          const int tw = trow[l];
          const double *gv = &wtgeo[12 * (size_t)(tw - 1)];
          for (int sv = 0; sv < 4; sv++) {
            const double wx = gv[3 * sv], wy = gv[3 * sv + 1], wz = gv[3 * sv + 2];
            const double d = sqrt((x - wx) * (x - wx) + (y - wy) * (y - wy) + (z - wz) * (z - wz));
            if (d < mind) { mind = d; triwall = tw; }
          }
        }
      }
  if (triwall == 0) return;
  nearest[gq] = triwall;
  const double *w1 = &wxyz[3 * (size_t)(wtri[3 * (size_t)(triwall - 1)] - 1)];
  const double *n = &wnormal[3 * (size_t)(triwall - 1)];
  const double d = (x - w1[0]) * n[0] + (y - w1[1]) * n[1] + (z - w1[2]) * n[2];
  if (d < r.wall_d0 && fabs(d) < r.wall_dmax) {
    double f = (r.wall_d0 - d) / r.wall_dmax;
    if (d < r.wall_d0 - r.wall_dmax) f = 1.0;
    vrep[gq] += f * r.k_wall * n[0];
    vrep[NT + gq] += f * r.k_wall * n[1];
    vrep[2 * NT + gq] += f * r.k_wall * n[2];
  }
}

// Vertices and centroid of every wall triangle as one 12-double record (the
// centroid summed as w/3 vertex by vertex, the rounding the repulsion search
// compares against), and its bounding sphere in FP32: the centroid rounded
// to FP32 and a radius, rounded up, that reaches every vertex from it.
__global__ void k_wall_geom(const double *__restrict__ wxyz, const int32_t *__restrict__ wtri,
                            int nwt, double *__restrict__ wtgeo, float4 *__restrict__ wsph) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nwt) return;
  double *o = &wtgeo[12 * (size_t)m];
  double cx = 0, cy = 0, cz = 0;
  for (int s = 0; s < 3; s++) {
    const double *w = &wxyz[3 * (size_t)(wtri[3 * (size_t)m + s] - 1)];
    o[3 * s] = w[0];
    o[3 * s + 1] = w[1];
    o[3 * s + 2] = w[2];
    cx += w[0] / 3.0;
    cy += w[1] / 3.0;
    cz += w[2] / 3.0;
  }
  o[9] = cx;
  o[10] = cy;
  o[11] = cz;
  const float fx = (float)cx, fy = (float)cy, fz = (float)cz;
  double rad = 0;
  for (int s = 0; s < 3; s++) {
    const double dx = o[3 * s] - fx, dy = o[3 * s + 1] - fy, dz = o[3 * s + 2] - fz;
    rad = fmax(rad, sqrt(dx * dx + dy * dy + dz * dz));
  }
  // the centroid point itself lies |c - fc| from the FP32 centre
  const double dc = sqrt((cx - fx) * (cx - fx) + (cy - fy) * (cy - fy) + (cz - fz) * (cz - fz));
  rad = fmax(rad, dc) * (1.0 + 1e-12) + 1e-12;
  float rf = (float)rad;
  if ((double)rf < rad) rf = nextafterf(rf, INFINITY);
  wsph[m] = make_float4(fx, fy, fz, rf);
}

// Per wall node, its position in FP32 and, rounded up, the largest distance
// from it to its triangles' vertices and centroids (wrad) plus the distance
// from its FP32 position to its exact one; and a unit vector n in FP32 (the
// normalised sum of its triangles' normals) with, rounded up, the largest
// |n . (p - w)| over those points p (w its FP32 position).  For any unit n,
// |x - p| >= |n . (x - w)| - |n . (p - w)|: on a wall that is nearly flat
// around the node this bounds the distance to its triangles far more
// tightly than |x - w| - wrad.
__global__ void k_wall_pos(const double *__restrict__ wxyz, const double *__restrict__ wrad, int nwn,
                           const int32_t *__restrict__ wnode_tri, const double *__restrict__ wnormal,
                           const double *__restrict__ wtgeo, float4 *__restrict__ wpos,
                           float4 *__restrict__ wnrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwn) return;
  const double x = wxyz[3 * (size_t)i], y = wxyz[3 * (size_t)i + 1], z = wxyz[3 * (size_t)i + 2];
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  const double off = sqrt((x - fx) * (x - fx) + (y - fy) * (y - fy) + (z - fz) * (z - fz));
  const double rad = wrad[i] + off;
  float rf = (float)rad;
  if ((double)rf < rad) rf = nextafterf(rf, INFINITY);
  wpos[i] = make_float4(fx, fy, fz, rf);
  const int32_t *row = &wnode_tri[13 * (size_t)i];
  const int cnt = min(row[0], 12);
  double sx = 0, sy = 0, sz = 0;
  for (int u = 0; u < cnt; u++) {
    const double *n = &wnormal[3 * (size_t)(row[1 + u] - 1)];
    sx += n[0];
    sy += n[1];
    sz += n[2];
  }
  const double len = sqrt(sx * sx + sy * sy + sz * sz);
  if (!(len > 1e-6)) {  // no usable normal: fall back to |n . (p - w)| <= |p - w| <= wrad
    wnrm[i] = make_float4(1.f, 0.f, 0.f, rf);
    return;
  }
  // the FP32 unit vector, shortened by a hair so that |n| <= 1 however it
  // rounded (the bound needs |n| <= 1)
  const float nx = (float)(sx / len * (1.0 - 1e-6)), ny = (float)(sy / len * (1.0 - 1e-6)),
              nz = (float)(sz / len * (1.0 - 1e-6));
  double h = 0;
  for (int u = 0; u < cnt; u++) {
    const double *tgp = &wtgeo[12 * (size_t)(row[1 + u] - 1)];
    for (int v = 0; v < 4; v++)
      h = fmax(h, fabs((double)nx * (tgp[3 * v] - fx) + (double)ny * (tgp[3 * v + 1] - fy) +
                       (double)nz * (tgp[3 * v + 2] - fz)));
  }
  h = h * (1.0 + 1e-9) + 1e-9;
  float hf = (float)h;
  if ((double)hf < h) hf = nextafterf(hf, INFINITY);
  wnrm[i] = make_float4(nx, ny, nz, hf);
}

// Per wall node, for each of its (up to 12) triangle slots in row order, the
// triangle's bounding sphere and its 12-double record, so that the
// wall-repulsion search reads them by (node, slot) with no load of the
// triangle number first; unused slots hold a sphere at infinity.
__global__ void k_wall_node_recs(const int32_t *__restrict__ wnode_tri, int nwn,
                                 const double *__restrict__ wtgeo, const float4 *__restrict__ wsph,
                                 float4 *__restrict__ wn_sph, double *__restrict__ wn_geo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 12 * nwn) return;
  const int wn = i / 12, u = i % 12;
  const int32_t *row = &wnode_tri[13 * (size_t)wn];
  if (u >= min(row[0], 12)) {
    wn_sph[i] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
    return;
  }
  const int tw = row[1 + u];
  wn_sph[i] = wsph[tw - 1];
  for (int v = 0; v < 12; v++) wn_geo[12 * (size_t)i + v] = wtgeo[12 * (size_t)(tw - 1) + v];
}

// W_index: every lattice node in a wall node's window keeps the nearest wall
// node, compared as (FP32 distance, node) through one 64-bit atomicMin, which
// is the reference's serial strict-< rule over FP32 distances (ties to the
// lower wall node).
__global__ void k_windex(const double *__restrict__ wxyz, int nwn, Grid g,
                         unsigned long long *__restrict__ key) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nwn) return;
  const double wx = wxyz[3 * (size_t)l], wy = wxyz[3 * (size_t)l + 1], wz = wxyz[3 * (size_t)l + 2];
  int lo[3], hi[3];
  window(g, wx, wy, wz, lo, hi);
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int k = lo[2]; k <= hi[2]; k++) {
        const double dx = wx - (g.x0 + (i - 1) * g.dx), dy = wy - (g.y0 + (j - 1) * g.dx),
                     dz = wz - (g.z0 + (k - 1) * g.dx);
        const float d = (float)sqrt(dx * dx + dy * dy + dz * dz);
        if (!(d < 5000.0f)) continue;
        const unsigned long long kv =
            ((unsigned long long)__float_as_uint(d) << 32) | (unsigned long long)(l + 1);
        atomicMin(&key[bq(g, i, j, k)], kv);
      }
}

__global__ void k_windex_final(const unsigned long long *__restrict__ key, size_t box,
                               int32_t *__restrict__ windex, float *__restrict__ wdist) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= box) return;
  const unsigned long long kv = key[q];
  windex[q] = kv == ~0ull ? 0 : (int32_t)(kv & 0xffffffffu);
  wdist[q] = kv == ~0ull ? __builtin_inff() : __uint_as_float((unsigned)(kv >> 32));
}

// Wall-layer velocity Uvoid at box node q (ZebraflowWT.c:4788-4851), zero
// where the node does not carry one.  FP32 distances as the reference's
// float coordinates give them (wall node coordinates and centroids in FP32,
// the centroid (x1 + x2 + x3)/3 as :4660 forms it), strict < in triangle
// order, vertices before the centroid.
struct VoidParams {
  double speed, d0, d1;
  int k_lo, k_hi, select;
  float px, py, pz;
};

__device__ bool uvoid_at(size_t q, const Grid &g, const int32_t *__restrict__ b,
                         const int32_t *__restrict__ windex, const float *__restrict__ wdist,
                         const double *__restrict__ wxyz, const int32_t *__restrict__ wtri,
                         const double *__restrict__ wnormal, const int32_t *__restrict__ wnode_tri,
                         const VoidParams &vp, double u[3]) {
  const int k = (int)(q % g.km) + 1;
  const int j = (int)((q / g.km) % g.jm) + 1;
  const int i = (int)(q / ((size_t)g.km * g.jm)) + 1;
  const int nw = windex[q];
  if (nw <= 0 || b[q] > 0 || k < vp.k_lo || k > g.km - vp.k_hi) return false;
  float px = vp.px, py = vp.py, pz = vp.pz;
  if (!vp.select) {
    px = (float)(g.x0 + (i - 1) * g.dx);
    py = (float)(g.y0 + (j - 1) * g.dx);
    pz = (float)(g.z0 + (k - 1) * g.dx);
  }
  const int32_t *row = &wnode_tri[13 * (size_t)(nw - 1)];
  const int cnt = min(row[0], 12);
  float mindis = 5000.0f;
  int triwall = 0;
  for (int l = 1; l <= cnt; l++) {
    const int tri = row[l];
    float cx = 0, cy = 0, cz = 0;
    for (int sv = 0; sv < 3; sv++) {
      const double *w = &wxyz[3 * (size_t)(wtri[3 * (size_t)(tri - 1) + sv] - 1)];
      const float x = (float)w[0], y = (float)w[1], z = (float)w[2];
      const float dx = px - x, dy = py - y, dz = pz - z;
      const float d = sqrtf(dx * dx + dy * dy + dz * dz);
      if (d < mindis) { mindis = d; triwall = tri; }
      if (sv == 0) { cx = x; cy = y; cz = z; }
      else { cx += x; cy += y; cz += z; }
    }
    cx = (float)((double)cx / 3.0);
    cy = (float)((double)cy / 3.0);
    cz = (float)((double)cz / 3.0);
    const float dx = px - cx, dy = py - cy, dz = pz - cz;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    if (d < mindis) { mindis = d; triwall = tri; }
  }
  if (triwall == 0) return false;
  const double wd = (double)wdist[q];
  const double w1 = wd <= vp.d0 ? 0.0 : wd <= vp.d1 ? (wd - vp.d0) / (vp.d1 - vp.d0) : 1.0;
  if (w1 == 0.0) return false;
  const double *n = &wnormal[3 * (size_t)(triwall - 1)];
  u[0] = w1 * vp.speed * n[0];
  u[1] = w1 * vp.speed * n[1];
  u[2] = w1 * vp.speed * n[2];
  return true;
}

__global__ void k_uvoid_flag(size_t box, Grid g, const int32_t *__restrict__ b,
                             const int32_t *__restrict__ windex, const float *__restrict__ wdist,
                             const double *__restrict__ wxyz, const int32_t *__restrict__ wtri,
                             const double *__restrict__ wnormal,
                             const int32_t *__restrict__ wnode_tri, VoidParams vp,
                             int32_t *__restrict__ flag) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= box) {
    if (q == box) flag[q] = 0;  // the scan's last element: the total lands in slot[box]
    return;
  }
  double u[3];
  flag[q] = uvoid_at(q, g, b, windex, wdist, wxyz, wtri, wnormal, wnode_tri, vp, u) ? 1 : 0;
}

// slot = exclusive prefix count of the flags, so the slots follow box order
__global__ void k_uvoid_fill(size_t box, Grid g, const int32_t *__restrict__ b,
                             const int32_t *__restrict__ windex, const float *__restrict__ wdist,
                             const double *__restrict__ wxyz, const int32_t *__restrict__ wtri,
                             const double *__restrict__ wnormal,
                             const int32_t *__restrict__ wnode_tri, VoidParams vp,
                             int32_t *__restrict__ slot_io, double *__restrict__ uvoid) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= box) return;
  double u[3];
  if (uvoid_at(q, g, b, windex, wdist, wxyz, wtri, wnormal, wnode_tri, vp, u)) {
    const int sl = slot_io[q];
    uvoid[3 * (size_t)sl] = u[0];
    uvoid[3 * (size_t)sl + 1] = u[1];
    uvoid[3 * (size_t)sl + 2] = u[2];
  } else {
    slot_io[q] = -1;
  }
}

__global__ void k_uvoid_export(size_t box, const int32_t *__restrict__ slot,
                               const double *__restrict__ uvoid, double *__restrict__ out) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= box) return;
  const int sl = slot[q];
  for (int a = 0; a < 3; a++) out[(size_t)a * box + q] = sl >= 0 ? uvoid[3 * (size_t)sl + a] : 0.0;
}

__global__ void k_fill_u64(unsigned long long *a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = ~0ull;
}

int require_rep(zf_ctx *ctx) {
  if (!ctx->have_rep) return fail(ctx, ZF_ERR_STATE, "zf_repulsion_setup not called");
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "zf_ibm_setup not called");
  return ZF_OK;
}

void free_void(zf_ctx *ctx) {
  dev_free(ctx, (void **)&ctx->d_vslot);
  dev_free(ctx, (void **)&ctx->d_uvoid);
  ctx->have_void = false;
  ctx->n_void = 0;
}

void free_wall(zf_ctx *ctx) {
  void **bufs[] = {(void **)&ctx->d_wxyz, (void **)&ctx->d_wnormal, (void **)&ctx->d_wtri,
                   (void **)&ctx->d_wnode_tri, (void **)&ctx->d_windex, (void **)&ctx->d_wdist,
                   (void **)&ctx->d_wtgeo, (void **)&ctx->d_wsph, (void **)&ctx->d_wn_sph,
                   (void **)&ctx->d_wn_geo, (void **)&ctx->d_wpos, (void **)&ctx->d_wnrm, (void **)&ctx->d_wrad};
  for (auto b : bufs) dev_free(ctx, b);
  free_void(ctx);
  ctx->have_wall = false;
}

}  // namespace

void free_cells(zf_ctx *ctx) {
  void **bufs[] = {(void **)&ctx->d_tkeys[0], (void **)&ctx->d_tkeys[1], (void **)&ctx->d_crec,
                   (void **)&ctx->d_bin_cnt, (void **)&ctx->d_bin_off, &ctx->d_bin_scan_tmp,
                   (void **)&ctx->d_npair, (void **)&ctx->d_pairs, (void **)&ctx->d_over, (void **)&ctx->d_rep_lists, &ctx->d_rsel_tmp,
                   (void **)&ctx->d_ochunk, (void **)&ctx->d_rep_pool,
                   (void **)&ctx->d_tsort_tmp,
                   (void **)&ctx->d_tbin, (void **)&ctx->d_bin_range, (void **)&ctx->d_near,
                   (void **)&ctx->d_rep_bounds, (void **)&ctx->d_row_bd};
  for (auto b : bufs) dev_free(ctx, b);
  free_wall(ctx);
  ctx->have_rep = false;
}

// CSR offsets of the cell list: the first key of every bin (empty bins
// included) and the key count past the last bin.
static int bin_offsets(zf_ctx *ctx) {
  const int nb1 = ctx->dommax + 1;
  hipLaunchKernelGGL(k_bin_counts, dim3(nb(nb1, 256)), dim3(256), 0, ctx->stream, ctx->d_bin_range,
                     ctx->dommax, ctx->d_bin_cnt);
  size_t tmp = ctx->bin_scan_tmp_bytes;
  ZF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(ctx->d_bin_scan_tmp, tmp, ctx->d_bin_cnt, ctx->d_bin_off,
                                               nb1, ctx->stream));
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int cell_list_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, TT = ncap * ctx->nt;
  const int n = ctx->ncell * ctx->nt;
  hipLaunchKernelGGL(k_fill, dim3(nb(2 * (size_t)ctx->dommax, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_bin_range, 2 * (size_t)ctx->dommax, 0);
  if (n == 0) return bin_offsets(ctx);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, ctx->stream, (int32_t *)ctx->d_rep_bounds,
                     (size_t)4, 0);
  const size_t nrows = (size_t)ctx->cfg.im * ctx->cfg.jm;
  hipLaunchKernelGGL(k_fill, dim3(nb(nrows, 256)), dim3(256), 0, ctx->stream, (int32_t *)ctx->d_row_bd,
                     nrows, 0);
  const size_t NT = ncap * ctx->nn;
  hipLaunchKernelGGL(k_tri_bin, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_tgeo, TT,
                     ctx->d_x, NT, ctx->nn, ctx->d_tri, ctx->nt, ctx->ncell, ctx->d_status,
                     ctx->d_b, ctx->d_domid, grid_of(ctx), ctx->rcfg.bin_cutoff, ctx->dommax,
                     ctx->d_tbin, ctx->d_tkeys[0], ctx->d_rep_bounds, ctx->d_row_bd);
  ctx->bins_gen = ctx->geom_gen;
  int bits = 1;
  while ((1ull << bits) <= (uint64_t)ctx->dommax) bits++;
  size_t tmp = ctx->tsort_tmp_bytes;
  // k_tri_bin writes the key of triangle q at position q, so the keys come
  // in triangle order and a (stable) radix sort of the bin bits alone gives
  // the (bin, triangle) order
  ZF_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->d_tsort_tmp, tmp, ctx->d_tkeys[0],
                                                  ctx->d_tkeys[1], n, 32, 32 + bits, ctx->stream));
  hipLaunchKernelGGL(k_bin_ranges, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_tkeys[1], n,
                     (uint64_t)ctx->dommax, ctx->d_bin_range);
  return bin_offsets(ctx);
}

namespace {
WrapCells wrap_cells(const zf_ctx *ctx) {
  return WrapCells{ctx->wcfg.direction, ctx->wz0, ctx->wzm, ctx->wzm - ctx->wz0,
                   0.5 * (ctx->wz0 + ctx->wzm)};
}

RepParams rep_params_of(const zf_ctx *ctx) {
  return RepParams{ctx->rcfg.k_cell, ctx->rcfg.cutoff, ctx->rcfg.search, ctx->rcfg.k_wall,
                   ctx->rcfg.wall_d0, ctx->rcfg.wall_dmax, ctx->rcfg.wall_search, 0.0, 0};
}
}  // namespace

int wrap_cell_list_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, TT = ncap * ctx->nt;
  const int n = ctx->ncell * ctx->nt;
  hipLaunchKernelGGL(k_fill, dim3(nb(2 * (size_t)ctx->dommax, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_bin_range, 2 * (size_t)ctx->dommax, 0);
  if (n == 0) return ZF_OK;
  hipLaunchKernelGGL(k_wrap_tri_bin, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_tgeo, TT,
                     ctx->nt, ctx->ncell, ctx->d_status, ctx->d_b, ctx->d_domid, grid_of(ctx),
                     wrap_cells(ctx), ctx->dommax, ctx->d_tbin, ctx->d_tkeys[0]);
  int bits = 1;
  while ((1ull << bits) <= (uint64_t)ctx->dommax) bits++;
  size_t tmp = ctx->tsort_tmp_bytes;
  ZF_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->d_tsort_tmp, tmp, ctx->d_tkeys[0],
                                                  ctx->d_tkeys[1], n, 0, 32 + bits, ctx->stream));
  hipLaunchKernelGGL(k_bin_ranges, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_tkeys[1], n,
                     (uint64_t)ctx->dommax, ctx->d_bin_range);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int wrap_wall_impl(zf_ctx *ctx) {
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0 || !ctx->have_wall) return ZF_OK;
  hipLaunchKernelGGL(k_wrap_wall, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x, NT, ctx->nn,
                     ctx->ncell, ctx->d_status, ctx->d_windex, ctx->d_wnode_tri, ctx->d_wtri,
                     ctx->d_wxyz, ctx->d_wtgeo, ctx->d_wnormal, grid_of(ctx), wrap_cells(ctx),
                     rep_params_of(ctx), ctx->d_vrep, ctx->d_near);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int wrap_repulsion_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn, TT = ncap * ctx->nt;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  hipLaunchKernelGGL(k_wrap_cell_repulsion, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x, NT,
                     ctx->nn, ctx->nt, ctx->ncell, ctx->d_tri, ctx->d_status, ctx->d_nnormal,
                     ctx->d_tgeo, TT, ctx->d_domid, ctx->d_tkeys[1], ctx->d_bin_range, grid_of(ctx),
                     wrap_cells(ctx), rep_params_of(ctx), ctx->d_vrep);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int repulsion_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn, TT = ncap * ctx->nt;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  const double wall_reach = ctx->rcfg.wall_search + ctx->wall_lmax + kReachSlack;
  const RepParams r{ctx->rcfg.k_cell, ctx->rcfg.cutoff, ctx->rcfg.search, ctx->rcfg.k_wall,
                    ctx->rcfg.wall_d0, ctx->rcfg.wall_dmax, ctx->rcfg.wall_search,
                    wall_reach * wall_reach, ctx->bins_gen == ctx->geom_gen ? 1 : 0};
  hipLaunchKernelGGL(k_fill, dim3(nb(3 * NT * 2, 256)), dim3(256), 0, ctx->stream,
                     (int32_t *)ctx->d_vrep, 3 * NT * 2, 0);
  hipLaunchKernelGGL(k_fill, dim3(nb(NT, 256)), dim3(256), 0, ctx->stream, ctx->d_near, NT, 0);
  if (ctx->have_wall) {
    // the nodes near the wall, listed in node order, then their walk (the
    // flags and the list reuse the cell-cell walk's counts and overflow
    // list, both written again below)
    hipLaunchKernelGGL(k_wall_select, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_x, NT, ctx->nn,
                       ctx->ncell, ctx->d_status, ctx->d_wdist, grid_of(ctx), r, ctx->d_npair);
    size_t stmp = ctx->rsel_tmp_bytes;
    ZF_HIP(ctx, hipcub::DeviceSelect::Flagged(ctx->d_rsel_tmp, stmp, hipcub::CountingInputIterator<int32_t>(0),
                                              ctx->d_npair, ctx->d_over + 1, ctx->d_over, n, ctx->stream));
    hipLaunchKernelGGL(k_wall_repulsion, dim3(std::min<unsigned>((nb(n, 64) + 7) / 8 * 8, 16384u)), dim3(64), 0,
                       ctx->stream, ctx->d_x, NT, ctx->d_over, ctx->d_windex, ctx->d_wdist,
                       ctx->d_wnode_tri, ctx->d_wtri, ctx->d_wxyz, ctx->d_wpos, ctx->d_wnrm, ctx->d_wn_sph,
                       ctx->d_wn_geo, ctx->d_wnormal, grid_of(ctx), r, ctx->d_vrep, ctx->d_near);
  }
  const int nkeys = ctx->ncell * ctx->nt;
  hipLaunchKernelGGL(k_cand_records, dim3(nb(nkeys, 256)), dim3(256), 0, ctx->stream, ctx->d_tkeys[1],
                     nkeys, ctx->d_tgeo, TT, ctx->d_crec);
  // walk (lists each node's passing triangles), evaluation of the lists,
  // walk and evaluation again for the nodes whose list overflowed
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, ctx->stream, ctx->d_over, (size_t)1, 0);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, ctx->stream, ctx->d_rep_lists, (size_t)kRepListHead, 0);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, ctx->stream, ctx->d_rep_pool, (size_t)1, 0);
  const int pool_cap = rep_pool_chunks(NT);
  hipLaunchKernelGGL(k_cell_repulsion<1>, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x, NT,
                     ctx->nn, ctx->nt, ctx->ncell, ctx->d_tri, ctx->d_status, ctx->d_nnormal,
                     ctx->d_tgeo, TT, ctx->d_domid, ctx->d_crec, ctx->d_tkeys[1], ctx->d_bin_off,
                     ctx->d_rep_bounds, ctx->d_row_bd, grid_of(ctx), r, ctx->d_npair, ctx->d_pairs, ctx->d_ochunk,
                     ctx->d_rep_pool, pool_cap, ctx->d_rep_lists, ctx->d_over, ctx->d_vrep);
  hipLaunchKernelGGL(k_rep_eval, dim3(std::min<unsigned>(nb(n, 128), 2048u)), dim3(128), 0, ctx->stream, ctx->d_x,
                     NT, ctx->nn, ctx->nt, ctx->d_tri, ctx->d_npair, ctx->d_pairs, ctx->d_ochunk,
                     ctx->d_rep_pool, ctx->d_rep_lists, r,
                     ctx->d_vrep);
  hipLaunchKernelGGL(k_rep_eval_wide, dim3(std::min<unsigned>(nb(n, 64), 4096u)), dim3(64), 0, ctx->stream,
                     ctx->d_x, NT, ctx->nn, ctx->nt, ctx->d_tri, ctx->d_npair, ctx->d_pairs, ctx->d_ochunk,
                     ctx->d_rep_pool, ctx->d_rep_lists, r, ctx->d_vrep);
  hipLaunchKernelGGL(k_cell_repulsion<0>, dim3(std::min<unsigned>(nb(n, 128), 256u)), dim3(128), 0, ctx->stream,
                     ctx->d_x, NT, ctx->nn, ctx->nt, ctx->ncell, ctx->d_tri, ctx->d_status,
                     ctx->d_nnormal, ctx->d_tgeo, TT, ctx->d_domid, ctx->d_crec, ctx->d_tkeys[1],
                     ctx->d_bin_off, ctx->d_rep_bounds, ctx->d_row_bd, grid_of(ctx), r, ctx->d_npair, ctx->d_pairs,
                     ctx->d_ochunk, ctx->d_rep_pool, pool_cap, ctx->d_rep_lists, ctx->d_over, ctx->d_vrep);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

}  // namespace zf

using namespace zf;

extern "C" {

void zf_repulsion_config_reference(zf_repulsion_config *rc) {
  rc->bin_cutoff = 1.0;
  rc->k_cell = 8.333e-7;
  rc->cutoff = 1.0;
  rc->search = 2.0;
  rc->k_wall = 1e-4;
  rc->wall_d0 = 0.25;
  rc->wall_dmax = 1.0;
  rc->wall_search = 1.5;
}

int zf_repulsion_setup(zf_ctx *ctx, const zf_repulsion_config *rc) {
  if (!ctx || !rc) return ZF_ERR_ARG;
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "call zf_ibm_setup first");
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  free_cells(ctx);
  ctx->rcfg = *rc;
  const size_t ncap = ctx->mcfg.max_cells, TT = ncap * ctx->nt, NT = ncap * ctx->nn;
  int r;
  if ((r = dev_alloc(ctx, (void **)&ctx->d_tkeys[0], 8 * TT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_tkeys[1], 8 * TT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_crec, sizeof(float4) * TT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_tbin, 4 * TT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_bin_range, 8 * (size_t)std::max(ctx->dommax, 1))) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_bin_cnt, 4 * ((size_t)ctx->dommax + 1))) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_bin_off, 4 * ((size_t)ctx->dommax + 1))) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_near, 4 * NT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_npair, 4 * NT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_pairs, 4 * (size_t)kRepSlots * NT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_ochunk, 4 * NT)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_rep_pool,
                     4 * (kRepPoolHead + (size_t)kRepChunk * rep_pool_chunks(NT)))) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_over, 4 * (NT + 1))) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_rep_lists, 4 * (kRepListHead + kRepClasses * NT))) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_row_bd, 4 * (size_t)ctx->cfg.im * ctx->cfg.jm)) ||
      (r = dev_alloc(ctx, (void **)&ctx->d_rep_bounds, 2 * sizeof(unsigned long long))))
    return r;
  ctx->bins_gen = ~0ull;
  hipLaunchKernelGGL(k_fill, dim3(nb(2 * (size_t)std::max(ctx->dommax, 1), 256)), dim3(256), 0,
                     ctx->stream, ctx->d_bin_range, 2 * (size_t)std::max(ctx->dommax, 1), 0);
  size_t tmp = 0;
  ZF_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, ctx->d_tkeys[0], ctx->d_tkeys[1],
                                                  (int)TT, 0, 64, ctx->stream));
  ctx->tsort_tmp_bytes = tmp;
  if ((r = dev_alloc(ctx, &ctx->d_tsort_tmp, tmp))) return r;
  size_t stmp = 0;
  ZF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, stmp, ctx->d_bin_cnt, ctx->d_bin_off,
                                               ctx->dommax + 1, ctx->stream));
  ctx->bin_scan_tmp_bytes = stmp;
  if ((r = dev_alloc(ctx, &ctx->d_bin_scan_tmp, std::max<size_t>(stmp, 1)))) return r;
  size_t wtmp = 0;
  ZF_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, wtmp, hipcub::CountingInputIterator<int32_t>(0),
                                            ctx->d_npair, ctx->d_over + 1, ctx->d_over, (int)NT,
                                            ctx->stream));
  ctx->rsel_tmp_bytes = wtmp;
  if ((r = dev_alloc(ctx, &ctx->d_rsel_tmp, std::max<size_t>(wtmp, 1)))) return r;
  if ((r = bin_offsets(ctx))) return r;
  ctx->have_rep = true;
  return ZF_OK;
}

int zf_wall_setup(zf_ctx *ctx, int32_t nwn, const double *wall_xyz, int32_t nwt,
                  const int32_t *wall_tri, const double *wall_normal) {
  if (!ctx || nwn <= 0 || nwt <= 0 || !wall_xyz || !wall_tri || !wall_normal) return ZF_ERR_ARG;
  int rc = require_rep(ctx);
  if (rc) return rc;
  for (size_t i = 0; i < 3 * (size_t)nwt; i++)
    if (wall_tri[i] < 1 || wall_tri[i] > nwn) return fail(ctx, ZF_ERR_ARG, "wall triangle node out of range");
  free_wall(ctx);
  // node -> triangle rows (count, up to 12 triangles), ZebraflowWT.c:3900-3935
  std::vector<int32_t> rows(13 * (size_t)nwn, 0);
  for (int m = 1; m <= nwt; m++)
    for (int s = 0; s < 3; s++) {
      int32_t *row = &rows[13 * (size_t)(wall_tri[3 * (size_t)(m - 1) + s] - 1)];
      row[0]++;
      if (row[0] <= 12) row[row[0]] = m;
    }
  // largest distance from a wall triangle's vertex to its other vertices
  // and its centroid (the reach of a wall node's triangles), overall and
  // per wall node (the latter with a relative and absolute slack that
  // covers the rounding of the device's centroid and distances)
  double lmax = 0;
  std::vector<double> wrad(nwn, 0.0);
  for (int m = 0; m < nwt; m++) {
    const double *v[3];
    for (int s = 0; s < 3; s++) v[s] = &wall_xyz[3 * (size_t)(wall_tri[3 * (size_t)m + s] - 1)];
    double cen[3];
    for (int a = 0; a < 3; a++) cen[a] = (v[0][a] + v[1][a] + v[2][a]) / 3.0;
    for (int s = 0; s < 3; s++) {
      const double *pts[3] = {v[(s + 1) % 3], v[(s + 2) % 3], cen};
      double &rv = wrad[wall_tri[3 * (size_t)m + s] - 1];
      for (const double *p : pts) {
        const double dx = v[s][0] - p[0], dy = v[s][1] - p[1], dz = v[s][2] - p[2];
        const double dd = std::sqrt(dx * dx + dy * dy + dz * dz);
        lmax = std::max(lmax, dd);
        rv = std::max(rv, dd);
      }
    }
  }
  unsigned long long *key = nullptr;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_wxyz, 24 * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wnormal, 24 * (size_t)nwt)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wtri, 12 * (size_t)nwt)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wtgeo, 96 * (size_t)nwt)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wsph, sizeof(float4) * (size_t)nwt)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wn_sph, sizeof(float4) * 12 * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wpos, sizeof(float4) * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wnrm, sizeof(float4) * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wn_geo, 8 * 144 * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wrad, 8 * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wnode_tri, 52 * (size_t)nwn)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_windex, 4 * ctx->box)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_wdist, 4 * ctx->box)) ||
      (rc = dev_alloc(ctx, (void **)&key, 8 * ctx->box)))
    return rc;
  ZF_HIP(ctx, hipMemcpy(ctx->d_wxyz, wall_xyz, 24 * (size_t)nwn, hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_wnormal, wall_normal, 24 * (size_t)nwt, hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_wtri, wall_tri, 12 * (size_t)nwt, hipMemcpyHostToDevice));
  ZF_HIP(ctx, hipMemcpy(ctx->d_wnode_tri, rows.data(), 52 * (size_t)nwn, hipMemcpyHostToDevice));
  for (double &rv : wrad) rv = rv * (1.0 + 1e-12) + 1e-12;
  ZF_HIP(ctx, hipMemcpy(ctx->d_wrad, wrad.data(), 8 * (size_t)nwn, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_wall_geom, dim3(nb(nwt, 256)), dim3(256), 0, ctx->stream, ctx->d_wxyz,
                     ctx->d_wtri, nwt, ctx->d_wtgeo, ctx->d_wsph);
  hipLaunchKernelGGL(k_wall_pos, dim3(nb(nwn, 256)), dim3(256), 0, ctx->stream, ctx->d_wxyz, ctx->d_wrad,
                     nwn, ctx->d_wnode_tri, ctx->d_wnormal, ctx->d_wtgeo, ctx->d_wpos, ctx->d_wnrm);
  hipLaunchKernelGGL(k_wall_node_recs, dim3(nb(12 * (size_t)nwn, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_wnode_tri, nwn, ctx->d_wtgeo, ctx->d_wsph, ctx->d_wn_sph, ctx->d_wn_geo);
  hipLaunchKernelGGL(k_fill_u64, dim3(nb(ctx->box, 256)), dim3(256), 0, ctx->stream, key, ctx->box);
  hipLaunchKernelGGL(k_windex, dim3(nb(nwn, 64)), dim3(64), 0, ctx->stream, ctx->d_wxyz, nwn,
                     grid_of(ctx), key);
  hipLaunchKernelGGL(k_windex_final, dim3(nb(ctx->box, 256)), dim3(256), 0, ctx->stream, key,
                     ctx->box, ctx->d_windex, ctx->d_wdist);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= 8 * (int64_t)ctx->box;
  dev_free(ctx, (void **)&key);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "wall setup: %s", hipGetErrorString(e));
  ctx->nwn = nwn;
  ctx->nwt = nwt;
  ctx->wall_lmax = lmax;
  ctx->have_wall = true;
  return ZF_OK;
}

void zf_uvoid_config_reference(zf_uvoid_config *uc) {
  if (!uc) return;
  *uc = zf_uvoid_config{};
  uc->speed = 1e-3 / 1.0;  // 1e-3 m/s over VEL_SCALE = L_SCALE/T_SCALE = 1 m/s
  uc->d0 = 0.7;
  uc->d1 = 3.0;
  uc->k_lo = 3;
  uc->k_hi = 2;
  uc->select = 1;
  uc->select_point[0] = 78.5;
  uc->select_point[1] = 214.5;
  uc->select_point[2] = 392.5;
}

int zf_uvoid_setup(zf_ctx *ctx, const zf_uvoid_config *uc) {
  if (!ctx) return ZF_ERR_ARG;
  if (!uc) {
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_void(ctx);
    return ZF_OK;
  }
  if (!ctx->have_wall) return fail(ctx, ZF_ERR_STATE, "zf_wall_setup not called");
  if (!(uc->d1 > uc->d0) || uc->k_lo < 1 || uc->k_hi < 0 || (uc->select != 0 && uc->select != 1))
    return fail(ctx, ZF_ERR_ARG, "uvoid configuration");
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  free_void(ctx);
  const VoidParams vp{uc->speed, uc->d0, uc->d1, uc->k_lo, uc->k_hi, uc->select,
                      (float)uc->select_point[0], (float)uc->select_point[1],
                      (float)uc->select_point[2]};
  const Grid g = grid_of(ctx);
  const size_t box = ctx->box;
  int32_t *flag = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc;
  if ((rc = dev_alloc(ctx, (void **)&flag, 4 * (box + 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_vslot, 4 * (box + 1))))
    return rc;
  hipLaunchKernelGGL(k_uvoid_flag, dim3(nb(box + 1, 256)), dim3(256), 0, ctx->stream, box, g, ctx->d_b,
                     ctx->d_windex, ctx->d_wdist, ctx->d_wxyz, ctx->d_wtri, ctx->d_wnormal,
                     ctx->d_wnode_tri, vp, flag);
  ZF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, flag, ctx->d_vslot, (int)(box + 1),
                                               ctx->stream));
  if ((rc = dev_alloc(ctx, &tmp, tmp_bytes))) {
    dev_free(ctx, (void **)&flag);
    return rc;
  }
  ZF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flag, ctx->d_vslot, (int)(box + 1),
                                               ctx->stream));
  int32_t total = 0;
  ZF_HIP(ctx, hipMemcpyAsync(&total, ctx->d_vslot + box, 4, hipMemcpyDeviceToHost, ctx->stream));
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, &tmp);
  dev_free(ctx, (void **)&flag);
  ctx->device_bytes -= (int64_t)(4 * (box + 1) + tmp_bytes);
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_uvoid, 24 * (size_t)std::max(total, 1)))) return rc;
  hipLaunchKernelGGL(k_uvoid_fill, dim3(nb(box, 256)), dim3(256), 0, ctx->stream, box, g, ctx->d_b,
                     ctx->d_windex, ctx->d_wdist, ctx->d_wxyz, ctx->d_wtri, ctx->d_wnormal,
                     ctx->d_wnode_tri, vp, ctx->d_vslot, ctx->d_uvoid);
  ZF_HIP(ctx, hipGetLastError());
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->n_void = total;
  ctx->have_void = true;
  return ZF_OK;
}

int zf_download_uvoid(zf_ctx *ctx, double *uvw) {
  if (!ctx || !uvw) return ZF_ERR_ARG;
  const size_t box = ctx->box;
  if (!ctx->have_void) {
    for (size_t i = 0; i < 3 * box; i++) uvw[i] = 0.0;
    return ZF_OK;
  }
  double *tmp = nullptr;
  int rc = dev_alloc(ctx, (void **)&tmp, 24 * box);
  if (rc) return rc;
  hipLaunchKernelGGL(k_uvoid_export, dim3(nb(box, 256)), dim3(256), 0, ctx->stream, box,
                     ctx->d_vslot, ctx->d_uvoid, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(uvw, tmp, 24 * box, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= (int64_t)(24 * box);
  dev_free(ctx, (void **)&tmp);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "uvoid download: %s", hipGetErrorString(e));
  return ZF_OK;
}

int zf_download_windex(zf_ctx *ctx, int32_t *windex) {
  if (!ctx || !windex) return ZF_ERR_ARG;
  if (!ctx->have_wall) return fail(ctx, ZF_ERR_STATE, "no wall set up");
  ZF_HIP(ctx, hipMemcpy(windex, ctx->d_windex, 4 * ctx->box, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_cell_list(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_rep(ctx);
  return rc ? rc : cell_list_impl(ctx);
}

int zf_download_cell_list(zf_ctx *ctx, int32_t *bin) {
  if (!ctx || !bin) return ZF_ERR_ARG;
  int rc = require_rep(ctx);
  if (rc) return rc;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->ncell)
    ZF_HIP(ctx, hipMemcpy(bin, ctx->d_tbin, 4 * (size_t)ctx->ncell * ctx->nt, hipMemcpyDeviceToHost));
  return ZF_OK;
}

int zf_repulsion(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_rep(ctx);
  if (rc) return rc;
  // the kernel walks the bins of the last cell list
  if (ctx->bins_gen == ~0ull) return fail(ctx, ZF_ERR_STATE, "build the cell list first (zf_cell_list)");
  return repulsion_impl(ctx);
}

int zf_download_wall_neighbour(zf_ctx *ctx, int32_t *nearest) {
  if (!ctx || !nearest) return ZF_ERR_ARG;
  int rc = require_rep(ctx);
  if (rc) return rc;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->ncell)
    ZF_HIP(ctx, hipMemcpy(nearest, ctx->d_near, 4 * (size_t)ctx->ncell * ctx->nn, hipMemcpyDeviceToHost));
  return ZF_OK;
}

}  // extern "C"
