// zf_ibm.hip -- immersed-boundary coupling and the membrane position update.
//
// Replaces, for the lattice of this context, ZebraflowWT.c:
//   velocity interpolation   :8702-8996
//   force spreading          :10746-10867
//   position update          :11975-11989
// and drives one whole iteration in the reference's order (zf_step).
//
// Spreading without atomics: membrane nodes are sorted by the 4x4x4 lattice
// tile holding them (64-bit keys tile<<32 | node, radix sort, so the order is
// deterministic); one 64-lane wave per tile then gathers, for its 64 lattice
// nodes, the contributions of every membrane node in the 27 surrounding
// tiles, staged through LDS in chunks of 64.  The kernel support (radius 2
// lattice units) never reaches past the neighbouring tiles.
#include "zf_internal.h"

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

namespace zf {

int step(zf_ctx *ctx, int iter, int rbc_step, hipEvent_t *ev);
int membrane_forces_impl(zf_ctx *ctx);

namespace {

constexpr int kTileW = 4;
constexpr double kPi = 3.14159265358979;  // PI, ZebraflowWT.c:33

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// U,V,W of every listed node without the force shift (ZebraflowWT.c:6606-6614).
template <class T, bool DEV16, bool FULL, class U>
__global__ void k_macro_vel(const T *__restrict__ f, size_t stride, int n, U *__restrict__ u) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  constexpr int EX[19] = {1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0, 0};
  constexpr int EY[19] = {0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0};
  constexpr int EZ[19] = {0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0};
  double drho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int c = 0; c < 19; c++) {
    const double w = c < 6 ? 1.0 / 18 : c < 18 ? 1.0 / 36 : 1.0 / 3;
    double g = (double)f[(size_t)c * stride + s];
    if (FULL) g -= w;          // FP64 storage holds f
    if (DEV16) g *= w;         // FP16 storage holds (f - w)/w
    drho += g;
    jx += EX[c] * g;
    jy += EY[c] * g;
    jz += EZ[c] * g;
  }
  const double rho = 1.0 + drho;
  u[s] = (U)(jx / rho);
  u[stride + s] = (U)(jy / rho);
  u[2 * stride + s] = (U)(jz / rho);
}

struct IbmParams {
  double x0, y0, z0, dx, vel_scale, force_scale, gamma_p, dt_advance, vmax;
  int im, jm, km, ntx, nty, ntz;
  int cube;  // ZF_IBM_CUBE_SUPPORT: no spherical cut-off
};

// Outside the kernel support: the sphere |d| > 2 (reference), or the cube
// max |d_a| >= 2; d in lattice units.
template <class F>
__device__ __forceinline__ bool off_support(F dx, F dy, F dz, int cube) {
  if (cube) return fabs(dx) >= (F)2 || fabs(dy) >= (F)2 || fabs(dz) >= (F)2;
  return dx * dx + dy * dy + dz * dz > (F)4;
}

__device__ __forceinline__ double kern1(double d, double width) {
  return 1.0 + cos(kPi * fabs(d) / width);
}

// Interpolation, one lane per membrane node (ZebraflowWT.c:8950-8996):
// the lattice nodes within the cut-off (|d| <= 2 DX, or the cube |d_a| <
// 2 DX) of the node, in (i, j, k) order.  They lie within 2 DX along every
// axis, so among the 5 planes per axis from ceil(u - 2) on (u the node's
// 1-based lattice coordinate, taken 1e-9 lower to cover rounding), all
// inside the window the reference scans around its (int)((x-X0)/DX + 1.01)
// index.  The kernel factor of each plane is evaluated once, and a row's 5
// points are read in three batches (B, then the storage slots, then the
// velocities) instead of one dependent chain per point; the terms are the
// same expressions, summed in the same order.  With a wall layer (vslot
// non-null), window nodes that are not fluid add the same weight times
// their Uvoid (ZebraflowWT.c:8980-8996), in the window's order like the
// fluid terms.
template <class U>
__global__ __launch_bounds__(128) void k_interp(const double *__restrict__ X, size_t NT, int nn, int ncell,
                         const int32_t *__restrict__ status, const int32_t *__restrict__ b,
                         const int32_t *__restrict__ sbox, const U *__restrict__ umac,
                         size_t stride, IbmParams p, const int32_t *__restrict__ vslot,
                         const double *__restrict__ uvoid, double *__restrict__ V) {
  const int g = xcd_range_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  if (status[g / nn] <= kStatusDead) return;
  const double xp = X[g], yp = X[NT + g], zp = X[2 * NT + g];
  const int i = (int)((xp - p.x0) / p.dx + 1.01), j = (int)((yp - p.y0) / p.dx + 1.01),
            k = (int)((zp - p.z0) / p.dx + 1.01);
  const int i1 = max(1, i - 3), i2 = min(p.im, i + 3), j1 = max(1, j - 3), j2 = min(p.jm, j + 3),
            k1 = max(1, k - 3), k2 = min(p.km, k + 3);
  const int ilo = (int)ceil((xp - p.x0) / p.dx + 1.0 - 2.0 - 1e-9),
            jlo = (int)ceil((yp - p.y0) / p.dx + 1.0 - 2.0 - 1e-9),
            klo = (int)ceil((zp - p.z0) / p.dx + 1.0 - 2.0 - 1e-9);
  const double width = 2.0 * p.dx, w2 = width * width;
  double dzv[5], kzv[5];
#pragma unroll
  for (int t = 0; t < 5; t++) {
    dzv[t] = p.z0 + (klo + t - 1) * p.dx - zp;
    kzv[t] = kern1(dzv[t], width);
  }
  double v0 = 0, v1 = 0, v2 = 0;
#pragma unroll 1
  for (int a = 0; a < 5; a++) {
    const int ii = ilo + a;
    const double dxv = p.x0 + (ii - 1) * p.dx - xp;
    const double kxv = kern1(dxv, width);
    const bool iok = ii >= i1 && ii <= i2;
#pragma unroll 1
    for (int c = 0; c < 5; c++) {
      const int jj = jlo + c;
      const double dyc = p.y0 + (jj - 1) * p.dx - yp;
      bool ok[5];
      size_t q[5];
      int bv[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
        const int kk = klo + t;
        ok[t] = iok && jj >= j1 && jj <= j2 && kk >= k1 && kk <= k2 &&
                !(p.cube ? (fabs(dxv) >= width || fabs(dyc) >= width || fabs(dzv[t]) >= width)
                         : dxv * dxv + dyc * dyc + dzv[t] * dzv[t] > w2);
        q[t] = ok[t] ? ((size_t)(ii - 1) * p.jm + (jj - 1)) * p.km + (kk - 1) : 0;
        if (ok[t]) bv[t] = b[q[t]];
      }
      if (!__any(ok[0] || ok[1] || ok[2] || ok[3] || ok[4])) continue;
      const double kyc = kern1(dyc, width);
      bool fl[5];
      int sv[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
        fl[t] = ok[t] && bv[t] > 0 && bv[t] < 4998;
        sv[t] = -1;
        if (fl[t]) sv[t] = sbox[q[t]];
        else if (ok[t] && vslot) sv[t] = vslot[q[t]];
      }
      double u0[5], u1[5], u2[5];
#pragma unroll
      for (int t = 0; t < 5; t++) {
        if (fl[t]) {
          u0[t] = (double)umac[sv[t]];
          u1[t] = (double)umac[stride + sv[t]];
          u2[t] = (double)umac[2 * stride + sv[t]];
        } else if (ok[t] && sv[t] >= 0) {
          u0[t] = uvoid[3 * (size_t)sv[t]];
          u1[t] = uvoid[3 * (size_t)sv[t] + 1];
          u2[t] = uvoid[3 * (size_t)sv[t] + 2];
        }
      }
#pragma unroll
      for (int t = 0; t < 5; t++) {
        if (!ok[t] || sv[t] < 0) continue;
        const double w = kxv * kyc * kzv[t] / 64.0;
        v0 += w * p.vel_scale * u0[t];
        v1 += w * p.vel_scale * u1[t];
        v2 += w * p.vel_scale * u2[t];
      }
    }
  }
  V[g] = v0;
  V[NT + g] = v1;
  V[2 * NT + g] = v2;
}

__device__ __forceinline__ int tile_of(double pos, double origin, double dx, int nt) {
  int t = (int)floor((pos - origin) / dx / kTileW);
  return t < 0 ? 0 : t >= nt ? nt - 1 : t;
}

// Spreading bin of a position: its tile times 8 plus its octant in the
// tile, bit 2 (x), 1 (y), 0 (z) set in the upper half along that axis
// (tile-relative position >= 2 lattice units).  A node within 2 lattice
// units of a neighbouring tile, the reach of the kernel, lies in the half
// facing it.
__device__ __forceinline__ uint64_t spread_bin(double x, double y, double z, const IbmParams &p) {
  const int tx = tile_of(x, p.x0, p.dx, p.ntx), ty = tile_of(y, p.y0, p.dx, p.nty),
            tz = tile_of(z, p.z0, p.dx, p.ntz);
  const int hx = (x - p.x0) / p.dx - kTileW * tx >= 2.0, hy = (y - p.y0) / p.dx - kTileW * ty >= 2.0,
            hz = (z - p.z0) / p.dx - kTileW * tz >= 2.0;
  return ((((uint64_t)tx * p.nty + ty) * p.ntz + tz) << 3) | (uint64_t)(hx * 4 + hy * 2 + hz);
}

// A clone node within its band (band[2c] the plane, band[2c+1] the
// direction of travel) moves with its parent cell and does not spread
// (ZebraflowWT.c:11002-11015).
__device__ __forceinline__ bool clone_band_node(const double *band, int c, int st, double z) {
  if (!band || !(st < 0 && st > kStatusDead)) return false;
  const double plane = band[2 * c], dir = band[2 * c + 1];
  return dir > 0 ? !(z > plane) : dir < 0 ? !(z < plane) : false;
}

// v += v_rep; F_fluid = gamma_p v_rep + F_Total (ZebraflowWT.c:10749-10754);
// sort key = spreading bin << 32 | node (dead cells and clone band nodes
// sort past the last bin).  id_base: added to the node index in the key (the
// nucleus nodes follow the membrane's, from id_base = cap * nn).
template <class U>
__global__ void k_node_prep(const double *__restrict__ X, double *__restrict__ V,
                            const double *__restrict__ vrep, const double *__restrict__ F,
                            size_t NT, int nn, int ncell, const int32_t *__restrict__ status,
                            IbmParams p, U *__restrict__ ffl, uint64_t *__restrict__ keys,
                            uint32_t id_base, const double *__restrict__ band) {
  const int g = xcd_range_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  const uint64_t nbins = (uint64_t)p.ntx * p.nty * p.ntz * 8;
  const uint64_t id = (uint64_t)id_base + (uint64_t)g;
  const int st = status[g / nn];
  if (st <= kStatusDead || clone_band_node(band, g / nn, st, X[2 * NT + g])) {
    keys[g] = (nbins << 32) | id;
    return;
  }
  double fl[3];
  for (int a = 0; a < 3; a++) {
    const double vr = vrep[a * NT + g];
    V[a * NT + g] += vr;
    fl[a] = p.gamma_p * vr + F[a * NT + g];
  }
  ffl[g] = (U)fl[0];
  ffl[NT + g] = (U)fl[1];
  ffl[2 * NT + g] = (U)fl[2];
  keys[g] = (spread_bin(X[g], X[NT + g], X[2 * NT + g], p) << 32) | id;
}

__global__ void k_fill_i32(int32_t *a, size_t n, int32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// [begin, end) of every spreading bin in the sorted keys (cleared to empty
// beforehand), and occ[tile] = 1 for every tile holding a node.
__global__ void k_bin_ranges(const uint64_t *__restrict__ keys, int n, uint64_t nbins,
                             int32_t *__restrict__ range, int32_t *__restrict__ occ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t t = keys[i] >> 32;
  if (t >= nbins) return;
  if (i == 0 || (keys[i - 1] >> 32) != t) {
    range[2 * t] = i;
    occ[t >> 3] = 1;
  }
  if (i == n - 1 || (keys[i + 1] >> 32) != t) range[2 * t + 1] = i + 1;
}

// act[t] = 1 for the tiles with a node in one of their 27 neighbours
// (themselves included), 0 for the others, whose force is zero.
__global__ void k_spread_tile_flags(const int32_t *__restrict__ occ, int ntx, int nty, int ntz,
                                    int32_t *__restrict__ act) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntx * nty * ntz) return;
  const int tz = t % ntz, ty = (t / ntz) % nty, tx = t / (ntz * nty);
  bool any = false;
  for (int a = max(tx - 1, 0); a <= min(tx + 1, ntx - 1); a++)
    for (int c = max(ty - 1, 0); c <= min(ty + 1, nty - 1); c++)
      for (int e = max(tz - 1, 0); e <= min(tz + 1, ntz - 1); e++) any |= occ[(a * nty + c) * ntz + e] != 0;
  act[t] = any ? 1 : 0;
}

// The bins a tile gathers from, in gathering order: the 27 neighbouring
// tiles in (x, y, z) order, and in each the octants in the halves facing
// the tile (both halves along an axis where the neighbour is level with
// it), ascending.  8 + 6 x 4 + 12 x 2 + 8 x 1 = 64 entries, one per lane.
struct SpreadSub {
  signed char dx, dy, dz, oct;
};
struct SpreadSubs {
  SpreadSub e[64];
};
constexpr bool faces(int d, int h) { return d == 0 || (d < 0 ? h == 1 : h == 0); }
constexpr SpreadSubs make_spread_subs() {
  SpreadSubs t{};
  int n = 0;
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++)
        for (int o = 0; o < 8; o++)
          if (faces(dx, o >> 2) && faces(dy, (o >> 1) & 1) && faces(dz, o & 1))
            t.e[n++] = SpreadSub{(signed char)dx, (signed char)dy, (signed char)dz, (signed char)o};
  return t;
}
constexpr int count_spread_subs() {
  int n = 0;
  for (int o = 0; o < 27 * 8; o++)
    n += faces(o / 72 - 1, (o % 8) >> 2) && faces((o / 24) % 3 - 1, ((o % 8) >> 1) & 1) &&
         faces((o / 8) % 3 - 1, o & 1);
  return n;
}
static_assert(count_spread_subs() == 64, "one bin per lane");
__device__ constexpr SpreadSubs kSpreadSubs = make_spread_subs();

__device__ __forceinline__ float kcos_pi(float x) { return cospif(x); }
__device__ __forceinline__ double kcos_pi(double x) { return cospi(x); }

// One wave per listed tile (the list in tile order, a contiguous part per
// XCD); lane = lattice node (4x4x4).  Gathers the
// membrane nodes of the 64 bins of kSpreadSubs in (bin, node) order, in the
// field's type F (positions relative to the tile origin, weights and
// sums): the bins' runs of the sorted keys are taken as one list, staged 64
// nodes at a time.  The
// kernel weight is a product of one factor per axis, and a lane's offset
// from a membrane node along an axis takes one of 4 values (its row in the
// tile), so the lane staging a node evaluates its 12 factors once and the
// other lanes read theirs from LDS (same expressions, same values).
// Node ids from NT on are nucleus nodes (Xn, ffln, stride NTn).
// WRAP (a parent lattice's periodic window, zf_wrap.hip): bits 30-31 of a
// key's node id mark a periodic image of the node, seen L lower (1) or
// higher (2); the reference's kmap sends an image's share only to planes
// k >= 2 (1) or k <= KM - 1 (2), plane KM and plane 1 staying distinct
// (ZebraflowWT.c:11220-11222).
struct WrapSpread {
  double L;
  int km;
};

template <class F, bool WRAP = false>
__global__ __launch_bounds__(64) void k_spread_gather(
    const uint64_t *__restrict__ keys, const int32_t *__restrict__ range,
    const int32_t *__restrict__ tlist, const double *__restrict__ X, const F *__restrict__ ffl,
    size_t NT, const double *__restrict__ Xn, const F *__restrict__ ffln, size_t NTn,
    const int32_t *__restrict__ b, const int32_t *__restrict__ sbox, IbmParams p,
    F *__restrict__ force, size_t stride, WrapSpread ws = WrapSpread{0.0, 0}) {
  __shared__ F sx[64], sy[64], sz[64], sf0[64], sf1[64], sf2[64];
  __shared__ int scode[WRAP ? 64 : 1];
  __shared__ F kx[4][64], ky[4][64], kz[4][64];  // (1 + cos(pi |r - s| / 2)) for tile row r
  __shared__ int soff[64], sbeg[64];             // each bin's offset in the list, first key
  const int lane = threadIdx.x;
  const int ix = lane / 16, iy = (lane / 4) % 4, iz = lane % 4;
  const F rx = (F)ix, ry = (F)iy, rz = (F)iz;
  int tbeg, tend;
  xcd_list_span(tlist[0], tbeg, tend);
  for (int it = tbeg + (int)blockIdx.x / 8; it < tend; it += (int)gridDim.x / 8) {
    const int tile = tlist[1 + it];
    const int tz = tile % p.ntz, ty = (tile / p.ntz) % p.nty, tx = tile / (p.ntz * p.nty);
    // lattice node of this lane (0-based), and the tile origin
    const int li = tx * kTileW + ix, lj = ty * kTileW + iy, lk = tz * kTileW + iz;
    const double ox = p.x0 + (double)(tx * kTileW) * p.dx, oy = p.y0 + (double)(ty * kTileW) * p.dx,
                 oz = p.z0 + (double)(tz * kTileW) * p.dx;
    // this lane's bin, its run of keys, and the run's offset in the list
    const SpreadSub se = kSpreadSubs.e[lane];
    const int nx = tx + se.dx, ny = ty + se.dy, nz = tz + se.dz;
    int beg = 0, len = 0;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < p.ntx && ny < p.nty && nz < p.ntz) {
      const size_t bin = (((size_t)nx * p.nty + ny) * p.ntz + nz) * 8 + se.oct;
      beg = range[2 * bin];
      len = range[2 * bin + 1] - beg;
    }
    int incl = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(incl, d);
      incl += lane >= d ? v : 0;
    }
    const int total = __shfl(incl, 63);
    soff[lane] = incl - len;
    sbeg[lane] = beg;
    __syncthreads();
    F a0 = 0, a1 = 0, a2 = 0;
    for (int c0 = 0; c0 < total; c0 += 64) {
      // staged nodes are compacted in order to those within 2 of this
      // tile along every axis: any other one is off the support of all 64
      // lanes, so the sums are unchanged
      bool rel = false;
      F px = 0, py = 0, pz = 0;
      size_t g = 0;
      const double *Xs = X;
      const F *fs = ffl;
      size_t st = NT;
      int code = 0;
      const int v = c0 + lane;
      if (v < total) {
        // the bin holding list entry v: the last with offset <= v (an
        // empty bin shares its offset with the next one)
        int lo = 0, hi = 64;
#pragma unroll
        for (int step = 0; step < 6; step++) {
          const int mid = (lo + hi) >> 1;
          if (soff[mid] <= v) lo = mid;
          else hi = mid;
        }
        g = (size_t)(keys[sbeg[lo] + (v - soff[lo])] & 0xffffffffu);
        if (WRAP) {
          code = (int)(g >> 30);
          g &= 0x3fffffffu;
        }
        if (g >= NT) {
          g -= NT;
          Xs = Xn;
          fs = ffln;
          st = NTn;
        }
        double zz = Xs[2 * st + g];
        if (WRAP) zz += code == 1 ? -ws.L : code == 2 ? ws.L : 0.0;
        px = (F)((Xs[g] - ox) / p.dx);
        py = (F)((Xs[st + g] - oy) / p.dx);
        pz = (F)((zz - oz) / p.dx);
        rel = px >= (F)-2 && px <= (F)5 && py >= (F)-2 && py <= (F)5 && pz >= (F)-2 && pz <= (F)5;
      }
      const unsigned long long relm = __ballot(rel);
      const int cnt = __popcll(relm);
      if (rel) {
        const int sl = __popcll(relm & ((1ull << lane) - 1ull));
        sx[sl] = px;
        sy[sl] = py;
        sz[sl] = pz;
        sf0[sl] = fs[g];
        sf1[sl] = fs[st + g];
        sf2[sl] = fs[2 * st + g];
        if (WRAP) scode[sl] = code;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          // (1 + cos(pi |d| / 2)) per axis, d in lattice units (width 2 DX)
          kx[r][sl] = (F)1 + kcos_pi((F)0.5 * fabs((F)r - px));
          ky[r][sl] = (F)1 + kcos_pi((F)0.5 * fabs((F)r - py));
          kz[r][sl] = (F)1 + kcos_pi((F)0.5 * fabs((F)r - pz));
        }
      }
      __syncthreads();
      for (int m = 0; m < cnt; m++) {
        const F dx = rx - sx[m], dy = ry - sy[m], dz = rz - sz[m];
        if (off_support(dx, dy, dz, p.cube)) continue;
        if (WRAP && ((scode[m] == 1 && lk == 0) || (scode[m] == 2 && lk == ws.km - 1))) continue;
        const F w = kx[ix][m] * ky[iy][m] * kz[iz][m] * ((F)1 / (F)64);
        a0 += w * sf0[m];
        a1 += w * sf1[m];
        a2 += w * sf2[m];
      }
      __syncthreads();
    }
    if (total > 0 && li < p.im && lj < p.jm && lk < p.km) {
      const size_t q = ((size_t)li * p.jm + lj) * p.km + lk;
      if (b[q] >= 0) {
        const int s = sbox[q];
        const F fsc = (F)p.force_scale;
        force[s] = a0 * fsc;
        force[stride + s] = a1 * fsc;
        force[2 * stride + s] = a2 * fsc;
      }
    }
  }
}

// ---- a parent lattice's cells: the periodic window along k (zf_wrap.hip)
struct WrapIbm {
  int dir, km;
  double z0, zm, L, mid;
};

// kmap of the interpolation and spreading windows (ZebraflowWT.c:7574-7576).
__device__ __forceinline__ int kmap_ibm(int k, int km) {
  return k < 1 ? k + (km - 1) : k > km ? k - (km - 1) : k;
}

// Interpolation of a parent's cells (:7541-7732 for status 2, :8101-8297
// for status > 2), one lane per node: a crossing cell's window is centred
// at zpmap (its nodes behind the window's middle seen past the exit);
// window planes wrap through kmap at their unwrapped place.  The cut-off
// of a fluid node is taken at zpmap and its z factor at the node's own z
// (:8145-8147), those of the wall layer both at its own z (:8153-8155),
// so a node seen through zpmap takes no Uvoid term.  Fluid is B > 0.
template <class U>
__global__ void k_wrap_interp(const double *__restrict__ X, size_t NT, int nn, int ncell,
                              const int32_t *__restrict__ status, const int32_t *__restrict__ b,
                              const int32_t *__restrict__ sbox, const U *__restrict__ umac,
                              size_t stride, IbmParams p, WrapIbm w, const int32_t *__restrict__ vslot,
                              const double *__restrict__ uvoid, double *__restrict__ V) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  const int st = status[g / nn];
  if (st < 2) return;
  const double xp = X[g], yp = X[NT + g], zp = X[2 * NT + g];
  double ze = zp;
  if (st > 2) {
    if (w.dir > 0 && zp < w.mid) ze = zp + w.L;
    if (w.dir < 0 && zp > w.mid) ze = zp - w.L;
  }
  const int i = (int)((xp - p.x0) / p.dx + 1.01), j = (int)((yp - p.y0) / p.dx + 1.01),
            k = (int)((ze - p.z0) / p.dx + 1.01);
  const int i1 = max(1, i - 3), i2 = min(p.im, i + 3), j1 = max(1, j - 3), j2 = min(p.jm, j + 3);
  const double width = 2.0 * p.dx, w2 = width * width;
  double v0 = 0, v1 = 0, v2 = 0;
  for (int ii = i1; ii <= i2; ii++) {
    const double dxv = p.x0 + (ii - 1) * p.dx - xp;
    for (int jj = j1; jj <= j2; jj++) {
      const double dyv = p.y0 + (jj - 1) * p.dx - yp;
      for (int kk = k - 3; kk <= k + 3; kk++) {
        const double zc = p.z0 + (kk - 1) * p.dx;
        const size_t q = ((size_t)(ii - 1) * p.jm + (jj - 1)) * p.km + (kmap_ibm(kk, p.km) - 1);
        const double dzv = zc - zp;
        if (b[q] > 0) {
          const double dze = zc - ze;
          if (dxv * dxv + dyv * dyv + dze * dze > w2) continue;
          const double wt = kern1(dxv, width) * kern1(dyv, width) * kern1(dzv, width) / 64.0;
          const int s = sbox[q];
          v0 += wt * p.vel_scale * (double)umac[s];
          v1 += wt * p.vel_scale * (double)umac[stride + s];
          v2 += wt * p.vel_scale * (double)umac[2 * stride + s];
        } else if (vslot) {
          const int sl = vslot[q];
          if (sl < 0 || dxv * dxv + dyv * dyv + dzv * dzv > w2) continue;
          const double wt = kern1(dxv, width) * kern1(dyv, width) * kern1(dzv, width) / 64.0;
          v0 += wt * p.vel_scale * uvoid[3 * (size_t)sl];
          v1 += wt * p.vel_scale * uvoid[3 * (size_t)sl + 1];
          v2 += wt * p.vel_scale * uvoid[3 * (size_t)sl + 2];
        }
      }
    }
  }
  V[g] = v0;
  V[NT + g] = v1;
  V[2 * NT + g] = v2;
}

// Spreading keys of a parent's cells: each live node at its own (wrapped)
// z, plus one periodic image when it lies within 3 DX of an end of the
// window (keys_img; otherwise a key past the last tile).
template <class U>
__global__ void k_wrap_node_prep(const double *__restrict__ X, double *__restrict__ V,
                                 const double *__restrict__ vrep, const double *__restrict__ F,
                                 size_t NT, int nn, int ncell, const int32_t *__restrict__ status,
                                 IbmParams p, WrapIbm w, U *__restrict__ ffl,
                                 uint64_t *__restrict__ keys, uint64_t *__restrict__ keys_img,
                                 uint32_t id_base) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  const uint64_t nbins = (uint64_t)p.ntx * p.nty * p.ntz * 8;
  const uint64_t id = (uint64_t)id_base + (uint64_t)g;
  keys_img[g] = (nbins << 32) | id;
  if (status[g / nn] < 2) {
    keys[g] = (nbins << 32) | id;
    return;
  }
  double fl[3];
  for (int a = 0; a < 3; a++) {
    const double vr = vrep[a * NT + g];
    V[a * NT + g] += vr;
    fl[a] = p.gamma_p * vr + F[a * NT + g];
  }
  ffl[g] = (U)fl[0];
  ffl[NT + g] = (U)fl[1];
  ffl[2 * NT + g] = (U)fl[2];
  const double x = X[g], y = X[NT + g], z = X[2 * NT + g];
  keys[g] = (spread_bin(x, y, z, p) << 32) | id;
  int code = 0;
  double zi = z;
  if (z > w.zm - 3.0 * p.dx) { code = 1; zi = z - w.L; }
  else if (z < w.z0 + 3.0 * p.dx) { code = 2; zi = z + w.L; }
  if (code) keys_img[g] = (spread_bin(x, y, zi, p) << 32) | ((uint64_t)code << 30) | id;
}

__global__ void k_speed(const double *__restrict__ V, size_t NT, int nn, int ncell,
                        const int32_t *__restrict__ status, double vmax, int32_t *__restrict__ flag) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  if (status[g / nn] < 0) return;
  const double v0 = V[g], v1 = V[NT + g], v2 = V[2 * NT + g];
  const double s = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
  if (s > vmax || isnan(s) || isinf(s)) atomicOr(&flag[g / nn], 1);
}

__global__ void k_move(double *__restrict__ X, const double *__restrict__ V, size_t NT, int nn,
                       int ncell, int32_t *__restrict__ status, const int32_t *__restrict__ flag,
                       double dt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncell * nn) return;
  const int c = g / nn;
  if (status[c] < 0 || flag[c]) return;
  for (int a = 0; a < 3; a++) X[a * NT + g] += V[a * NT + g] * dt;
}

__global__ void k_remove(int32_t *__restrict__ status, const int32_t *__restrict__ flag, int ncell,
                         int32_t *__restrict__ count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  if (status[c] >= 0 && flag[c]) {
    status[c] = -5000;
    atomicAdd(count, 1);
  }
}

__global__ void k_import_soa(const double *__restrict__ h, int n, size_t NT, double *__restrict__ A) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  for (int a = 0; a < 3; a++) A[a * NT + g] = h[3 * (size_t)g + a];
}

__global__ void k_export_soa(const double *__restrict__ A, int n, size_t NT, double *__restrict__ h) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  for (int a = 0; a < 3; a++) h[3 * (size_t)g + a] = A[a * NT + g];
}

__global__ void k_zero_f64(double *a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

IbmParams params(const zf_ctx *ctx) {
  const zf_ibm_config &c = ctx->icfg;
  return IbmParams{c.x0, c.y0, c.z0, c.dx, c.vel_scale, c.force_scale, c.gamma_p, c.dt_advance,
                   c.vmax, ctx->cfg.im, ctx->cfg.jm, ctx->cfg.km, ctx->ntx, ctx->nty, ctx->ntz,
                   (c.flags & ZF_IBM_CUBE_SUPPORT) ? 1 : 0};
}

int require_ibm(zf_ctx *ctx) {
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "zf_ibm_setup not called");
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  if (!ctx->have_lattice) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  return ZF_OK;
}

}  // namespace

int ensure_force_field(zf_ctx *ctx) {
  if (ctx->d_force) return ZF_OK;
  const size_t nw = 3 * ctx->stride * field_elem(ctx) / 4;
  int rc = dev_alloc(ctx, (void **)&ctx->d_force, 4 * nw);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb(nw, 256)), dim3(256), 0, ctx->stream,
                     static_cast<int32_t *>(ctx->d_force), nw, 0);
  return ZF_OK;
}

void free_ibm(zf_ctx *ctx) {
  free_wrap(ctx);
  void **bufs[] = {(void **)&ctx->d_umac, (void **)&ctx->d_vrep,
                   (void **)&ctx->d_ffl,   (void **)&ctx->d_keys[0], (void **)&ctx->d_keys[1],
                   (void **)&ctx->d_sort_tmp, (void **)&ctx->d_tile_range, (void **)&ctx->d_flag,
                   (void **)&ctx->d_sbin_range, (void **)&ctx->d_tile_occ, (void **)&ctx->d_spread_tiles,
                   (void **)&ctx->d_tile_act, &ctx->d_sel_tmp};
  for (auto b : bufs) dev_free(ctx, b);
  ctx->keys_cap = 0;
  ctx->have_ibm = false;
}

int ensure_sort_keys(zf_ctx *ctx, size_t n) {
  if (n <= ctx->keys_cap) return ZF_OK;
  if (n >= (1ull << 32)) return fail(ctx, ZF_ERR_ARG, "too many immersed-boundary points");
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dev_free(ctx, (void **)&ctx->d_keys[0]);
  dev_free(ctx, (void **)&ctx->d_keys[1]);
  dev_free(ctx, &ctx->d_sort_tmp);
  ctx->keys_cap = 0;
  int rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_keys[0], 8 * n)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_keys[1], 8 * n)))
    return rc;
  size_t tmp = 0;
  ZF_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, ctx->d_keys[0], ctx->d_keys[1],
                                                  (int)n, 0, 64, ctx->stream));
  ctx->sort_tmp_bytes = tmp;
  if ((rc = dev_alloc(ctx, &ctx->d_sort_tmp, tmp))) return rc;
  ctx->keys_cap = n;
  return ZF_OK;
}

int ibm_interpolate_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  // macro velocities of the current populations
  const void *f = ctx->d_f[ctx->cur];
  const unsigned g = nb(ctx->dommax, 256);
  float *um = static_cast<float *>(ctx->d_umac);
  switch (ctx->cfg.storage) {
    case ZF_STORE_F64:
      hipLaunchKernelGGL((k_macro_vel<double, false, true, double>), dim3(g), dim3(256), 0,
                         ctx->stream, (const double *)f, ctx->stride, ctx->dommax,
                         static_cast<double *>(ctx->d_umac));
      hipLaunchKernelGGL(k_interp<double>, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x, NT,
                         ctx->nn, ctx->ncell, ctx->d_status, ctx->d_b, ctx->d_sbox,
                         static_cast<const double *>(ctx->d_umac), ctx->stride, params(ctx),
                         ctx->have_void ? ctx->d_vslot : nullptr, ctx->d_uvoid, ctx->d_v);
      if (ctx->have_nuc && ctx->nnuc > 0)
        hipLaunchKernelGGL(k_interp<double>, dim3(nb(ctx->ncell * ctx->nnuc, 128)), dim3(128), 0,
                           ctx->stream, ctx->d_xn, ncap * ctx->nnuc, ctx->nnuc, ctx->ncell,
                           ctx->d_status, ctx->d_b, ctx->d_sbox,
                           static_cast<const double *>(ctx->d_umac), ctx->stride, params(ctx),
                           ctx->have_void ? ctx->d_vslot : nullptr, ctx->d_uvoid, ctx->d_vn);
      ZF_HIP(ctx, hipGetLastError());
      return ZF_OK;
    case ZF_STORE_F32DEV:
      hipLaunchKernelGGL((k_macro_vel<float, false, false, float>), dim3(g), dim3(256), 0, ctx->stream,
                         (const float *)f, ctx->stride, ctx->dommax, um);
      break;
    default:
      hipLaunchKernelGGL((k_macro_vel<_Float16, true, false, float>), dim3(g), dim3(256), 0,
                         ctx->stream, (const _Float16 *)f, ctx->stride, ctx->dommax, um);
  }
  hipLaunchKernelGGL(k_interp<float>, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x, NT,
                     ctx->nn, ctx->ncell, ctx->d_status, ctx->d_b, ctx->d_sbox, um, ctx->stride,
                     params(ctx), ctx->have_void ? ctx->d_vslot : nullptr, ctx->d_uvoid, ctx->d_v);
  if (ctx->have_nuc && ctx->nnuc > 0)
    hipLaunchKernelGGL(k_interp<float>, dim3(nb(ctx->ncell * ctx->nnuc, 128)), dim3(128), 0,
                       ctx->stream, ctx->d_xn, ncap * ctx->nnuc, ctx->nnuc, ctx->ncell, ctx->d_status,
                       ctx->d_b, ctx->d_sbox, um, ctx->stride, params(ctx),
                       ctx->have_void ? ctx->d_vslot : nullptr, ctx->d_uvoid, ctx->d_vn);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

// Sorts the nkeys spreading keys (bin << 32 | node) and builds the bins'
// runs, the occupied tiles and the list of tiles to gather.  in_order: the
// keys come in node order, so a (stable) radix sort of the bin bits alone
// gives the (bin, node) order.
int spread_bins(zf_ctx *ctx, int nkeys, uint64_t ntiles, bool in_order) {
  const uint64_t nbins = 8 * ntiles;
  // end bit: bins need ceil(log2(nbins+1)) bits above the 32 node bits
  int bbits = 1;
  while ((1ull << bbits) <= nbins) bbits++;
  size_t tmp = ctx->sort_tmp_bytes;
  ZF_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->d_sort_tmp, tmp, ctx->d_keys[0], ctx->d_keys[1],
                                                  nkeys, in_order ? 32 : 0, 32 + bbits, ctx->stream));
  hipLaunchKernelGGL(k_fill_i32, dim3(nb(2 * nbins, 256)), dim3(256), 0, ctx->stream, ctx->d_sbin_range,
                     (size_t)(2 * nbins), 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb(ntiles, 256)), dim3(256), 0, ctx->stream, ctx->d_tile_occ,
                     (size_t)ntiles, 0);
  if (nkeys > 0)
    hipLaunchKernelGGL(k_bin_ranges, dim3(nb(nkeys, 256)), dim3(256), 0, ctx->stream, ctx->d_keys[1],
                       nkeys, nbins, ctx->d_sbin_range, ctx->d_tile_occ);
  hipLaunchKernelGGL(k_spread_tile_flags, dim3(nb(ntiles, 256)), dim3(256), 0, ctx->stream,
                     ctx->d_tile_occ, ctx->ntx, ctx->nty, ctx->ntz, ctx->d_tile_act);
  // the flagged tiles in tile order: count at [0], tiles from [1] on
  size_t stmp = ctx->sel_tmp_bytes;
  ZF_HIP(ctx, hipcub::DeviceSelect::Flagged(ctx->d_sel_tmp, stmp, hipcub::CountingInputIterator<int32_t>(0),
                                            ctx->d_tile_act, ctx->d_spread_tiles + 1, ctx->d_spread_tiles,
                                            (int)ntiles, ctx->stream));
  return ZF_OK;
}

// Workgroups of the gather (one wave each, over the listed tiles), a
// multiple of 8.
unsigned spread_grid(uint64_t ntiles) { return (unsigned)std::min<uint64_t>((ntiles + 7) / 8 * 8, 16384); }

int ibm_spread_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  int rc = ensure_force_field(ctx);
  if (rc) return rc;
  const IbmParams p = params(ctx);
  const uint64_t ntiles = (uint64_t)ctx->ntx * ctx->nty * ctx->ntz;
  // the force field is rebuilt from zero (cleared by a fill kernel)
  const size_t nw = 3 * ctx->stride * field_elem(ctx) / 4;
  const bool f64 = field_elem(ctx) == 8;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb(nw, 256)), dim3(256), 0, ctx->stream,
                     static_cast<int32_t *>(ctx->d_force), nw, 0);
  if (n == 0) return ZF_OK;
  if (f64)
    hipLaunchKernelGGL(k_node_prep<double>, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_x,
                       ctx->d_v, ctx->d_vrep, ctx->d_mforce, NT, ctx->nn, ctx->ncell, ctx->d_status, p,
                       static_cast<double *>(ctx->d_ffl), ctx->d_keys[0], 0u, ctx->d_band);
  else
    hipLaunchKernelGGL(k_node_prep<float>, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_x,
                       ctx->d_v, ctx->d_vrep, ctx->d_mforce, NT, ctx->nn, ctx->ncell, ctx->d_status, p,
                       static_cast<float *>(ctx->d_ffl), ctx->d_keys[0], 0u, ctx->d_band);
  // nucleus nodes (F_fluid = gamma_p NODE_WALLREPUL_N + F_Total_N,
  // ZebraflowWT.c:10879-10884), keyed after the membrane's
  const size_t NTn = ctx->have_nuc ? (size_t)ctx->mcfg.max_cells * ctx->nnuc : 0;
  const int nnuc_pts = ctx->have_nuc ? ctx->ncell * ctx->nnuc : 0;
  if (nnuc_pts > 0) {
    if (f64)
      hipLaunchKernelGGL(k_node_prep<double>, dim3(nb(nnuc_pts, 256)), dim3(256), 0, ctx->stream,
                         ctx->d_xn, ctx->d_vn, ctx->d_vrepn, ctx->d_fn, NTn, ctx->nnuc, ctx->ncell,
                         ctx->d_status, p, static_cast<double *>(ctx->d_ffln), ctx->d_keys[0] + n,
                         (uint32_t)NT, ctx->d_band);
    else
      hipLaunchKernelGGL(k_node_prep<float>, dim3(nb(nnuc_pts, 256)), dim3(256), 0, ctx->stream,
                         ctx->d_xn, ctx->d_vn, ctx->d_vrepn, ctx->d_fn, NTn, ctx->nnuc, ctx->ncell,
                         ctx->d_status, p, static_cast<float *>(ctx->d_ffln), ctx->d_keys[0] + n,
                         (uint32_t)NT, ctx->d_band);
  }
  const int nkeys = n + nnuc_pts;
  // membrane keys by node, then the nucleus keys by node from NT on
  if ((rc = spread_bins(ctx, nkeys, ntiles, true))) return rc;
  const unsigned grid = spread_grid(ntiles);
  if (f64)
    hipLaunchKernelGGL(k_spread_gather<double>, dim3(grid), dim3(64), 0, ctx->stream, ctx->d_keys[1],
                       ctx->d_sbin_range, ctx->d_spread_tiles, ctx->d_x,
                       static_cast<const double *>(ctx->d_ffl), NT, ctx->d_xn,
                       static_cast<const double *>(ctx->d_ffln), NTn, ctx->d_b, ctx->d_sbox, p,
                       static_cast<double *>(ctx->d_force), ctx->stride);
  else
    hipLaunchKernelGGL(k_spread_gather<float>, dim3(grid), dim3(64), 0, ctx->stream, ctx->d_keys[1],
                       ctx->d_sbin_range, ctx->d_spread_tiles, ctx->d_x,
                       static_cast<const float *>(ctx->d_ffl), NT, ctx->d_xn,
                       static_cast<const float *>(ctx->d_ffln), NTn, ctx->d_b, ctx->d_sbox, p,
                       static_cast<float *>(ctx->d_force), ctx->stride);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

WrapIbm wrap_ibm(const zf_ctx *ctx) {
  return WrapIbm{ctx->wcfg.direction, ctx->cfg.km, ctx->wz0, ctx->wzm, ctx->wzm - ctx->wz0,
                 0.5 * (ctx->wz0 + ctx->wzm)};
}

int wrap_interpolate_impl(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  const void *f = ctx->d_f[ctx->cur];
  const unsigned g = nb(ctx->dommax, 256);
  const WrapIbm w = wrap_ibm(ctx);
  const int32_t *vs = ctx->have_void ? ctx->d_vslot : nullptr;
  const int nnp = ctx->have_nuc ? ctx->ncell * ctx->nnuc : 0;
  if (ctx->cfg.storage == ZF_STORE_F64) {
    double *um = static_cast<double *>(ctx->d_umac);
    hipLaunchKernelGGL((k_macro_vel<double, false, true, double>), dim3(g), dim3(256), 0, ctx->stream,
                       (const double *)f, ctx->stride, ctx->dommax, um);
    hipLaunchKernelGGL(k_wrap_interp<double>, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x,
                       NT, ctx->nn, ctx->ncell, ctx->d_status, ctx->d_b, ctx->d_sbox, um, ctx->stride,
                       params(ctx), w, vs, ctx->d_uvoid, ctx->d_v);
    if (nnp > 0)
      hipLaunchKernelGGL(k_wrap_interp<double>, dim3(nb(nnp, 128)), dim3(128), 0, ctx->stream,
                         ctx->d_xn, ncap * ctx->nnuc, ctx->nnuc, ctx->ncell, ctx->d_status, ctx->d_b,
                         ctx->d_sbox, um, ctx->stride, params(ctx), w, vs, ctx->d_uvoid, ctx->d_vn);
  } else {
    float *um = static_cast<float *>(ctx->d_umac);
    if (ctx->cfg.storage == ZF_STORE_F32DEV)
      hipLaunchKernelGGL((k_macro_vel<float, false, false, float>), dim3(g), dim3(256), 0, ctx->stream,
                         (const float *)f, ctx->stride, ctx->dommax, um);
    else
      hipLaunchKernelGGL((k_macro_vel<_Float16, true, false, float>), dim3(g), dim3(256), 0,
                         ctx->stream, (const _Float16 *)f, ctx->stride, ctx->dommax, um);
    hipLaunchKernelGGL(k_wrap_interp<float>, dim3(nb(n, 128)), dim3(128), 0, ctx->stream, ctx->d_x,
                       NT, ctx->nn, ctx->ncell, ctx->d_status, ctx->d_b, ctx->d_sbox, um, ctx->stride,
                       params(ctx), w, vs, ctx->d_uvoid, ctx->d_v);
    if (nnp > 0)
      hipLaunchKernelGGL(k_wrap_interp<float>, dim3(nb(nnp, 128)), dim3(128), 0, ctx->stream,
                         ctx->d_xn, ncap * ctx->nnuc, ctx->nnuc, ctx->ncell, ctx->d_status, ctx->d_b,
                         ctx->d_sbox, um, ctx->stride, params(ctx), w, vs, ctx->d_uvoid, ctx->d_vn);
  }
  // the repulsion velocities of this iteration start from zero: the wall
  // term of the same loop (zf_cells.hip), then contact and cell-cell
  hipLaunchKernelGGL(k_zero_f64, dim3(nb(3 * NT, 256)), dim3(256), 0, ctx->stream, ctx->d_vrep, 3 * NT);
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

template <class F>
static int wrap_spread_typed(zf_ctx *ctx) {
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  const IbmParams p = params(ctx);
  const WrapIbm w = wrap_ibm(ctx);
  const uint64_t ntiles = (uint64_t)ctx->ntx * ctx->nty * ctx->ntz;
  const size_t NTn = ctx->have_nuc ? ncap * ctx->nnuc : 0;
  const int nnp = ctx->have_nuc ? ctx->ncell * ctx->nnuc : 0;
  const int half = n + nnp;  // originals, then images
  uint64_t *k0 = ctx->d_keys[0];
  hipLaunchKernelGGL(k_wrap_node_prep<F>, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_x,
                     ctx->d_v, ctx->d_vrep, ctx->d_mforce, NT, ctx->nn, ctx->ncell, ctx->d_status, p, w,
                     static_cast<F *>(ctx->d_ffl), k0, k0 + half, 0u);
  if (nnp > 0)
    hipLaunchKernelGGL(k_wrap_node_prep<F>, dim3(nb(nnp, 256)), dim3(256), 0, ctx->stream, ctx->d_xn,
                       ctx->d_vn, ctx->d_vrepn, ctx->d_fn, NTn, ctx->nnuc, ctx->ncell, ctx->d_status, p,
                       w, static_cast<F *>(ctx->d_ffln), k0 + n, k0 + half + n, (uint32_t)NT);
  const int nkeys = 2 * half;
  int rc = spread_bins(ctx, nkeys, ntiles, false);  // images after the originals: full sort
  if (rc) return rc;
  hipLaunchKernelGGL((k_spread_gather<F, true>), dim3(spread_grid(ntiles)), dim3(64), 0, ctx->stream,
                     ctx->d_keys[1], ctx->d_sbin_range, ctx->d_spread_tiles, ctx->d_x,
                     static_cast<const F *>(ctx->d_ffl), NT, ctx->d_xn,
                     static_cast<const F *>(ctx->d_ffln), NTn, ctx->d_b, ctx->d_sbox, p,
                     static_cast<F *>(ctx->d_force), ctx->stride, WrapSpread{w.L, ctx->cfg.km});
  ZF_HIP(ctx, hipGetLastError());
  return ZF_OK;
}

int wrap_spread_impl(zf_ctx *ctx) {
  int rc = ensure_force_field(ctx);
  if (rc) return rc;
  const size_t nw = 3 * ctx->stride * field_elem(ctx) / 4;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb(nw, 256)), dim3(256), 0, ctx->stream,
                     static_cast<int32_t *>(ctx->d_force), nw, 0);
  if (ctx->ncell * ctx->nn == 0) return ZF_OK;
  return field_elem(ctx) == 8 ? wrap_spread_typed<double>(ctx) : wrap_spread_typed<float>(ctx);
}

int advance_impl(zf_ctx *ctx, int32_t *n_removed) {
  ctx->geom_gen++;  // positions change: repulsion bounds of the cell list expire
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const int n = ctx->ncell * ctx->nn, nc = ctx->ncell;
  if (n == 0) {
    if (n_removed) *n_removed = 0;
    return ZF_OK;
  }
  // flag[nc] is the removal counter
  hipLaunchKernelGGL(k_fill_i32, dim3(nb(nc + 1, 256)), dim3(256), 0, ctx->stream, ctx->d_flag,
                     (size_t)nc + 1, 0);
  hipLaunchKernelGGL(k_speed, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_v, NT, ctx->nn,
                     nc, ctx->d_status, ctx->icfg.vmax, ctx->d_flag);
  hipLaunchKernelGGL(k_move, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_x, ctx->d_v, NT,
                     ctx->nn, nc, ctx->d_status, ctx->d_flag, ctx->icfg.dt_advance);
  // the nucleus nodes move with their cell's update (:11950-11956, no
  // speed check of their own)
  if (ctx->have_nuc && ctx->nnuc > 0)
    hipLaunchKernelGGL(k_move, dim3(nb(nc * ctx->nnuc, 256)), dim3(256), 0, ctx->stream, ctx->d_xn,
                       ctx->d_vn, ncap * ctx->nnuc, ctx->nnuc, nc, ctx->d_status, ctx->d_flag,
                       ctx->icfg.dt_advance);
  hipLaunchKernelGGL(k_remove, dim3(nb(nc, 256)), dim3(256), 0, ctx->stream, ctx->d_status,
                     ctx->d_flag, nc, ctx->d_flag + nc);
  ZF_HIP(ctx, hipGetLastError());
  if (n_removed) {
    ZF_HIP(ctx, hipMemcpyAsync(n_removed, ctx->d_flag + nc, 4, hipMemcpyDeviceToHost, ctx->stream));
    ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return ZF_OK;
}

int rbc_update(zf_ctx *ctx, bool advance, int iter) {
  int rc;
  if ((rc = membrane_forces_impl(ctx))) return rc;
  if (ctx->have_nuc && (rc = nucleus_forces_impl(ctx))) return rc;
  if (ctx->have_rep && (rc = cell_list_impl(ctx))) return rc;
  if ((rc = ibm_interpolate_impl(ctx))) return rc;
  if (ctx->have_rep) {
    if ((rc = repulsion_impl(ctx))) return rc;
  } else {
    const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
    hipLaunchKernelGGL(k_zero_f64, dim3(nb(3 * NT, 256)), dim3(256), 0, ctx->stream, ctx->d_vrep, 3 * NT);
  }
  if (ctx->have_nuc && (rc = nucleus_contact_impl(ctx, iter))) return rc;
  if ((rc = ibm_spread_impl(ctx))) return rc;
  return advance ? advance_impl(ctx, nullptr) : ZF_OK;
}

}  // namespace zf

using namespace zf;

extern "C" {

// Reference scales: X0=-1, Y0=29, Z0=-27 (ZebraflowWT.c:95-97), DX=0.5,
// VEL_SCALE = L_SCALE/T_SCALE = 1, UNITMASS = DX^3 * 1e4 kg/m^3,
// LBM_BFORCE_SCALE = L_SCALE/T_SCALE^2/3, gamma_p = 1e-7 (:124),
// DT*SOLVERSTAGGER = 25 (:41, :151), removal speed 5e-3 (:11978).
void zf_ibm_config_reference(zf_ibm_config *ic) {
  *ic = zf_ibm_config{};
  const double l = 0.5e-6, t = 0.5e-6;
  ic->x0 = -1.0;
  ic->y0 = 29.0;
  ic->z0 = -27.0;
  ic->dx = 0.5;
  ic->vel_scale = l / t;
  const double unitmass = l * l * l * 10.0 * 1000.0;
  const double bforce = 1.0 / 3.0 * l / t / t;
  ic->force_scale = 1.0 / unitmass / bforce;
  ic->gamma_p = 1e-7;
  ic->dt_advance = 0.5 * 50;
  ic->vmax = 5e-3;
  ic->rbc_interval = 50;
}

int zf_ibm_setup(zf_ctx *ctx, const zf_ibm_config *ic) {
  if (!ctx || !ic || ic->dx <= 0 || ic->rbc_interval < 1) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "call zf_membrane_setup first");
  if (!ctx->have_lattice) return fail(ctx, ZF_ERR_STATE, "call zf_upload_lattice first");
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  free_cells(ctx);
  free_ibm(ctx);
  ctx->icfg = *ic;
  ctx->ntx = (ctx->cfg.im + kTileW - 1) / kTileW;
  ctx->nty = (ctx->cfg.jm + kTileW - 1) / kTileW;
  ctx->ntz = (ctx->cfg.km + kTileW - 1) / kTileW;
  const size_t ncap = ctx->mcfg.max_cells, NT = ncap * ctx->nn;
  const size_t ntiles = (size_t)ctx->ntx * ctx->nty * ctx->ntz;
  if (NT >= (1ull << 32) || ntiles >= (1ull << 31)) return fail(ctx, ZF_ERR_ARG, "too many nodes or tiles");
  int rc;
  if ((rc = dev_alloc(ctx, (void **)&ctx->d_umac, 3 * field_elem(ctx) * ctx->stride)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_vrep, 24 * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_ffl, 3 * field_elem(ctx) * NT)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_tile_range, 8 * ntiles)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_sbin_range, 64 * ntiles)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_tile_occ, 4 * ntiles)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_spread_tiles, 4 * (ntiles + 1))) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_tile_act, 4 * ntiles)) ||
      (rc = dev_alloc(ctx, (void **)&ctx->d_flag, 4 * (ncap + 1))))
    return rc;
  size_t stmp = 0;
  ZF_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, stmp, hipcub::CountingInputIterator<int32_t>(0),
                                            ctx->d_tile_act, ctx->d_spread_tiles + 1, ctx->d_spread_tiles,
                                            (int)ntiles, ctx->stream));
  ctx->sel_tmp_bytes = stmp;
  if ((rc = dev_alloc(ctx, &ctx->d_sel_tmp, stmp))) return rc;
  // sort keys for the membrane nodes and, with a nucleus, its nodes after them
  if ((rc = ensure_sort_keys(ctx, NT + (ctx->have_nuc ? ncap * ctx->nnuc : 0)))) return rc;
  hipLaunchKernelGGL(k_zero_f64, dim3(nb(3 * NT, 256)), dim3(256), 0, ctx->stream, ctx->d_vrep, 3 * NT);
  ZF_HIP(ctx, hipGetLastError());
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_ibm = true;
  return ZF_OK;
}

int zf_ibm_interpolate(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_ibm(ctx);
  return rc ? rc : ibm_interpolate_impl(ctx);
}

int zf_ibm_spread(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_ibm(ctx);
  return rc ? rc : ibm_spread_impl(ctx);
}

int zf_advance_cells(zf_ctx *ctx, int32_t *n_removed) {
  if (!ctx) return ZF_ERR_ARG;
  int rc = require_ibm(ctx);
  return rc ? rc : advance_impl(ctx, n_removed);
}

static int upload_node_field(zf_ctx *ctx, const double *h, double *dst) {
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  double *tmp = nullptr;
  int rc = dev_alloc(ctx, (void **)&tmp, 24 * (size_t)n);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(tmp, h, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_import_soa, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, tmp, n, NT, dst);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= 24 * (int64_t)n;
  dev_free(ctx, (void **)&tmp);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "node field upload: %s", hipGetErrorString(e));
  return ZF_OK;
}

int zf_upload_membrane_force(zf_ctx *ctx, const double *force) {
  if (!ctx || !force) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  return upload_node_field(ctx, force, ctx->d_mforce);
}

int zf_upload_repulsion(zf_ctx *ctx, const double *vrep) {
  if (!ctx || !vrep) return ZF_ERR_ARG;
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "zf_ibm_setup not called");
  return upload_node_field(ctx, vrep, ctx->d_vrep);
}

int zf_download_repulsion(zf_ctx *ctx, double *vrep) {
  if (!ctx || !vrep) return ZF_ERR_ARG;
  if (!ctx->have_ibm) return fail(ctx, ZF_ERR_STATE, "zf_ibm_setup not called");
  const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
  const int n = ctx->ncell * ctx->nn;
  if (n == 0) return ZF_OK;
  double *tmp = nullptr;
  int rc = dev_alloc(ctx, (void **)&tmp, 24 * (size_t)n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_export_soa, dim3(nb(n, 256)), dim3(256), 0, ctx->stream, ctx->d_vrep, n, NT, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(vrep, tmp, 24 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= 24 * (int64_t)n;
  dev_free(ctx, (void **)&tmp);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "repulsion download: %s", hipGetErrorString(e));
  return ZF_OK;
}

int zf_step(zf_ctx *ctx, int32_t iter) {
  if (!ctx) return ZF_ERR_ARG;
  const bool rbc = ctx->have_ibm && ctx->ncell > 0 && iter % ctx->icfg.rbc_interval == 0;
  int rc = zf_lbm_step(ctx, iter, rbc ? 1 : 0);
  if (rc || !rbc) return rc;
  return rbc_update(ctx, true, iter);
}

int zf_run(zf_ctx *ctx, int32_t iter0, int32_t nsteps) {
  if (!ctx || nsteps < 0) return ZF_ERR_ARG;
  const bool pairs = nsteps >= 2 && pair_eligible(ctx);
  auto rbc_iter = [&](int it) {
    return ctx->have_ibm && ctx->ncell > 0 && it % ctx->icfg.rbc_interval == 0;
  };
  for (int t = 1; t <= nsteps;) {
    const int it = iter0 + t;
    int rc;
    if (pairs && t < nsteps && !rbc_iter(it)) {
      // two LBM steps in one launch; the cell stages of the second, if it
      // is an RBC iteration, follow as in zf_step
      const bool rbc2 = rbc_iter(it + 1);
      if ((rc = step_pair(ctx, it, rbc2 ? 1 : 0))) return rc;
      if (rbc2 && (rc = rbc_update(ctx, true, it + 1))) return rc;
      t += 2;
    } else {
      if ((rc = zf_step(ctx, it))) return rc;
      t++;
    }
  }
  return ZF_OK;
}

}  // extern "C"
