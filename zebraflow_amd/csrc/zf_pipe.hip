// zf_pipe.hip -- two LBM time steps per launch (temporal blocking).
//
// The one-step kernel (zf_lbm.hip) moves every population through HBM
// twice per step: read by the pull, written by the collision.  Beyond the
// 256 MB last-level cache that traffic sets the step time.  This launch
// advances the lattice by two steps, t -> t+1 -> t+2, so that the state t+1
// never goes to HBM: it lives in a small ring of population slots that stays
// in the L2 / last-level cache, and only state t is read from and state t+2
// written to the population buffers.  The arithmetic per node is the
// one-step kernel's (bgk_guo, write_bc): the populations after the launch are
// the same bits as after two one-step launches.
//
// Work items: blocks of 64 consecutive interior storage indices (one wave
// each); phase 0
// (the step t -> t+1 of one block) and phase 1 (t+1 -> t+2).  A phase-1 item
// needs the phase-0 items of every block its nodes pull from (the chords
// reach at most `ghost` blocks away) and of the fused writers of the
// boundary nodes they pull.  A phase-0 item reuses the ring slot of the item
// `ring` places earlier and waits for that slot's readers.
//
// Each XCD runs its own sequence: the interior blocks are split into one
// contiguous range per XCD, and its phase-0 items cover that range widened
// by `ghost` blocks on both sides (those blocks' t+1 states are computed
// twice, by both neighbouring XCDs, with identical bits), so every item
// waits only for items of its own XCD.  The waves are persistent: each reads
// the XCD it runs on (hardware register XCC_ID), takes the next ticket of
// that XCD's sequence with an atomic counter, waits for the tickets it
// depends on (all earlier in the sequence, so taken by running waves),
// processes the item and publishes its completion flag.  Waves never wait
// for each other at a barrier: the item, the ticket and every branch around
// a wait are wave-uniform, and the one-lane operations (ticket, flag,
// count) are atomics every lane joins with a neutral operand, so no lane of
// a wave can run ahead of the others.  Data written by one wave is read by
// another of the same XCD through that XCD's L2: producers drain their
// stores (s_waitcnt) before the flag, and consumers load ring slots with
// the sc0 policy (a miss in the CU's L1, a hit in the L2).  Flags and
// tickets are agent-scope atomics.  Every wait is bounded in time; a wait
// that runs out sets an error word that zf_sync reports.
#include "zf_d3q19.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace zf {

using namespace d3q19;

namespace {

constexpr int kPipeNodes = 64;    // interior nodes per item (one wave)
constexpr int kPipeShift = 6;
constexpr int kPipeMaxXcd = 8;
// Control words of one XCD, in a 256-byte block of its own (words that
// different XCDs update must not share a cache line): ticket, items
// processed, error word, items expected.
constexpr int kCtlStride = 64;
constexpr int kCtlTick = 0, kCtlDone = 16, kCtlErr = 32, kCtlExpect = 48;
// longest wait, in ticks of the 100 MHz s_memrealtime clock (50 ms)
constexpr unsigned long long kPipeWaitTicks = 5000000ull;
// Tickets by which phase-0 items run ahead of the phase-1 items that need
// them (beyond the dependency itself): about the waves one XCD keeps in
// flight (32 CUs x 32 waves), so that a phase-1 item rarely finds the
// items it reads still running.  The ring then holds twice that plus the
// pull reach, so that a phase-0 item rarely waits for the readers of the
// slot it reuses.
constexpr int kPipeLead = 1024;
// tickets taken by one atomic, run in order by the wave (2 and 4 were
// measured slower than 1)
constexpr int kPipeChunk = 1;

struct PipeXcd {
  int seq_off, nitems;  // tickets of this XCD in seq
  int b1_0, n1;         // phase-0 items: blocks b1_0 .. b1_0 + n1 - 1
  int bx0, n2;          // phase-1 items: blocks bx0 .. bx0 + n2 - 1
  int f1_off, f1_p;     // completion flags of its phase-0 items: item k at f1_off + (k & 255) * f1_p + (k >> 8)
  int f2_off, f2_p;     // the same for its phase-1 items
};

// Flag of item k in a region: consecutive items on different cache lines
// (p >= 32 ints), so the waves publishing neighbouring items and those
// polling them do not queue on one line.
__host__ __device__ __forceinline__ int flag_slot(int off, int p, int k) {
  return off + (k & 255) * p + (k >> 8);
}

struct PipeArgs {
  int n_int, n_fused, nx;
  size_t stride;
  const uint32_t *word;
  const int32_t *block0, *chords;
  const BcEntry *bc;
  const int32_t *fused_s, *att_start;
  void *scratch;                // rings (one per XCD) then the side buffer of BC nodes
  int ring_elems;               // slots per direction in one XCD's ring (ring * 64)
  int ring_blocks;              // ring (power of two)
  int side_elems;               // per direction in the side buffer (stride - n_int)
  // per ticket: item << 1 | phase, and the items of the other phase it
  // waits for (an empty range: hi < lo)
  const int4 *tix;
  int *flags;  // completion flags (the launch's epoch when done)
  int epoch;
  int *ctl;  // kCtlStride ints per XCD
  int xrank[kPipeMaxXcd];
  PipeXcd px[kPipeMaxXcd];
};

__device__ __forceinline__ int xcc_id() {
  // HW_REG_XCC_ID (hardware register 20), bits 0..3
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));
}

// The whole wave: until the flags of items [lo, hi] of a region hold
// `epoch` (flags only grow: a later launch's epoch is larger).  Bounded in
// time: a wait that runs out, or that sees the XCD's error word set by
// another one, stops waiting (the launch's results are then void and
// zf_sync reports it), so every launch ends.  All exits are wave-uniform.
__device__ void wait_flags(const int *flags, int off, int p, int lo, int hi, int epoch, int *err) {
  const int lane = threadIdx.x & 63;
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  for (int base = lo; base <= hi; base += 64) {
    const int k = base + lane;
    const int *f = flags + flag_slot(off, p, k <= hi ? k : lo);
    for (int spins = 1;; spins++) {
      const int v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__all(v >= epoch)) break;
      if ((spins & 15) == 0) {
        const int e = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (e) return;
        if (__builtin_amdgcn_s_memrealtime() - t0 > kPipeWaitTicks) {
          __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          return;
        }
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

// Load of one ring or side-buffer element (element offset e in the scratch
// buffer) written by another workgroup of this XCD: sc0, a miss in the L1
// and a hit in the L2.
template <class T>
__device__ __forceinline__ T ld_l2(__amdgpu_buffer_rsrc_t r, int e) {
  const int off = e * (int)sizeof(T);
  if constexpr (sizeof(T) == 2)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 1));
  else if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 1));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 1));
}

// Ring slot of interior storage index s in the ring of an XCD whose first
// phase-0 block is b1_0.
__device__ __forceinline__ int ring_slot(int s, int b1_0, int ring_blocks) {
  return ((((s >> kPipeShift) - b1_0) & (ring_blocks - 1)) << kPipeShift) | (s & (kPipeNodes - 1));
}

template <int S, bool FIELD, int KT, bool ACC>
__global__ __launch_bounds__(256) void k_stream_pair(
    const typename Store<S>::T *__restrict__ A, typename Store<S>::T *__restrict__ B,
    const typename Store<S>::F *__restrict__ force, StepParams prm1, BcParams bp1,
    StepParams prm2, BcParams bp2, PipeArgs a) {
  using R = typename Store<S>::R;
  using Sx = Store<S>;
  using T = typename Sx::T;
  using CT = typename std::conditional<ACC, double, R>::type;
  constexpr bool NT_ALL = KT == 0;
  const int lane = threadIdx.x & 63;
  const int xcc = xcc_id();
  const int xr = xcc < kPipeMaxXcd ? a.xrank[xcc] : -1;
  if (xr < 0) return;  // an XCD without a sequence (not seen at set-up)
  const PipeXcd px = a.px[xr];
  T *ring = static_cast<T *>(a.scratch) + (size_t)xr * Q * a.ring_elems;
  T *side = static_cast<T *>(a.scratch) + (size_t)a.nx * Q * a.ring_elems;
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(a.scratch, 0, 0x7fffffff, 0x00020000);
  const int side0 = a.nx * Q * a.ring_elems;  // element offset of the side buffer
  const int ring0 = xr * Q * a.ring_elems;      // element offset of this XCD's ring
  const size_t stride = a.stride;
  // tickets of this XCD, kPipeChunk at a time (lane 0 adds, the others add
  // zero), run in order; the next chunk is taken while this one runs, so
  // its latency is hidden
  auto take = [&]() {
    return __hip_atomic_fetch_add(a.ctl + xr * kCtlStride + kCtlTick, lane == 0 ? kPipeChunk : 0,
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  int t0 = __builtin_amdgcn_readfirstlane(take());
  int done_items = 0;
  while (t0 < px.nitems) {
   const int tn = take();
   for (int j = 0; j < kPipeChunk && t0 + j < px.nitems; j++) {
    const int4 ti = a.tix[px.seq_off + t0 + j];
    const int item = __builtin_amdgcn_readfirstlane(ti.x);
    const int wlo = __builtin_amdgcn_readfirstlane(ti.z), whi = __builtin_amdgcn_readfirstlane(ti.w);
    const int phase = item & 1, idx = item >> 1;
    const int blk = phase ? px.bx0 + idx : px.b1_0 + idx;
    const int s = blk * kPipeNodes + lane;
    const bool live = s < a.n_int;
    // the node's word and chord do not depend on the wait: loaded before it
    const uint32_t w = live ? a.word[s] : 0u;
    int d[20];
    {
      const int ch = live ? a.block0[s >> 6] + (int)(w >> kWordChordShift) : 0;
      const int4 *cr = reinterpret_cast<const int4 *>(a.chords + (size_t)ch * kChordInts);
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const int4 x = cr[v];
        d[4 * v] = x.x; d[4 * v + 1] = x.y; d[4 * v + 2] = x.z; d[4 * v + 3] = x.w;
      }
    }
    if (phase)
      wait_flags(a.flags, px.f1_off, px.f1_p, wlo, whi, a.epoch, a.ctl + xr * kCtlStride + kCtlErr);
    else
      wait_flags(a.flags, px.f2_off, px.f2_p, wlo, whi, a.epoch, a.ctl + xr * kCtlStride + kCtlErr);
    if (live) {
      CT g[Q];
      if (phase == 0) {
        for_dirs([&](auto cc) {
          constexpr int c = decltype(cc)::value;
          size_t id = (size_t)c * stride + s;
          if constexpr (c < 18) id = ((w >> c) & 1u) ? (size_t)(c ^ 1) * stride + s : id + d[c];
          const T raw = ld_policy<(c >= KT)>(A + id);
          g[c] = (CT)Sx::ld(&raw, 0, WQ[c]);
        });
      } else {
        // t+1 sources: interior nodes in this XCD's ring, boundary nodes in
        // the side buffer (element offsets within the scratch buffer)
        const int self = ring0 + ring_slot(s, px.b1_0, a.ring_blocks);
        for_dirs([&](auto cc) {
          constexpr int c = decltype(cc)::value;
          int e;
          if constexpr (c == 18) {
            e = self + c * a.ring_elems;
          } else {
            const int s2 = s + d[c];
            const int ei = ring0 + c * a.ring_elems + ring_slot(s2, px.b1_0, a.ring_blocks);
            const int eb = side0 + c * a.side_elems + (s2 - a.n_int);
            e = ((w >> c) & 1u) ? self + (c ^ 1) * a.ring_elems : (s2 < a.n_int ? ei : eb);
          }
          const T raw = ld_l2<T>(rsrc, e);
          g[c] = (CT)Sx::ld(&raw, 0, WQ[c]);
        });
      }
      const StepParams &prm = phase ? prm2 : prm1;
      CT fx = (CT)prm.g[0], fy = (CT)prm.g[1], fz = (CT)prm.g[2];
      if (FIELD) {
        fx += (CT)force[s];
        fy += (CT)force[stride + s];
        fz += (CT)force[2 * stride + s];
      }
      CT drho, ux, uy, uz;
      const int rs = ring_slot(s, px.b1_0, a.ring_blocks);
      bgk_guo<CT>(g, fx, fy, fz, (CT)prm.dt, (CT)prm.visc_scale, d[18], drho, ux, uy, uz,
                  [&](auto cc, CT v) {
                    constexpr int c = decltype(cc)::value;
                    if (phase == 0) ring[c * a.ring_elems + rs] = Sx::st((R)v, WQ[c]);
                    else st_policy<(c >= KT)>(B + (size_t)c * stride + s, Sx::st((R)v, WQ[c]));
                  });
      if (w & kWordMacro) {
        int lo = 0, hi = a.n_fused - 1, k = -1;
        while (lo <= hi) {
          const int mid = (lo + hi) >> 1, v = a.fused_s[mid];
          if (v == s) { k = mid; break; }
          if (v < s) lo = mid + 1; else hi = mid - 1;
        }
        if (k >= 0)
          for (int t = a.att_start[k]; t < a.att_start[k + 1]; t++) {
            if (phase == 0)
              write_bc<S>(side, a.side_elems, a.bc[t], (R)drho, (R)ux, (R)uy, (R)uz, bp1, false, a.n_int);
            else
              write_bc<S>(B, stride, a.bc[t], (R)drho, (R)ux, (R)uy, (R)uz, bp2, NT_ALL);
          }
      }
    }
    // publish: a phase-0 item's ring and side stores have reached the L2
    // (its readers are other waves), then its completion bit (an atomic OR
    // every lane joins, only lane 0 with the bit); a phase-1 item's bit only
    // says that its ring reads are done
    if (phase == 0) __builtin_amdgcn_s_waitcnt(0);
    const int own = phase ? flag_slot(px.f2_off, px.f2_p, idx) : flag_slot(px.f1_off, px.f1_p, idx);
    __hip_atomic_fetch_max(a.flags + own, lane == 0 ? a.epoch : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    done_items++;
   }
   t0 = __builtin_amdgcn_readfirstlane(tn);
  }
  __hip_atomic_fetch_add(a.ctl + xr * kCtlStride + kCtlDone, lane == 0 ? done_items : 0, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// Before each launch: the item count of the previous launch checked, the
// tickets and counts cleared (a kernel, so the sequence can be captured in
// a graph).
__global__ void k_pipe_reset(int *ctl) {
  const int x = threadIdx.x;
  if (x >= kPipeMaxXcd) return;
  int *c = ctl + x * kCtlStride;
  if (c[kCtlDone] != c[kCtlExpect]) c[kCtlErr] |= 2;
  c[kCtlTick] = 0;
  c[kCtlDone] = 0;
}

// Listed nodes that interior nodes pull but no step writes (their value is
// the same in both buffers): copied from state t into the side buffer
// before each launch, so that phase 1 reads every t+1 source from the ring
// or the side buffer.
template <class T>
__global__ void k_side_fill(const T *__restrict__ A, size_t stride, const int32_t *__restrict__ list,
                            int n, T *__restrict__ side, int side_elems, int n_int) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = list[i];
#pragma unroll
  for (int c = 0; c < Q; c++) side[(size_t)c * side_elems + (x - n_int)] = A[(size_t)c * stride + x];
}

// Set-up probe: the XCD each workgroup of a wide launch runs on.
__global__ void k_xcc_probe(int *out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

__global__ void k_fill_i32_pipe(int *a, size_t n, int v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

inline unsigned nblk_of(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// Host tables of the two-step launch, built once per lattice.
struct PipeState {
  bool tried = false, ok = false;
  std::string why;  // when not ok
  int nx = 0, ghost = 0, ring_blocks = 0, nblk = 0;
  int xrank[kPipeMaxXcd];
  PipeXcd px[kPipeMaxXcd];
  int4 *d_tix = nullptr;  // per ticket: item, own flag, wait range
  int32_t *d_fill = nullptr;  // pulled, never written listed nodes
  int n_fill = 0;
  void *d_scratch = nullptr;
  int *d_ctl = nullptr;  // tick[8], processed[8], err[8], expected[8], then flags
  int *d_flags = nullptr;  // completion flags of every XCD's items
  size_t nflags = 0;
  int epoch = 0;
  int grid = 0;
  size_t ring_elems = 0, side_elems = 0;
};

void free_pipe(zf_ctx *ctx) {
  PipeState *p = ctx->pipe;
  if (!p) return;
  dev_free(ctx, (void **)&p->d_tix);
  dev_free(ctx, (void **)&p->d_fill);
  dev_free(ctx, &p->d_scratch);
  dev_free(ctx, (void **)&p->d_ctl);
  dev_free(ctx, (void **)&p->d_flags);
  delete p;
  ctx->pipe = nullptr;
}

namespace {

template <class T>
std::vector<T> download(const T *d, size_t n) {
  std::vector<T> h(n);
  if (n) (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
  return h;
}

// Builds the tables; on a lattice the two-step launch cannot take, records
// why and leaves ok false (the one-step kernel is used).
int pipe_prepare(zf_ctx *ctx) {
  if (!ctx->pipe) ctx->pipe = new PipeState();
  PipeState &P = *ctx->pipe;
  if (P.tried) return ZF_OK;
  P.tried = true;
  auto no = [&](const char *why) { P.why = why; P.ok = false; return ZF_OK; };
  if (ctx->n_int <= 0) return no("no interior nodes");
  if (ctx->n_bc != ctx->n_bc_fused) return no("boundary entries with a non-interior neighbour");
  const int n_int = ctx->n_int, nblk = (n_int + kPipeNodes - 1) / kPipeNodes;
  P.nblk = nblk;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));

  // XCDs this device's workgroups run on
  {
    const int nprobe = 4096;
    int *d = nullptr;
    int rc = dev_alloc(ctx, (void **)&d, sizeof(int) * nprobe);
    if (rc) return rc;
    hipLaunchKernelGGL(k_xcc_probe, dim3(nprobe), dim3(64), 0, ctx->stream, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    std::vector<int> h(nprobe, -1);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, sizeof(int) * nprobe, hipMemcpyDeviceToHost);
    dev_free(ctx, (void **)&d);
    if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "XCD probe: %s", hipGetErrorString(e));
    bool seen[16] = {};
    for (int v : h) {
      if (v < 0 || v >= kPipeMaxXcd) return no("unexpected XCC_ID");
      seen[v] = true;
    }
    P.nx = 0;
    for (int x = 0; x < kPipeMaxXcd; x++) P.xrank[x] = seen[x] ? P.nx++ : -1;
  }
  const int nx = P.nx;

  // node tables back on the host
  const std::vector<uint32_t> word = download(ctx->d_word, (size_t)n_int);
  const std::vector<int32_t> block0 = download(ctx->d_block_chord0, (size_t)(n_int + 63) / 64 + 1);
  const std::vector<int32_t> chords = download(ctx->d_chords, (size_t)ctx->n_chords * kChordInts);
  const std::vector<BcEntry> bce = download(ctx->d_bc, (size_t)ctx->n_bc);
  const size_t nside = ctx->stride - (size_t)n_int;
  // writer block of each boundary node written in phase 0 (fused entries)
  std::vector<int32_t> writer(nside, -1);
  for (int t = 0; t < ctx->n_bc_fused; t++) {
    const BcEntry &e = bce[t];
    if (e.s < n_int || (size_t)(e.s - n_int) >= nside) return no("boundary entry inside the interior range");
    writer[e.s - n_int] = e.sm >> kPipeShift;
  }
  std::vector<int2> dep(nblk);
  std::vector<uint8_t> pulled(nside, 0);  // pulled and never written
  int ghost = 0;
  for (int b = 0; b < nblk; b++) dep[b] = make_int2(b, b);
  for (int s = 0; s < n_int; s++) {
    const uint32_t w = word[s];
    const int ch = block0[s >> 6] + (int)(w >> kWordChordShift);
    const int b = s >> kPipeShift;
    for (int c = 0; c < 18; c++) {
      if ((w >> c) & 1u) continue;
      const int s2 = s + chords[(size_t)ch * kChordInts + c];
      int src;
      if (s2 < n_int) src = s2 >> kPipeShift;
      else if (writer[s2 - n_int] >= 0) src = writer[s2 - n_int];
      else {
        pulled[s2 - n_int] = 1;  // never written: copied from state t
        continue;
      }
      dep[b].x = std::min(dep[b].x, src);
      dep[b].y = std::max(dep[b].y, src);
      ghost = std::max(ghost, std::abs(src - b));
    }
  }
  P.ghost = ghost;
  if (ghost > nblk) return no("pull distance beyond the lattice");

  // per-XCD ranges and ticket sequences
  std::vector<int32_t> seq;
  int f1 = 0, f2 = 0;
  int ring = 1;
  while (ring < 2 * kPipeLead + 4 * ghost + 64) ring *= 2;
  for (;; ring *= 2) {
    if (ring > (1 << 16)) return no("no ring size fits the schedule");
    seq.clear();
    f1 = f2 = 0;
    bool fits = true;
    for (int x = 0; x < nx && fits; x++) {
      PipeXcd &q = P.px[x];
      const int bx0 = (int)((int64_t)nblk * x / nx), bx1 = (int)((int64_t)nblk * (x + 1) / nx);
      const int b1_0 = std::max(0, bx0 - ghost), b1_1 = std::min(nblk, bx1 + ghost);
      q.seq_off = (int)seq.size();
      q.b1_0 = b1_0;
      q.n1 = b1_1 - b1_0;
      q.bx0 = bx0;
      q.n2 = bx1 - bx0;
      // flag regions: 256 columns of p >= 32 ints (one cache line apart)
      q.f1_p = std::max(32, (q.n1 + 255) / 256);
      q.f2_p = std::max(32, (q.n2 + 255) / 256);
      q.f1_off = f1;
      q.f2_off = f2;
      f1 += 256 * q.f1_p;
      f2 += 256 * q.f2_p;
      int u = 0, v = 0;  // phase-0 and phase-1 items issued so far
      while (u < q.n1 || v < q.n2) {
        // phase 1 of block bx0 + v: every phase-0 item it reads issued
        // (hard), and preferably kPipeLead tickets before it (soft)
        const int need = v < q.n2 ? dep[bx0 + v].y - b1_0 : 0;
        const bool v_hard = v < q.n2 && need < u;
        const bool v_soft = v < q.n2 && (u >= q.n1 || need + kPipeLead < u);
        // phase 0 item u reuses the ring slot of item u - ring: every
        // phase-1 reader of that item (blocks within ghost of it) issued
        const bool u_ok = u < q.n1 && (u < ring || v > std::min(b1_0 + u - ring + ghost - bx0, q.n2 - 1));
        if (v_soft) {
          seq.push_back((v << 1) | 1);
          v++;
        } else if (u_ok) {
          seq.push_back(u << 1);
          u++;
        } else if (v_hard) {
          seq.push_back((v << 1) | 1);
          v++;
        } else {
          fits = false;  // the ring is too small for this schedule
          break;
        }
      }
      q.nitems = (int)seq.size() - q.seq_off;
    }
    if (fits) break;
  }
  P.ring_blocks = ring;
  // check: every dependency of every ticket is an earlier ticket of its XCD
  for (int x = 0; x < nx; x++) {
    const PipeXcd &q = P.px[x];
    std::vector<int> pos1(q.n1, -1), pos2(q.n2, -1);
    for (int t = 0; t < q.nitems; t++) {
      const int it = seq[q.seq_off + t];
      if (it & 1) pos2[it >> 1] = t; else pos1[it >> 1] = t;
    }
    for (int t = 0; t < q.nitems; t++) {
      const int it = seq[q.seq_off + t], i = it >> 1;
      if (it & 1) {
        const int2 dr = dep[q.bx0 + i];
        if (dr.x < q.b1_0 || dr.y >= q.b1_0 + q.n1) return no("dependency outside the XCD's range");
        for (int u = dr.x - q.b1_0; u <= dr.y - q.b1_0; u++)
          if (pos1[u] < 0 || pos1[u] >= t) return no("schedule check failed (phase 1)");
      } else if (i >= ring) {
        const int bold = q.b1_0 + i - ring;
        const int lo = std::max(bold - ghost - q.bx0, 0), hi = std::min(bold + ghost - q.bx0, q.n2 - 1);
        for (int v = lo; v <= hi; v++)
          if (pos2[v] < 0 || pos2[v] >= t) return no("schedule check failed (ring reuse)");
      }
    }
  }
  P.ring_elems = (size_t)ring * kPipeNodes;
  P.side_elems = nside;
  const size_t eb = store_bytes(ctx->cfg.storage);
  const size_t scratch_elems = (size_t)nx * Q * P.ring_elems + (size_t)Q * nside;
  if (scratch_elems * eb >= (1ull << 31)) return no("ring and side buffer beyond 2 GiB");

  std::vector<int32_t> fill;
  for (size_t i = 0; i < nside; i++)
    if (pulled[i]) fill.push_back(n_int + (int32_t)i);

  int rc;
  // per-ticket records: the item, and the items of the other phase it
  // waits for (an empty range: hi < lo); the phase-1 flag regions follow
  // the phase-0 ones
  for (int x = 0; x < nx; x++) P.px[x].f2_off += f1;
  std::vector<int4> tix(seq.size());
  for (int x = 0; x < nx; x++) {
    const PipeXcd &q = P.px[x];
    for (int t = 0; t < q.nitems; t++) {
      const int it = seq[q.seq_off + t], i = it >> 1;
      int4 r = make_int4(it, 0, 0, -1);
      if (it & 1) {
        const int2 dr = dep[q.bx0 + i];
        r.z = dr.x - q.b1_0;
        r.w = dr.y - q.b1_0;
      } else if (i >= ring) {
        const int bold = q.b1_0 + i - ring;
        r.z = std::max(bold - ghost - q.bx0, 0);
        r.w = std::min(bold + ghost - q.bx0, q.n2 - 1);
      }
      tix[q.seq_off + t] = r;
    }
  }
  if ((rc = dev_alloc(ctx, (void **)&P.d_tix, sizeof(int4) * tix.size()))) return rc;
  if ((rc = dev_alloc(ctx, (void **)&P.d_fill, sizeof(int32_t) * std::max<size_t>(fill.size(), 1))))
    return rc;
  if ((rc = dev_alloc(ctx, &P.d_scratch, scratch_elems * eb))) return rc;
  const size_t nctl = (size_t)kCtlStride * kPipeMaxXcd;
  if ((rc = dev_alloc(ctx, (void **)&P.d_ctl, sizeof(int) * nctl))) return rc;
  P.nflags = (size_t)f1 + (size_t)f2;
  if ((rc = dev_alloc(ctx, (void **)&P.d_flags, sizeof(int) * P.nflags))) return rc;
  hipLaunchKernelGGL(k_fill_i32_pipe, dim3(nblk_of(P.nflags, 256)), dim3(256), 0, ctx->stream, P.d_flags,
                     P.nflags, 0);
  ZF_HIP(ctx, hipMemcpy(P.d_tix, tix.data(), sizeof(int4) * tix.size(), hipMemcpyHostToDevice));
  if (!fill.empty())
    ZF_HIP(ctx, hipMemcpy(P.d_fill, fill.data(), sizeof(int32_t) * fill.size(), hipMemcpyHostToDevice));
  P.n_fill = (int)fill.size();
  hipLaunchKernelGGL(k_fill_i32_pipe, dim3(nblk_of(nctl, 256)), dim3(256), 0, ctx->stream, P.d_ctl, nctl, 0);
  ZF_HIP(ctx, hipGetLastError());
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // item counts per XCD: expected, and as processed by a launch before the
  // first, so that the first reset's check passes
  std::vector<int> ctl((size_t)kCtlStride * kPipeMaxXcd, 0);
  for (int x = 0; x < nx; x++)
    ctl[(size_t)x * kCtlStride + kCtlDone] = ctl[(size_t)x * kCtlStride + kCtlExpect] = P.px[x].nitems;
  ZF_HIP(ctx, hipMemcpy(P.d_ctl, ctl.data(), sizeof(int) * ctl.size(), hipMemcpyHostToDevice));
  P.epoch = 0;

  // persistent grid: every workgroup slot of the device
  hipDeviceProp_t prop;
  ZF_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
  P.grid = prop.multiProcessorCount * 8;
  P.ok = true;
  return ZF_OK;
}

template <int S, bool FIELD, int KT, bool ACC>
void launch_pair(zf_ctx *ctx, const void *src, void *dst, const StepParams &p1, const BcParams &b1,
                 const StepParams &p2, const BcParams &b2) {
  PipeState &P = *ctx->pipe;
  PipeArgs a{};
  a.n_int = ctx->n_int;
  a.n_fused = ctx->n_fused_nodes;
  a.nx = P.nx;
  a.stride = ctx->stride;
  a.word = ctx->d_word;
  a.block0 = ctx->d_block_chord0;
  a.chords = ctx->d_chords;
  a.bc = ctx->d_bc;
  a.fused_s = ctx->d_fused_s;
  a.att_start = ctx->d_att_start;
  a.scratch = P.d_scratch;
  a.ring_elems = (int)P.ring_elems;
  a.ring_blocks = P.ring_blocks;
  a.side_elems = (int)P.side_elems;
  a.tix = P.d_tix;
  a.ctl = P.d_ctl;
  a.flags = P.d_flags;
  a.epoch = P.epoch;
  for (int x = 0; x < kPipeMaxXcd; x++) a.xrank[x] = P.xrank[x];
  for (int x = 0; x < P.nx; x++) a.px[x] = P.px[x];
  using T = typename Store<S>::T;
  if (P.n_fill)
    hipLaunchKernelGGL(k_side_fill<T>, dim3(nblk_of(P.n_fill, 256)), dim3(256), 0, ctx->stream,
                       static_cast<const T *>(src), ctx->stride, P.d_fill, P.n_fill,
                       static_cast<T *>(P.d_scratch) + (size_t)P.nx * Q * P.ring_elems,
                       (int)P.side_elems, ctx->n_int);
  hipLaunchKernelGGL((k_stream_pair<S, FIELD, KT, ACC>), dim3(P.grid), dim3(256), 0, ctx->stream,
                     static_cast<const T *>(src), static_cast<T *>(dst),
                     FIELD ? static_cast<const typename Store<S>::F *>(ctx->d_force) : nullptr, p1, b1,
                     p2, b2, a);
}

template <int S, bool FIELD, bool ACC>
void launch_pair_kt(zf_ctx *ctx, int kt, const void *src, void *dst, const StepParams &p1,
                    const BcParams &b1, const StepParams &p2, const BcParams &b2) {
#define ZF_PAIR_KT(K) \
  case K: launch_pair<S, FIELD, K, ACC>(ctx, src, dst, p1, b1, p2, b2); break;
  switch (kt) {
    ZF_PAIR_KT(0) ZF_PAIR_KT(2) ZF_PAIR_KT(4) ZF_PAIR_KT(6) ZF_PAIR_KT(8) ZF_PAIR_KT(10)
    ZF_PAIR_KT(12) ZF_PAIR_KT(14) ZF_PAIR_KT(16) ZF_PAIR_KT(18)
    default: launch_pair<S, FIELD, Q, ACC>(ctx, src, dst, p1, b1, p2, b2); break;
  }
#undef ZF_PAIR_KT
}

}  // namespace

bool pair_eligible(zf_ctx *ctx) {
  const uint32_t fl = (uint32_t)ctx->cfg.flags;
  if (!ctx->have_lattice || !ctx->d_f[0]) return false;
  if (fl & ZF_FLAG_SINGLE_STEPS) return false;
  if (ctx->link_main || !ctx->h_export.empty() || !ctx->linked.empty()) return false;
  if (!(fl & ZF_FLAG_PAIRED_STEPS) && !pair_default(ctx)) return false;
  if (pipe_prepare(ctx) != ZF_OK) return false;
  return ctx->pipe && ctx->pipe->ok;
}

int step_pair(zf_ctx *ctx, int iter1, int rbc2) {
  PipeState &P = *ctx->pipe;
  StepParams p1, p2;
  BcParams b1, b2;
  step_params(ctx, iter1, p1, b1);
  step_params(ctx, iter1 + 1, p2, b2);
  hipLaunchKernelGGL(k_pipe_reset, dim3(1), dim3(64), 0, ctx->stream, P.d_ctl);
  P.epoch++;
  const void *src = ctx->d_f[ctx->cur];
  void *dst = ctx->d_f[ctx->cur ^ 1];
  const int kt = pair_cached_dirs(ctx);
  const bool acc = ctx->cfg.storage != ZF_STORE_F64 && (ctx->cfg.flags & ZF_FLAG_F64_ARITH);
  const bool field = ctx->d_force != nullptr;
  switch (ctx->cfg.storage) {
    case ZF_STORE_F64:
      if (field) launch_pair_kt<ZF_STORE_F64, true, false>(ctx, kt, src, dst, p1, b1, p2, b2);
      else launch_pair_kt<ZF_STORE_F64, false, false>(ctx, kt, src, dst, p1, b1, p2, b2);
      break;
    case ZF_STORE_F32DEV:
      if (field) {
        if (acc) launch_pair_kt<ZF_STORE_F32DEV, true, true>(ctx, kt, src, dst, p1, b1, p2, b2);
        else launch_pair_kt<ZF_STORE_F32DEV, true, false>(ctx, kt, src, dst, p1, b1, p2, b2);
      } else {
        if (acc) launch_pair_kt<ZF_STORE_F32DEV, false, true>(ctx, kt, src, dst, p1, b1, p2, b2);
        else launch_pair_kt<ZF_STORE_F32DEV, false, false>(ctx, kt, src, dst, p1, b1, p2, b2);
      }
      break;
    default:
      if (field) {
        if (acc) launch_pair_kt<ZF_STORE_F16DEV, true, true>(ctx, kt, src, dst, p1, b1, p2, b2);
        else launch_pair_kt<ZF_STORE_F16DEV, true, false>(ctx, kt, src, dst, p1, b1, p2, b2);
      } else {
        if (acc) launch_pair_kt<ZF_STORE_F16DEV, false, true>(ctx, kt, src, dst, p1, b1, p2, b2);
        else launch_pair_kt<ZF_STORE_F16DEV, false, false>(ctx, kt, src, dst, p1, b1, p2, b2);
      }
      break;
  }
  if (rbc2 && ctx->d_force) {
    const size_t nw = 3 * ctx->stride * field_elem(ctx) / 4;
    hipLaunchKernelGGL(k_fill_i32_pipe, dim3(nblk_of(nw, 256)), dim3(256), 0, ctx->stream,
                       static_cast<int *>(ctx->d_force), nw, 0);
  }
  ZF_HIP(ctx, hipGetLastError());
  ctx->cur ^= 1;
  ctx->pipe_used = true;
  return ZF_OK;
}

// The error words of the two-step launches since the last check (after a
// synchronisation of the stream): a bounded wait that ran out, or a launch
// that did not process all its items.
int pipe_check(zf_ctx *ctx) {
  if (!ctx->pipe_used || !ctx->pipe || !ctx->pipe->ok) return ZF_OK;
  ctx->pipe_used = false;
  PipeState &P = *ctx->pipe;
  std::vector<int> h((size_t)kCtlStride * kPipeMaxXcd);
  ZF_HIP(ctx, hipMemcpy(h.data(), P.d_ctl, sizeof(int) * h.size(), hipMemcpyDeviceToHost));
  int err = 0;
  for (int x = 0; x < P.nx; x++) {
    err |= h[(size_t)x * kCtlStride + kCtlErr];
    if (h[(size_t)x * kCtlStride + kCtlDone] != P.px[x].nitems) err |= 2;
  }
  if (!err) return ZF_OK;
  // clear the error words, so a later check reports only new failures
  for (int x = 0; x < kPipeMaxXcd; x++) h[(size_t)x * kCtlStride + kCtlErr] = 0;
  ZF_HIP(ctx, hipMemcpy(P.d_ctl, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
  return fail(ctx, ZF_ERR_HIP, "two-step launch: %s", (err & 1) ? "a dependency wait ran out"
                                                                 : "items left unprocessed");
}

}  // namespace zf

using namespace zf;

extern "C" int zf_pair_info(zf_ctx *ctx, int32_t info[6]) {
  if (!ctx || !info) return ZF_ERR_ARG;
  for (int k = 0; k < 6; k++) info[k] = 0;
  if (!ctx->have_lattice || !ctx->d_f[0]) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  const int rc = pipe_prepare(ctx);
  if (rc) return rc;
  const PipeState &P = *ctx->pipe;
  info[0] = pair_eligible(ctx) ? 1 : 0;
  info[1] = P.ok ? 1 : 0;
  info[2] = P.nx;
  info[3] = P.ghost;
  info[4] = P.ring_blocks;
  info[5] = P.grid;
  return ZF_OK;
}
