// zf_wss.hip -- wall shear stress and wall pressure at wall-mesh points.
//
// Replaces the reference's post-processing operator
// (PostProWSSandPressure.c:267-369), computed on the device from the
// current populations instead of from a FLUIDdomain CSV:
//   wall pressure      nearest lattice node with B >= 0 closer than 1.42 DX,
//                      p = (D - 1) * LBM_PRESSURE_SCALE - Pgauge
//   wall shear stress  probe x + DX n, unnormalised cosine kernel over
//                      |d| <= 2 DX on nodes 0.11 < B < 4997,
//                      tau = -(mu/DX) (u - (u.n) n)
// with the post-processor's window: index (int)((x-X0)/DX + 1.01) +- 3, i
// clipped to [3, IM] and j, k to [1, JM], [1, KM] (its "else if" pairs clip
// one side only; indices outside the box are skipped).  One lane per point,
// FP64 throughout.  The velocities are the unshifted macro velocities of the
// listed nodes (ZebraflowWT.c:6606-6614), the ones the reference writes out.
#include "zf_wss_dev.h"

#include <cmath>
#include <mutex>

namespace zf {
namespace {

using namespace d3q19;
using namespace wssdev;

inline unsigned nb(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Scratch of the allocation-free entry point zf_wss_noalloc: per point the
// inputs (xyz, normal: 6 doubles), the outputs (wss, pressure: 4 doubles)
// and the valid flag (4 bytes), 84 bytes, for up to ZF_WSS_NOALLOC_MAX_POINTS
// points.  Each device holds its own copy for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the device (the array's name is not a device address in host code).
constexpr size_t kWssPointBytes = 84;
constexpr size_t kWssScratchBytes = kWssPointBytes * ZF_WSS_NOALLOC_MAX_POINTS;
__device__ __align__(256) unsigned char scratch[kWssScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

template <int S>
__global__ void k_wss(const typename Store<S>::T *__restrict__ f, size_t stride,
                      const int32_t *__restrict__ b, const int32_t *__restrict__ sbox, WGrid g,
                      int n, const double *__restrict__ xyz, const double *__restrict__ nrm,
                      double mu_over_dx, double pressure_scale, double pgauge,
                      double *__restrict__ wss, double *__restrict__ press,
                      int32_t *__restrict__ valid) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double x = xyz[3 * m], y = xyz[3 * m + 1], z = xyz[3 * m + 2];
  const double nx = nrm[3 * m], ny = nrm[3 * m + 1], nz = nrm[3 * m + 2];
  int lo[3], hi[3];
  wss_window(g, x, y, z, lo, hi);
  double mind = 1.42 * g.dx;
  size_t best = ~(size_t)0;
  for (int i = lo[0]; i <= hi[0]; i++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int k = lo[2]; k <= hi[2]; k++) {
        const size_t q = wq(g, i, j, k);
        if (b[q] < 0) continue;
        const double dx = x - (g.x0 + (i - 1) * g.dx), dy = y - (g.y0 + (j - 1) * g.dx),
                     dz = z - (g.z0 + (k - 1) * g.dx);
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (d < mind) { mind = d; best = q; }
      }
  if (best != ~(size_t)0) {
    double drho, jm[3];
    node_moments<S>(f, stride, sbox[best], drho, jm);
    valid[m] = 1;
    press[m] = drho * pressure_scale - pgauge;
  } else {
    valid[m] = 0;
    press[m] = 0.0;
  }
  double u0 = 0, u1 = 0, u2 = 0;
  wss_stencil(g, b, sbox, x + g.dx * nx, y + g.dx * ny, z + g.dx * nz, [&](int sq, double w) {
    double drho, jv[3];
    node_moments<S>(f, stride, sq, drho, jv);
    const double r = 1.0 + drho;
    u0 += w * (jv[0] / r);
    u1 += w * (jv[1] / r);
    u2 += w * (jv[2] / r);
  });
  const double un = u0 * nx + u1 * ny + u2 * nz;
  wss[3 * m] = -mu_over_dx * (u0 - un * nx);
  wss[3 * m + 1] = -mu_over_dx * (u1 - un * ny);
  wss[3 * m + 2] = -mu_over_dx * (u2 - un * nz);
}

template <int S>
void launch_wss(const zf_ctx *ctx, const WGrid &g, int n, const double *d_xyz, const double *d_nrm,
                double mu_over_dx, double pgauge, double *d_wss, double *d_p, int32_t *d_valid) {
  const dim3 grid(nb((size_t)n, 64)), blk(64);
  hipLaunchKernelGGL(k_wss<S>, grid, blk, 0, ctx->stream,
                     static_cast<const typename Store<S>::T *>(ctx->d_f[ctx->cur]), ctx->stride,
                     ctx->d_b, ctx->d_sbox, g, n, d_xyz, d_nrm, mu_over_dx,
                     ctx->cfg.pressure_scale, pgauge, d_wss, d_p, d_valid);
}

void launch_wss_any(const zf_ctx *ctx, const WGrid &g, int n, const double *d_xyz,
                    const double *d_nrm, double mu_over_dx, double pgauge, double *d_wss,
                    double *d_p, int32_t *d_valid) {
  switch (ctx->cfg.storage) {
    case ZF_STORE_F64:
      launch_wss<ZF_STORE_F64>(ctx, g, n, d_xyz, d_nrm, mu_over_dx, pgauge, d_wss, d_p, d_valid);
      break;
    case ZF_STORE_F32DEV:
      launch_wss<ZF_STORE_F32DEV>(ctx, g, n, d_xyz, d_nrm, mu_over_dx, pgauge, d_wss, d_p, d_valid);
      break;
    default:
      launch_wss<ZF_STORE_F16DEV>(ctx, g, n, d_xyz, d_nrm, mu_over_dx, pgauge, d_wss, d_p, d_valid);
  }
}

WGrid wss_grid(const zf_ctx *ctx) {
  const GridGeom gp = grid_geom(ctx);
  return WGrid{gp.x0, gp.y0, gp.z0, gp.dx, ctx->cfg.im, ctx->cfg.jm, ctx->cfg.km};
}

}  // namespace

// Device address of the allocation-free scratch on ctx's device, looked up
// once per device with that device current (zf_create calls this).
int wss_scratch_init(zf_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(ctx, ZF_ERR_ARG, "device %d beyond the scratch table", ctx->device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[ctx->device]) return ZF_OK;
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(ctx, ZF_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch_ptr[ctx->device] = ptr;
  return ZF_OK;
}

GridGeom grid_geom(const zf_ctx *ctx) {
  if (ctx->have_ibm) return GridGeom{ctx->icfg.x0, ctx->icfg.y0, ctx->icfg.z0, ctx->icfg.dx};
  return GridGeom{0.0, 0.0, 0.0, 1.0};
}

}  // namespace zf

using namespace zf;

extern "C" {

int zf_wss(zf_ctx *ctx, int64_t n, const double *xyz, const double *normal, double mu_over_dx,
           double pgauge, double *wss, double *pressure, int32_t *valid) {
  if (!ctx || n < 0 || (n > 0 && (!xyz || !normal || !wss || !pressure || !valid)))
    return ZF_ERR_ARG;
  if (!ctx->have_lattice || !ctx->d_f[0]) return fail(ctx, ZF_ERR_STATE, "no lattice state");
  if (n == 0) return ZF_OK;
  if (n > (1ll << 30)) return fail(ctx, ZF_ERR_ARG, "too many wall points");
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  double *buf = nullptr;
  int32_t *vb = nullptr;
  int rc = dev_alloc(ctx, (void **)&buf, sizeof(double) * 10 * (size_t)n);
  if (rc) return rc;
  if ((rc = dev_alloc(ctx, (void **)&vb, sizeof(int32_t) * (size_t)n))) {
    dev_free(ctx, (void **)&buf);
    return rc;
  }
  double *d_xyz = buf, *d_nrm = buf + 3 * n, *d_wss = buf + 6 * n, *d_p = buf + 9 * n;
  const WGrid g = wss_grid(ctx);
  hipError_t e = hipMemcpyAsync(d_xyz, xyz, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_nrm, normal, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    launch_wss_any(ctx, g, (int)n, d_xyz, d_nrm, mu_over_dx, pgauge, d_wss, d_p, vb);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(wss, d_wss, 24 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pressure, d_p, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(valid, vb, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->device_bytes -= (int64_t)(84 * (size_t)n);
  dev_free(ctx, (void **)&buf);
  dev_free(ctx, (void **)&vb);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "wall shear stress: %s", hipGetErrorString(e));
  return ZF_OK;
}

// The allocation-free variant of zf_wss: the same kernel and results, with
// the inputs and outputs staged in the file-scope scratch instead of
// buffers allocated per call.  Calls on one device share that scratch, so
// overlapping calls on one device (from several host threads or contexts)
// race; calls on one context are ordered by its stream.  Every staged word
// is written before it is read, so nothing is cleared between calls.
int zf_wss_noalloc(zf_ctx *ctx, int64_t n, const double *xyz, const double *normal,
                   double mu_over_dx, double pgauge, double *wss, double *pressure,
                   int32_t *valid) {
  if (!ctx || n < 0 || (n > 0 && (!xyz || !normal || !wss || !pressure || !valid)))
    return ZF_ERR_ARG;
  if (!ctx->have_lattice || !ctx->d_f[0]) return fail(ctx, ZF_ERR_STATE, "no lattice state");
  if (n > ZF_WSS_NOALLOC_MAX_POINTS)
    return fail(ctx, ZF_ERR_ARG, "zf_wss_noalloc: %lld wall points exceed its limit of %d",
                (long long)n, (int)ZF_WSS_NOALLOC_MAX_POINTS);
  if (n == 0) return ZF_OK;
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  }
  if (!base) return fail(ctx, ZF_ERR_STATE, "zf_wss_noalloc: scratch of device %d not set up", ctx->device);
  ZF_HIP(ctx, hipSetDevice(ctx->device));
  // layout for this call's n: xyz, normal, wss (3n doubles each), pressure
  // (n doubles), valid (n ints)
  double *d_xyz = static_cast<double *>(base), *d_nrm = d_xyz + 3 * n, *d_wss = d_xyz + 6 * n,
         *d_p = d_xyz + 9 * n;
  int32_t *d_valid = reinterpret_cast<int32_t *>(d_xyz + 10 * n);
  hipError_t e = hipMemcpyAsync(d_xyz, xyz, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_nrm, normal, 24 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    launch_wss_any(ctx, wss_grid(ctx), (int)n, d_xyz, d_nrm, mu_over_dx, pgauge, d_wss, d_p, d_valid);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(wss, d_wss, 24 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pressure, d_p, 8 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(valid, d_valid, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "wall shear stress (no alloc): %s", hipGetErrorString(e));
  return ZF_OK;
}

}  // extern "C"
