// zf_core.hip -- context lifetime, configuration defaults, errors, events.
#include "zf_internal.h"

#include <cstring>

namespace zf {

int fail(zf_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

int dev_alloc(zf_ctx *ctx, void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(ctx, ZF_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  ctx->device_bytes += (int64_t)bytes;
  return ZF_OK;
}

void dev_free(zf_ctx *ctx, void **p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

}  // namespace zf

using namespace zf;

extern "C" {

const char *zf_version(void) { return "zebraflow-mi355x 0.1 (gfx950)"; }

// Reference constants: DX = DT = 0.5 (ZebraflowWT.c:40-41), MU_SCALE*MU =
// 3*T_SCALE/L_SCALE^2 * 1.2e-7 = 0.72 (:47-49, :65-66), LBM_PRESSURE_SCALE =
// CS^2*DEN_SCALE*DENphys = 10000/3 (:69).
void zf_config_reference(zf_config *cfg, int32_t im, int32_t jm, int32_t km) {
  std::memset(cfg, 0, sizeof *cfg);
  cfg->im = im; cfg->jm = jm; cfg->km = km;
  cfg->storage = ZF_STORE_F32DEV;
  cfg->wall_mode = ZF_WALL_EQUILIBRIUM;
  const double t_scale = 0.5e-6, l_scale = 0.5e-6;
  const double mu = 0.0012 / 1000.0 / 10.0;
  cfg->visc_scale = 3.0 * t_scale / l_scale / l_scale * mu;
  cfg->force_dt = 0.5;
  cfg->inlet_period = 1.0;
  cfg->inlet_dir[0] = 1.0;
  cfg->outlet_rho = 1.0;
  cfg->t_scale = t_scale;
  cfg->pressure_scale = (1.0 / 3.0) * 10.0 * 1000.0;
}

void zf_config_lattice(zf_config *cfg, int32_t im, int32_t jm, int32_t km, double tau) {
  zf_config_reference(cfg, im, jm, km);
  cfg->visc_scale = tau - 0.5;
  cfg->force_dt = 1.0;
  cfg->wall_mode = ZF_WALL_HALFWAY_BB;
}

int zf_create(const zf_config *cfg, zf_ctx **out) {
  if (!cfg || !out) return ZF_ERR_ARG;
  *out = nullptr;
  if (cfg->im <= 0 || cfg->jm <= 0 || cfg->km <= 0) return ZF_ERR_ARG;
  if (cfg->storage < ZF_STORE_F64 || cfg->storage > ZF_STORE_F16DEV) return ZF_ERR_ARG;
  if (cfg->wall_mode != ZF_WALL_EQUILIBRIUM && cfg->wall_mode != ZF_WALL_HALFWAY_BB)
    return ZF_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ZF_ERR_HIP;
  if (cfg->device < 0 || cfg->device >= ndev) return ZF_ERR_ARG;
  zf_ctx *ctx = new zf_ctx();
  ctx->cfg = *cfg;
  ctx->device = cfg->device;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return ZF_ERR_HIP;
  }
  for (auto &e : ctx->events) {
    if (hipEventCreate(&e) != hipSuccess) {
      zf_destroy(ctx);
      return ZF_ERR_HIP;
    }
  }
  if (wss_scratch_init(ctx) != ZF_OK) {
    zf_destroy(ctx);
    return ZF_ERR_HIP;
  }
  *out = ctx;
  return ZF_OK;
}

int zf_destroy(zf_ctx *ctx) {
  if (!ctx) return ZF_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_parent(ctx);
  free_adjoint(ctx);
  free_state_slots(ctx);
  free_cells(ctx);
  free_nucleus(ctx);
  free_ibm(ctx);
  free_membrane(ctx);
  free_state(ctx);
  free_lattice(ctx);
  for (auto &e : ctx->events)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ZF_OK;
}

const char *zf_last_error(const zf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int zf_get_config(const zf_ctx *ctx, zf_config *cfg) {
  if (!ctx || !cfg) return ZF_ERR_ARG;
  *cfg = ctx->cfg;
  return ZF_OK;
}

int zf_sync(zf_ctx *ctx) {
  if (!ctx) return ZF_ERR_ARG;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return pipe_check(ctx);
}

}  // extern "C"

namespace {
// Plain copy, one 16-byte word per thread (the fastest of the simple copy
// shapes on MI355X at sizes beyond the last-level cache).
__global__ __launch_bounds__(256) void k_copy16(const float4 *__restrict__ a, float4 *__restrict__ b,
                                                size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}
}  // namespace

extern "C" {

int zf_copy_bandwidth(zf_ctx *ctx, int64_t bytes, int32_t reps, double *gbps) {
  if (!ctx || !gbps || bytes < 16 || reps < 1) return ZF_ERR_ARG;
  void *a = nullptr, *b = nullptr;
  int rc = dev_alloc(ctx, &a, (size_t)bytes);
  if (rc) return rc;
  if ((rc = dev_alloc(ctx, &b, (size_t)bytes))) {
    dev_free(ctx, &a);
    return rc;
  }
  const size_t n = (size_t)bytes / 16;
  const dim3 grid((unsigned)((n + 255) / 256));
  hipError_t e = hipMemsetAsync(a, 0, (size_t)bytes, ctx->stream);
  // warm-up, then the timed copies
  hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, ctx->stream, (const float4 *)a, (float4 *)b, n);
  if (e == hipSuccess) e = hipEventRecord(ctx->events[14], ctx->stream);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, ctx->stream, (const float4 *)a, (float4 *)b, n);
  if (e == hipSuccess) e = hipEventRecord(ctx->events[15], ctx->stream);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipEventSynchronize(ctx->events[15]);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->events[14], ctx->events[15]);
  ctx->device_bytes -= 2 * bytes;
  dev_free(ctx, &a);
  dev_free(ctx, &b);
  if (e != hipSuccess) return fail(ctx, ZF_ERR_HIP, "copy bandwidth: %s", hipGetErrorString(e));
  *gbps = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
  return ZF_OK;
}

int zf_event_record(zf_ctx *ctx, int32_t slot) {
  if (!ctx || slot < 0 || slot >= 16) return ZF_ERR_ARG;
  ZF_HIP(ctx, hipEventRecord(ctx->events[slot], ctx->stream));
  return ZF_OK;
}

int zf_event_elapsed(zf_ctx *ctx, int32_t a, int32_t b, float *ms) {
  if (!ctx || !ms || a < 0 || a >= 16 || b < 0 || b >= 16) return ZF_ERR_ARG;
  ZF_HIP(ctx, hipEventSynchronize(ctx->events[b]));
  ZF_HIP(ctx, hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
  return ZF_OK;
}

}  // extern "C"
