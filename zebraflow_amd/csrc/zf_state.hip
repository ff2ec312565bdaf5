// zf_state.hip -- exact state checkpoints on the device, and the cell status
// upload a restart needs.
//
// The reference restarts from its macroscopic CSV output and re-equilibrates
// the populations (ZebraflowWT.c:5253-5262), which drops the non-equilibrium
// part; these checkpoints keep the populations as stored, the force field
// spread at the last RBC iteration, and the cells (positions, velocities,
// membrane force, repulsion velocity, status; with nuclei their positions,
// velocities, forces and contact lists; on a parent lattice the cells'
// flags, on a main lattice its clones' bands), so that a restored context
// continues bit-identically.
#include "zf_internal.h"

#include <vector>

namespace zf {

void free_state_slots(zf_ctx *ctx) {
  for (auto &s : ctx->slots) dev_free(ctx, &s.buf);
  ctx->slots.clear();
}

}  // namespace zf

using namespace zf;

namespace {

// Regions of the state, in slot order: (device pointer, bytes).
std::vector<std::pair<void *, size_t>> regions(zf_ctx *ctx, bool cells) {
  std::vector<std::pair<void *, size_t>> r;
  r.push_back({ctx->d_f[ctx->cur], store_bytes(ctx->cfg.storage) * Q * ctx->stride});
  if (ctx->d_force) r.push_back({ctx->d_force, field_elem(ctx) * 3 * ctx->stride});
  if (cells) {
    const size_t NT = (size_t)ctx->mcfg.max_cells * ctx->nn;
    r.push_back({ctx->d_x, 24 * NT});
    r.push_back({ctx->d_v, 24 * NT});
    r.push_back({ctx->d_mforce, 24 * NT});
    r.push_back({ctx->d_status, 4 * (size_t)ctx->mcfg.max_cells});
    if (ctx->d_vrep) r.push_back({ctx->d_vrep, 24 * NT});
    if (ctx->have_nuc) {  // nuclei: positions, velocities, forces, contact lists
      const size_t NTn = (size_t)ctx->mcfg.max_cells * ctx->nnuc;
      r.push_back({ctx->d_xn, 24 * NTn});
      r.push_back({ctx->d_vn, 24 * NTn});
      r.push_back({ctx->d_fn, 24 * NTn});
      r.push_back({ctx->d_nlist, 4 * NTn * (size_t)ctx->ncfg.list_cap});
      r.push_back({ctx->d_nlist_cnt, 4 * (NTn + 1)});
      r.push_back({ctx->d_nuc_stale, 4 * (size_t)ctx->mcfg.max_cells});
    }
    if (ctx->have_wrap) {  // a parent's cells: REGRESSFLAG and CLONEFLAG
      r.push_back({ctx->d_regress, 4 * (size_t)ctx->mcfg.max_cells});
      r.push_back({ctx->d_cloneflag, 4 * (size_t)ctx->mcfg.max_cells});
    }
    if (ctx->d_band) r.push_back({ctx->d_band, 16 * (size_t)ctx->mcfg.max_cells});  // clone bands
  }
  return r;
}

}  // namespace

extern "C" {

int zf_checkpoint_save(zf_ctx *ctx, int32_t slot) {
  if (!ctx || slot < 0 || slot > 1023) return ZF_ERR_ARG;
  if (!ctx->have_lattice || !ctx->d_f[0]) return fail(ctx, ZF_ERR_STATE, "no lattice uploaded");
  auto &sl = ctx->slots;
  if ((size_t)slot >= sl.size()) sl.resize(slot + 1);
  StateSlot &s = sl[slot];
  const bool cells = ctx->have_mesh;
  const auto reg = regions(ctx, cells);
  size_t bytes = 0;
  for (auto &r : reg) bytes += (r.second + 255) & ~size_t(255);
  if (s.bytes != bytes) {
    dev_free(ctx, &s.buf);
    s.bytes = 0;
    int rc = dev_alloc(ctx, &s.buf, bytes);
    if (rc) return rc;
    s.bytes = bytes;
  }
  char *p = static_cast<char *>(s.buf);
  for (auto &r : reg) {
    ZF_HIP(ctx, hipMemcpyAsync(p, r.first, r.second, hipMemcpyDeviceToDevice, ctx->stream));
    p += (r.second + 255) & ~size_t(255);
  }
  s.force = ctx->d_force != nullptr;
  s.ncell = cells ? ctx->ncell : -1;
  s.nuc = cells && ctx->have_nuc;
  s.nuc_lists_valid = ctx->nuc_lists_valid;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ZF_OK;
}

int zf_checkpoint_restore(zf_ctx *ctx, int32_t slot) {
  if (!ctx || slot < 0) return ZF_ERR_ARG;
  auto &sl = ctx->slots;
  if ((size_t)slot >= sl.size() || !sl[slot].buf) return fail(ctx, ZF_ERR_STATE, "slot %d is empty", slot);
  StateSlot &s = sl[slot];
  if (s.force != (ctx->d_force != nullptr) || (s.ncell >= 0) != ctx->have_mesh ||
      (s.ncell >= 0 && s.nuc != ctx->have_nuc))
    return fail(ctx, ZF_ERR_STATE, "slot %d holds a state of another configuration", slot);
  if (s.ncell >= 0) ctx->ncell = s.ncell;
  const auto reg = regions(ctx, s.ncell >= 0);
  size_t bytes = 0;
  for (auto &r : reg) bytes += (r.second + 255) & ~size_t(255);
  if (bytes != s.bytes) return fail(ctx, ZF_ERR_STATE, "slot %d holds a state of another size", slot);
  const char *p = static_cast<const char *>(s.buf);
  for (auto &r : reg) {
    ZF_HIP(ctx, hipMemcpyAsync(r.first, p, r.second, hipMemcpyDeviceToDevice, ctx->stream));
    p += (r.second + 255) & ~size_t(255);
  }
  ctx->geom_gen++;  // positions changed: the cell list's bounds expire
  if (s.nuc) ctx->nuc_lists_valid = s.nuc_lists_valid;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ZF_OK;
}

int zf_upload_cell_status(zf_ctx *ctx, const int32_t *status) {
  if (!ctx || !status) return ZF_ERR_ARG;
  if (!ctx->have_mesh) return fail(ctx, ZF_ERR_STATE, "no membrane mesh");
  if (ctx->ncell == 0) return ZF_OK;
  ZF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ZF_HIP(ctx, hipMemcpy(ctx->d_status, status, 4 * (size_t)ctx->ncell, hipMemcpyHostToDevice));
  return ZF_OK;
}

}  // extern "C"
