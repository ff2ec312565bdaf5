/*
 * zf_loop.c -- the reference's time loop over the engine's C ABI.
 *
 * ZebraflowWT.c:5646-12412 advance one iteration at a time (LBM step every
 * iteration; membrane, IBM and repulsion every SOLVERSTAGGER-th, in the
 * order of zf_step) and call OUTPUT() every outputfreq iterations
 * (:12408), which writes the fluid domain, the cells and the volume/area
 * history (:282-446); optionally an exact restart file every
 * checkpoint_every iterations.  Parent domains and nuclei are not part of
 * this engine.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zf_host.h"

static int output_fluid(zf_ctx *ctx, const int32_t *b, const zfh_run_config *rc, int iter) {
  zf_lattice_info info;
  int err = zf_get_lattice_info(ctx, &info);
  if (err) return err;
  const int64_t n = info.dommax;
  const size_t box = (size_t)rc->im * rc->jm * rc->km;
  double *m = malloc(sizeof(double) * 7 * (size_t)(n > 0 ? n : 1));
  int32_t *eul = calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
  int32_t *windex = rc->have_wall ? malloc(sizeof(int32_t) * box) : NULL;
  int32_t *bins = rc->ncell > 0 && rc->have_rep ? malloc(sizeof(int32_t) * (size_t)rc->ncell * rc->nt) : NULL;
  if (!m || !eul || (rc->have_wall && !windex) || (rc->ncell > 0 && rc->have_rep && !bins)) {
    err = ZF_ERR_NOMEM;
    goto done;
  }
  double *d = m, *u = m + n, *v = m + 2 * n, *w = m + 3 * n, *force = m + 4 * n;
  if ((err = zf_download_macro(ctx, d, u, v, w))) goto done;
  if (rc->ncell > 0 && (err = zf_download_force(ctx, force))) goto done;
  if (bins) {
    /* EulNeigh[0][o]: the number of triangles in the bin of node o */
    if ((err = zf_download_cell_list(ctx, bins))) goto done;
    for (int64_t q = 0; q < (int64_t)rc->ncell * rc->nt; q++)
      if (bins[q] > 0 && bins[q] <= n) eul[bins[q] - 1]++;
  }
  if (windex && (err = zf_download_windex(ctx, windex))) goto done;
  char path[1024];
  snprintf(path, sizeof path, "%s/FLUIDdomain-%09d.csv", rc->out_dir, iter);
  if (zfh_write_fluid_csv(path, b, rc->im, rc->jm, rc->km, rc->x0, rc->y0, rc->z0, rc->dx, d, u, v,
                          w, rc->ncell > 0 ? force : NULL, eul, windex, n))
    err = ZF_ERR_ARG;
done:
  free(m);
  free(eul);
  free(windex);
  free(bins);
  return err;
}

static int output_cells(zf_ctx *ctx, const zfh_run_config *rc, int iter) {
  const size_t nq = (size_t)rc->ncell * rc->nn;
  double *buf = malloc(sizeof(double) * (12 * nq + 5 * (size_t)rc->ncell));
  int32_t *ib = malloc(sizeof(int32_t) * (nq + 2 * (size_t)rc->ncell));
  int err = 0;
  if (!buf || !ib) {
    err = ZF_ERR_NOMEM;
    goto done;
  }
  double *xyz = buf, *vel = buf + 3 * nq, *force = buf + 6 * nq, *vrep = buf + 9 * nq;
  double *area = buf + 12 * nq, *vol = area + rc->ncell, *cen = vol + rc->ncell;
  int32_t *wall = ib, *status = ib + nq, *fail = status + rc->ncell;
  if ((err = zf_download_cells(ctx, xyz, vel, status))) goto done;
  if ((err = zf_download_membrane(ctx, force, NULL, area, vol, cen, fail))) goto done;
  if ((err = zf_download_repulsion(ctx, vrep))) goto done;
  if (rc->have_rep) {
    if ((err = zf_download_wall_neighbour(ctx, wall))) goto done;
  } else {
    memset(wall, 0, sizeof(int32_t) * nq);
  }
  char path[1024];
  snprintf(path, sizeof path, "%s/AllCellMemb-%09d.csv", rc->out_dir, iter);
  if (zfh_write_cells_csv(path, rc->ncell, rc->nn, xyz, vel, status, wall, vrep, force)) {
    err = ZF_ERR_ARG;
    goto done;
  }
  snprintf(path, sizeof path, "%s/VolumeAreaHistory.dat", rc->out_dir);
  if (zfh_append_volume_area(path, iter, rc->ncell, vol, area, cen, status)) {
    err = ZF_ERR_ARG;
    goto done;
  }
  if (rc->nnuc > 0) {
    const size_t nqn = (size_t)rc->ncell * rc->nnuc;
    double *nb = malloc(sizeof(double) * 9 * (nqn ? nqn : 1));
    if (!nb) {
      err = ZF_ERR_NOMEM;
      goto done;
    }
    if (!(err = zf_download_nuclei(ctx, nb, nb + 3 * nqn, nb + 6 * nqn))) {
      snprintf(path, sizeof path, "%s/AllCellNucleus-%09d.csv", rc->out_dir, iter);
      if (zfh_write_nuclei_csv(path, rc->ncell, rc->nnuc, nb, nb + 3 * nqn, nb + 6 * nqn, status))
        err = ZF_ERR_ARG;
    }
    free(nb);
  }
done:
  free(buf);
  free(ib);
  return err;
}

int zfh_run(zf_ctx *ctx, const int32_t *b, const zfh_run_config *rc) {
  if (!ctx || !b || !rc || rc->steps < 0) return ZF_ERR_ARG;
  for (int iter = rc->iter0 + 1; iter <= rc->iter0 + rc->steps; iter++) {
    int err = zf_step(ctx, iter);
    if (err) return err;
    if (rc->output_every > 0 && iter % rc->output_every == 0) {
      if ((err = output_fluid(ctx, b, rc, iter))) return err;
      if (rc->ncell > 0 && (err = output_cells(ctx, rc, iter))) return err;
    }
    if (rc->checkpoint_every > 0 && iter % rc->checkpoint_every == 0) {
      char path[1024];
      snprintf(path, sizeof path, "%s/ZFstate-%09d.bin", rc->out_dir, iter);
      if ((err = zfh_save_state(ctx, path, iter, rc->ncell, rc->nn))) return err;
    }
  }
  return zf_sync(ctx);
}
