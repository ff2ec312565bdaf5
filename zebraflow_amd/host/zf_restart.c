/*
 * zf_restart.c -- exact restart files.
 *
 * The reference restarts from its FLUIDdomain CSV (rho, u) and rebuilds the
 * populations at equilibrium (ZebraflowWT.c:5253-5262), which drops their
 * non-equilibrium part.  These files hold the state as the engine stores
 * it: the populations as stored words (zf_download_populations_raw: double
 * f, float f - w or half (f - w)/w, so no deviation is rounded on the way),
 * the force field spread at the last RBC iteration (stored in FP32 or FP64,
 * written as double, which holds either exactly), the cell positions and
 * status.  Node velocities, membrane forces and repulsion velocities are
 * not stored: every RBC iteration recomputes them before they are used.
 *
 * Layout (native byte order): "ZFSTATE2", then int32 iter, storage, dommax,
 * ncell, nn, has_force; populations[19][dommax] of the storage's element
 * (8, 4 or 2 bytes); if has_force, double force[3][dommax]; if ncell > 0,
 * double xyz[ncell][nn][3] and int32 status[ncell]; then, for nucleated
 * cells only, int32 nnuc, list_cap, lists_valid, double
 * xyz[ncell][nnuc][3], int32 list[ncell*nnuc][list_cap] and
 * count[ncell*nnuc] (the contact lists, so that the restarted run keeps
 * the continuous one's lists until their next rebuild).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zf_host.h"

static const char kMagic[8] = {'Z', 'F', 'S', 'T', 'A', 'T', 'E', '2'};

static size_t elem_bytes(int storage) {
  return storage == ZF_STORE_F64 ? 8 : storage == ZF_STORE_F32DEV ? 4 : 2;
}

int zfh_save_state(zf_ctx *ctx, const char *path, int32_t iter, int32_t ncell, int32_t nn) {
  if (!ctx || !path || ncell < 0 || (ncell > 0 && nn <= 0)) return ZF_ERR_ARG;
  zf_lattice_info info;
  int err = zf_get_lattice_info(ctx, &info);
  if (err) return err;
  zf_config cfg;
  if ((err = zf_get_config(ctx, &cfg))) return err;
  const size_t n = (size_t)info.dommax, nq = (size_t)ncell * nn, eb = elem_bytes(cfg.storage);
  void *f = malloc(eb * 19 * n + 1);
  double *force = malloc(sizeof(double) * (3 * n + 3 * nq + 1));
  int32_t *st = malloc(sizeof(int32_t) * ((size_t)ncell + 1));
  FILE *fp = NULL;
  double *nbuf = NULL;
  int32_t *lbuf = NULL;
  if (!f || !force || !st) {
    err = ZF_ERR_NOMEM;
    goto done;
  }
  double *xyz = force + 3 * n;
  if ((err = zf_download_populations_raw(ctx, f, (int64_t)(eb * 19 * n))) ||
      (err = zf_download_force(ctx, force)))
    goto done;
  int32_t has_force = 0;
  for (size_t i = 0; i < 3 * n && !has_force; i++) has_force = force[i] != 0.0;
  if (ncell > 0 && (err = zf_download_cells(ctx, xyz, NULL, st))) goto done;
  int32_t nuc[3] = {0, 0, 0};
  if ((err = zf_get_nucleus_info(ctx, &nuc[0], &nuc[1]))) goto done;
  const size_t nqn = ncell > 0 ? (size_t)ncell * nuc[0] : 0;
  if (nqn) {
    nbuf = malloc(sizeof(double) * 3 * nqn);
    lbuf = malloc(sizeof(int32_t) * nqn * ((size_t)nuc[1] + 1));
    if (!nbuf || !lbuf) {
      err = ZF_ERR_NOMEM;
      goto done;
    }
    if ((err = zf_download_nuclei(ctx, nbuf, NULL, NULL)) ||
        (err = zf_download_nucleus_lists(ctx, lbuf, lbuf + nqn * nuc[1], &nuc[2])))
      goto done;
  }
  fp = fopen(path, "wb");
  if (!fp) {
    err = ZF_ERR_ARG;
    goto done;
  }
  const int32_t hdr[6] = {iter, cfg.storage, (int32_t)n, ncell, nn, has_force};
  if (fwrite(kMagic, 1, 8, fp) != 8 || fwrite(hdr, sizeof hdr, 1, fp) != 1 ||
      fwrite(f, eb, 19 * n, fp) != 19 * n ||
      (has_force && fwrite(force, sizeof(double), 3 * n, fp) != 3 * n) ||
      (ncell > 0 && (fwrite(xyz, sizeof(double), 3 * nq, fp) != 3 * nq ||
                     fwrite(st, sizeof(int32_t), (size_t)ncell, fp) != (size_t)ncell)) ||
      (nqn && (fwrite(nuc, sizeof nuc, 1, fp) != 1 || fwrite(nbuf, sizeof(double), 3 * nqn, fp) != 3 * nqn ||
               fwrite(lbuf, sizeof(int32_t), nqn * (nuc[1] + 1), fp) != nqn * (nuc[1] + 1))))
    err = ZF_ERR_ARG;
done:
  if (fp && fclose(fp) && !err) err = ZF_ERR_ARG;
  free(f);
  free(force);
  free(st);
  free(nbuf);
  free(lbuf);
  return err;
}

int zfh_load_state(zf_ctx *ctx, const char *path, int32_t *iter, int32_t ncell, int32_t nn) {
  if (!ctx || !path || !iter) return ZF_ERR_ARG;
  FILE *fp = fopen(path, "rb");
  if (!fp) return ZF_ERR_ARG;
  char magic[8];
  int32_t hdr[6];
  void *f = NULL;
  double *force = NULL;
  int32_t *st = NULL, *lbuf = NULL;
  double *nbuf = NULL;
  int err = 0;
  zf_lattice_info info;
  zf_config cfg;
  if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, kMagic, 8) || fread(hdr, sizeof hdr, 1, fp) != 1) {
    err = ZF_ERR_ARG;
    goto done;
  }
  if ((err = zf_get_lattice_info(ctx, &info)) || (err = zf_get_config(ctx, &cfg))) goto done;
  /* the file must come from the same lattice, storage form and cell set-up */
  if (hdr[1] != cfg.storage || hdr[2] != info.dommax || hdr[3] != ncell || (ncell > 0 && hdr[4] != nn)) {
    err = ZF_ERR_ARG;
    goto done;
  }
  const size_t n = (size_t)hdr[2], nq = (size_t)ncell * nn, eb = elem_bytes(cfg.storage);
  f = malloc(eb * 19 * n + 1);
  force = malloc(sizeof(double) * (3 * n + 3 * nq + 1));
  st = malloc(sizeof(int32_t) * ((size_t)ncell + 1));
  if (!f || !force || !st) {
    err = ZF_ERR_NOMEM;
    goto done;
  }
  double *xyz = force + 3 * n;
  if (fread(f, eb, 19 * n, fp) != 19 * n ||
      (hdr[5] && fread(force, sizeof(double), 3 * n, fp) != 3 * n) ||
      (ncell > 0 && (fread(xyz, sizeof(double), 3 * nq, fp) != 3 * nq ||
                     fread(st, sizeof(int32_t), (size_t)ncell, fp) != (size_t)ncell))) {
    err = ZF_ERR_ARG;
    goto done;
  }
  if ((err = zf_upload_populations_raw(ctx, f, (int64_t)(eb * 19 * n)))) goto done;
  /* an all-zero field (no RBC iteration yet) is not stored and not uploaded */
  if (hdr[5] && (err = zf_upload_force(ctx, force))) goto done;
  if (ncell > 0 && ((err = zf_upload_cells(ctx, ncell, xyz)) || (err = zf_upload_cell_status(ctx, st))))
    goto done;
  /* nuclei: present in the file exactly when the context has them */
  int32_t nnuc = 0, cap = 0;
  if ((err = zf_get_nucleus_info(ctx, &nnuc, &cap))) goto done;
  const size_t nqn = ncell > 0 ? (size_t)ncell * nnuc : 0;
  if (nqn) {
    int32_t nuc[3];
    nbuf = malloc(sizeof(double) * 3 * nqn);
    lbuf = malloc(sizeof(int32_t) * nqn * ((size_t)cap + 1));
    if (!nbuf || !lbuf) {
      err = ZF_ERR_NOMEM;
      goto done;
    }
    if (fread(nuc, sizeof nuc, 1, fp) != 1 || nuc[0] != nnuc || nuc[1] != cap ||
        fread(nbuf, sizeof(double), 3 * nqn, fp) != 3 * nqn ||
        fread(lbuf, sizeof(int32_t), nqn * (cap + 1), fp) != nqn * (cap + 1)) {
      err = ZF_ERR_ARG;
      goto done;
    }
    if ((err = zf_upload_nuclei(ctx, nbuf)) ||
        (err = zf_upload_nucleus_lists(ctx, lbuf, lbuf + nqn * cap, nuc[2])))
      goto done;
  }
  char extra;
  if (fread(&extra, 1, 1, fp) == 1) { /* e.g. nuclei that the context lacks */
    err = ZF_ERR_ARG;
    goto done;
  }
  *iter = hdr[0];
done:
  fclose(fp);
  free(f);
  free(force);
  free(st);
  free(nbuf);
  free(lbuf);
  return err;
}
