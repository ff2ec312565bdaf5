/*
 * zf_host.c -- the C host program: reads a case in the reference's input
 * formats, runs the reference's step/output loop on one MI355X through
 * libzfgpu.so, and writes the reference's output files.
 *
 *   zf_host case.cfg
 *
 * The case file holds "key = value" lines (# comments):
 *   box = IM JM KM                 lattice box (ZebraflowWT.c:81-83)
 *   grid = X0 Y0 Z0 DX             node (i,j,k) at X0 + (i-1) DX, ... (:2508-2525)
 *   bindex = FLUIDdomainBindex.csv the B field (:2574-2619)
 *   units = lattice | reference    lattice: tau, force_dt 1, bounce-back walls;
 *                                  reference: the reference's constants (:34-74)
 *   tau = 0.8                      relaxation time of B = 1 nodes (lattice units)
 *   storage = f32dev | f64 | f16dev
 *   wall_mode = bounceback | equilibrium
 *   periodic = PI PJ PK
 *   body_force = GX GY GZ
 *   inlet = U0 AMP PERIOD DX DY DZ  velocity inlet u0 (1 + amp sin(2 pi t/period)) dir
 *   outlet_rho = 1.0
 *   rbc_nodes = ZFISHRBCNODES.dat  rbc_triangles = ZFISHRBCTRIANGLES.dat (:1601-1626)
 *   nucleus_nodes = NucleusNodes.dat  nucleus_edges = NucleusEdges.dat (:1903-1922)
 *   nuclei = AllCellNucStart.csv   the start cells' nuclei (:5530-5585); with
 *                                  these three the cells are nucleated and the
 *                                  loop also writes AllCellNucleus-%09d.csv
 *   cells = AllCellMembStart.csv   (AllCellMemb layout, :5327-5430)
 *   max_cells = N                  shear_modulus = ES   bending_modulus = KB
 *   rbc_interval = 50              gamma_p = G          repulsion = 0 | 1
 *   wall_nodes / wall_cells / wall_normals = WallMesh*.dat, or wall = surface
 *   steps = N   output_every = M   output_dir = SimulationOutputData
 *   checkpoint_every = C           exact restart file ZFstate-%09d.bin every C
 *   restart = ZFstate-000000100.bin   continue from such a file (steps more)
 *   uvoid = reference | node | off  wall-layer velocity Uvoid (:4788-4851; default
 *                                  reference with units = reference, else off)
 *   uvoid_speed = S                its speed (1e-3, units of U)
 *   fluid_start = FLUIDdomainStart.aa FLUIDdomainStart.ab ...
 *                                  the reference's start state (:2626-3387),
 *                                  populations at equilibrium (:5253-5262)
 * Outputs: FLUIDdomain-%09d.csv, AllCellMemb-%09d.csv, VolumeAreaHistory.dat
 * (:282-446).
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "zf_host.h"

#define MAXKEYS 64

typedef struct {
  char key[64];
  char val[960];
} kv_t;

static kv_t g_kv[MAXKEYS];
static int g_nkv = 0;

static const char *get(const char *key) {
  for (int i = 0; i < g_nkv; i++)
    if (strcmp(g_kv[i].key, key) == 0) return g_kv[i].val;
  return NULL;
}

static int read_case(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    char *h = strchr(line, '#');
    if (h) *h = 0;
    char *eq = strchr(line, '=');
    if (!eq) continue;
    *eq = 0;
    char k[64], v[960];
    if (sscanf(line, " %63s", k) != 1) continue;
    char *val = eq + 1;
    while (*val == ' ' || *val == '\t') val++;
    size_t n = strlen(val);
    while (n > 0 && (val[n - 1] == '\n' || val[n - 1] == '\r' || val[n - 1] == ' ')) val[--n] = 0;
    snprintf(v, sizeof v, "%s", val);
    if (g_nkv == MAXKEYS) { fclose(f); return -1; }
    snprintf(g_kv[g_nkv].key, sizeof g_kv[g_nkv].key, "%s", k);
    snprintf(g_kv[g_nkv].val, sizeof g_kv[g_nkv].val, "%s", v);
    g_nkv++;
  }
  fclose(f);
  return 0;
}

static double getd(const char *key, double def) {
  const char *v = get(key);
  return v ? atof(v) : def;
}

static int geti(const char *key, int def) {
  const char *v = get(key);
  return v ? atoi(v) : def;
}

static int die(zf_ctx *ctx, int rc, const char *what) {
  fprintf(stderr, "zf_host: %s failed (%d): %s\n", what, rc, ctx ? zf_last_error(ctx) : "");
  if (ctx) zf_destroy(ctx);
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: zf_host case.cfg\n");
    return 2;
  }
  if (read_case(argv[1])) {
    fprintf(stderr, "zf_host: cannot read %s\n", argv[1]);
    return 2;
  }
  int im = 0, jm = 0, km = 0;
  if (!get("box") || sscanf(get("box"), "%d %d %d", &im, &jm, &km) != 3 || im <= 0 || jm <= 0 || km <= 0) {
    fprintf(stderr, "zf_host: box = IM JM KM is required\n");
    return 2;
  }
  double x0 = 0, y0 = 0, z0 = 0, dx = 1;
  if (get("grid") && sscanf(get("grid"), "%lf %lf %lf %lf", &x0, &y0, &z0, &dx) != 4) {
    fprintf(stderr, "zf_host: grid = X0 Y0 Z0 DX\n");
    return 2;
  }
  const int reference = get("units") && strcmp(get("units"), "reference") == 0;
  zf_config cfg;
  if (reference) zf_config_reference(&cfg, im, jm, km);
  else zf_config_lattice(&cfg, im, jm, km, getd("tau", 0.8));
  const char *st = get("storage");
  if (st) cfg.storage = strcmp(st, "f64") == 0 ? ZF_STORE_F64 : strcmp(st, "f16dev") == 0 ? ZF_STORE_F16DEV : ZF_STORE_F32DEV;
  const char *wm = get("wall_mode");
  if (wm) cfg.wall_mode = strcmp(wm, "equilibrium") == 0 ? ZF_WALL_EQUILIBRIUM : ZF_WALL_HALFWAY_BB;
  if (get("periodic")) sscanf(get("periodic"), "%d %d %d", &cfg.periodic[0], &cfg.periodic[1], &cfg.periodic[2]);
  if (get("body_force"))
    sscanf(get("body_force"), "%lf %lf %lf", &cfg.body_force[0], &cfg.body_force[1], &cfg.body_force[2]);
  if (get("inlet"))
    sscanf(get("inlet"), "%lf %lf %lf %lf %lf %lf", &cfg.inlet_u0, &cfg.inlet_amp, &cfg.inlet_period,
           &cfg.inlet_dir[0], &cfg.inlet_dir[1], &cfg.inlet_dir[2]);
  cfg.outlet_rho = getd("outlet_rho", cfg.outlet_rho);

  /* lattice */
  const char *bpath = get("bindex");
  if (!bpath) {
    fprintf(stderr, "zf_host: bindex = FLUIDdomainBindex.csv is required\n");
    return 2;
  }
  const size_t box = (size_t)im * jm * km;
  int32_t *b = malloc(sizeof(int32_t) * box);
  if (!b) return 1;
  const int64_t rows = zfh_read_bindex_csv(bpath, b, im, jm, km, x0, y0, z0, dx);
  if (rows < 0) {
    fprintf(stderr, "zf_host: cannot read %s\n", bpath);
    return 1;
  }
  zf_ctx *ctx = NULL;
  int rc = zf_create(&cfg, &ctx);
  if (rc) return die(NULL, rc, "zf_create");
  int32_t dommax = 0, bcmax = 0, nint = 0;
  if ((rc = zf_upload_lattice(ctx, b, &dommax, &bcmax, &nint))) return die(ctx, rc, "zf_upload_lattice");
  printf("zf_host: %s, %lld rows, %d listed nodes, %d fluid, %d boundary\n", zf_version(),
         (long long)rows, dommax, nint, bcmax);

  /* fluid start state */
  if (get("fluid_start")) {
    char buf[960], *paths[64];
    snprintf(buf, sizeof buf, "%s", get("fluid_start"));
    int np = 0;
    for (char *t = strtok(buf, " \t"); t && np < 64; t = strtok(NULL, " \t")) paths[np++] = t;
    const size_t n = (size_t)dommax;
    double *mac = calloc(7 * (n ? n : 1), sizeof(double));
    if (!mac) return die(ctx, ZF_ERR_NOMEM, "fluid start state");
    for (size_t p = 0; p < n; p++) mac[p] = 1.0;  /* D = 1, U = 0 where no row sets them */
    const int64_t got = zfh_read_fluid_start(np, (const char *const *)paths, b, im, jm, km, x0, y0,
                                             z0, dx, dommax, mac, mac + n, mac + 2 * n, mac + 3 * n,
                                             mac + 4 * n);
    if (got < 0) return die(ctx, ZF_ERR_ARG, "reading fluid_start");
    if ((rc = zfh_init_from_macro(ctx, mac, mac + n, mac + 2 * n, mac + 3 * n, mac + 4 * n)))
      return die(ctx, rc, "fluid start state");
    free(mac);
    printf("zf_host: fluid start state from %d file(s), %lld rows\n", np, (long long)got);
  }

  /* cells */
  zfh_run_config run;
  memset(&run, 0, sizeof run);
  const char *nodes = get("rbc_nodes"), *tris = get("rbc_triangles"), *cells = get("cells");
  double *rest = NULL, *xyz = NULL;
  int32_t *tri = NULL, *status = NULL;
  if (nodes && tris && cells) {
    const int64_t nn = zfh_read_points(nodes, NULL, 0), nt = zfh_read_rbc_triangles(tris, NULL, 0);
    if (nn <= 0 || nt <= 0) return die(ctx, ZF_ERR_ARG, "reading the RBC mesh");
    rest = malloc(sizeof(double) * 3 * nn);
    tri = malloc(sizeof(int32_t) * 3 * nt);
    if (zfh_read_points(nodes, rest, nn) != nn || zfh_read_rbc_triangles(tris, tri, nt) != nt)
      return die(ctx, ZF_ERR_ARG, "reading the RBC mesh");
    const int cap = geti("max_cells", 1);
    xyz = malloc(sizeof(double) * 3 * nn * (size_t)cap);
    status = malloc(sizeof(int32_t) * (size_t)cap);
    const int ncell = zfh_read_cells_csv(cells, (int)nn, xyz, status, cap);
    if (ncell < 0) return die(ctx, ZF_ERR_ARG, "reading the cells (max_cells too small?)");
    zf_membrane_config mc;
    if (reference) {
      zf_membrane_config_reference(&mc);
      mc.max_cells = cap;
    } else {
      zfh_membrane_config_lattice(rest, (int)nn, tri, (int)nt, getd("shear_modulus", 5e-3),
                                  getd("bending_modulus", 2e-3), cap, &mc);
    }
    if ((rc = zf_membrane_setup(ctx, &mc, (int32_t)nn, (int32_t)nt, tri, rest)))
      return die(ctx, rc, "zf_membrane_setup");
    if ((rc = zf_upload_cells(ctx, ncell, xyz))) return die(ctx, rc, "zf_upload_cells");
    if (get("nucleus_nodes") && get("nucleus_edges") && get("nuclei")) {
      const int64_t nnuc = zfh_read_points(get("nucleus_nodes"), NULL, 0);
      const int64_t nen = zfh_read_lines(get("nucleus_edges"), NULL, 0);
      if (nnuc <= 0 || nen <= 0) return die(ctx, ZF_ERR_ARG, "reading the nucleus network");
      double *nrest = malloc(sizeof(double) * 3 * nnuc);
      int32_t *nedge = malloc(sizeof(int32_t) * 2 * nen);
      double *nxyz = malloc(sizeof(double) * 3 * nnuc * (size_t)cap);
      if (!nrest || !nedge || !nxyz || zfh_read_points(get("nucleus_nodes"), nrest, nnuc) != nnuc ||
          zfh_read_lines(get("nucleus_edges"), nedge, nen) != nen)
        return die(ctx, ZF_ERR_ARG, "reading the nucleus network");
      if (zfh_read_nuclei_csv(get("nuclei"), (int)nnuc, nxyz, cap) != ncell)
        return die(ctx, ZF_ERR_ARG, "reading the nuclei (one per start cell)");
      zf_nucleus_config nc;
      zf_nucleus_config_reference(&nc);
      if ((rc = zf_nucleus_setup(ctx, &nc, (int32_t)nnuc, (int32_t)nen, nedge, nrest)))
        return die(ctx, rc, "zf_nucleus_setup");
      if ((rc = zf_upload_nuclei(ctx, nxyz))) return die(ctx, rc, "zf_upload_nuclei");
      run.nnuc = (int32_t)nnuc;
      printf("zf_host: nucleated cells, %lld nucleus nodes, %lld filaments\n", (long long)nnuc,
             (long long)nen);
      free(nrest);
      free(nedge);
      free(nxyz);
    }
    zf_ibm_config ic;
    zf_ibm_config_reference(&ic);
    if (!reference) {
      ic.x0 = x0; ic.y0 = y0; ic.z0 = z0; ic.dx = dx;
      ic.vel_scale = 1.0; ic.force_scale = 1.0; ic.dt_advance = 1.0; ic.vmax = 1.0;
      ic.gamma_p = 0.5;
    }
    ic.gamma_p = getd("gamma_p", ic.gamma_p);
    ic.rbc_interval = geti("rbc_interval", ic.rbc_interval);
    if ((rc = zf_ibm_setup(ctx, &ic))) return die(ctx, rc, "zf_ibm_setup");
    run.ncell = ncell;
    run.nn = (int32_t)nn;
    run.nt = (int32_t)nt;
    if (geti("repulsion", 1)) {
      zf_repulsion_config rcfg;
      zf_repulsion_config_reference(&rcfg);
      if (!reference) {
        rcfg.cutoff = 1.0;
        rcfg.search = 2.0;
        rcfg.bin_cutoff = 2.0;
      }
      if ((rc = zf_repulsion_setup(ctx, &rcfg))) return die(ctx, rc, "zf_repulsion_setup");
      run.have_rep = 1;
      /* lumen wall: files in the reference's WallMesh format, or surface nets */
      int64_t nwn = 0, nwt = 0;
      double *wx = NULL, *wn = NULL;
      int32_t *wt = NULL;
      if (get("wall_nodes") && get("wall_cells") && get("wall_normals")) {
        nwn = zfh_read_points(get("wall_nodes"), NULL, 0);
        nwt = zfh_read_wall_cells(get("wall_cells"), NULL, 0);
        if (nwn <= 0 || nwt <= 0) return die(ctx, ZF_ERR_ARG, "reading the wall mesh");
        wx = malloc(sizeof(double) * 3 * nwn);
        wt = malloc(sizeof(int32_t) * 3 * nwt);
        wn = malloc(sizeof(double) * 3 * nwt);
        if (zfh_read_points(get("wall_nodes"), wx, nwn) != nwn ||
            zfh_read_wall_cells(get("wall_cells"), wt, nwt) != nwt ||
            zfh_read_wall_normals(get("wall_normals"), wn, nwt) != nwt)
          return die(ctx, ZF_ERR_ARG, "reading the wall mesh");
      } else if (get("wall") && strcmp(get("wall"), "surface") == 0) {
        if (zfh_wall_mesh(b, im, jm, km, x0, y0, z0, dx, NULL, 0, NULL, NULL, 0, &nwn, &nwt))
          return die(ctx, ZF_ERR_ARG, "wall mesh");
        wx = malloc(sizeof(double) * 3 * nwn);
        wt = malloc(sizeof(int32_t) * 3 * nwt);
        wn = malloc(sizeof(double) * 3 * nwt);
        if (zfh_wall_mesh(b, im, jm, km, x0, y0, z0, dx, wx, nwn, wt, wn, nwt, &nwn, &nwt))
          return die(ctx, ZF_ERR_ARG, "wall mesh");
      }
      if (wx) {
        if ((rc = zf_wall_setup(ctx, (int32_t)nwn, wx, (int32_t)nwt, wt, wn)))
          return die(ctx, rc, "zf_wall_setup");
        run.have_wall = 1;
        printf("zf_host: wall mesh %lld nodes, %lld triangles\n", (long long)nwn, (long long)nwt);
        /* wall-layer velocity: on by default with the reference's units */
        const char *uv = get("uvoid") ? get("uvoid") : reference ? "reference" : "off";
        if (strcmp(uv, "off") != 0) {
          zf_uvoid_config uc;
          zf_uvoid_config_reference(&uc);
          if (strcmp(uv, "node") == 0) uc.select = 0;
          uc.speed = getd("uvoid_speed", uc.speed);
          if ((rc = zf_uvoid_setup(ctx, &uc))) return die(ctx, rc, "zf_uvoid_setup");
          printf("zf_host: wall-layer velocity (%s)\n", uv);
        }
      }
      free(wx);
      free(wt);
      free(wn);
    }
    printf("zf_host: %d cells of %lld nodes\n", ncell, (long long)nn);
  }

  /* the loop */
  const char *out = get("output_dir") ? get("output_dir") : "SimulationOutputData";
  if (mkdir(out, 0755) && errno != EEXIST) {
    fprintf(stderr, "zf_host: cannot create %s\n", out);
    return die(ctx, ZF_ERR_ARG, "output directory");
  }
  run.out_dir = out;
  run.im = im; run.jm = jm; run.km = km;
  run.x0 = x0; run.y0 = y0; run.z0 = z0; run.dx = dx;
  run.iter0 = 0;
  run.steps = geti("steps", 1000);
  run.output_every = geti("output_every", 0);
  run.checkpoint_every = geti("checkpoint_every", 0);
  if (get("restart")) {
    int32_t it = 0;
    if ((rc = zfh_load_state(ctx, get("restart"), &it, run.ncell, run.nn)))
      return die(ctx, rc, "restart file");
    run.iter0 = it;
    printf("zf_host: restarted from %s at iteration %d\n", get("restart"), it);
  }
  if ((rc = zfh_run(ctx, b, &run))) return die(ctx, rc, "the time loop");
  printf("zf_host: %d iterations done\n", run.steps);
  zf_destroy(ctx);
  free(b);
  free(rest);
  free(tri);
  free(xyz);
  free(status);
  return 0;
}
