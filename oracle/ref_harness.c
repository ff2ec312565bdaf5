/*
 * ref_harness.c -- drives the reference's own setup functions to write
 * golden vectors (TEST INFRASTRUCTURE; runs only in the development
 * container, where the reference source exists).
 *
 * Linked against ZebraflowWT.c compiled from its original location with
 * -Dmain=zf_main (see oracle/Makefile).  The reference keeps all state in
 * static globals sized by compile-time macros (ZebraflowWT.c:81-122,
 * :171-277); this file only declares the ones it fills or reads.
 *
 * Modes:
 *   ref_harness lists <in.bin> <out.bin>
 *     in.bin  : int32 im,jm,km then im*jm*km int32 B values (i outer, k
 *               inner).  The block is placed at lattice (1,1,1) of the
 *               reference's 249x547x723 box; the rest is B=-1.
 *     out.bin : int32 DOMMAX, DOMLIST (DOMMAX*3), BCMAX, BCLIST (BCMAX*4),
 *               BCNLIST (BCMAX*3); all as the reference filled them.
 *     Calls LBM_DOM_SETUP (ZebraflowWT.c:1481) and LBM_BC_SETUP (:1506).
 *
 *   ref_harness mesh <workdir> <out.bin>
 *     Runs CGSM_ZFISHRBC_DATA_GEN (:1592) in <workdir> (which must hold
 *     InputData/ZFISHRBC*.dat, InputData/Nucleus*.dat and an empty
 *     CGSMDataStructure/), then the 51-pass CGSM_FILAMENT_PARAMETERS /
 *     CGSM_EQUILIBRIUM_METRICS loop exactly as main() does (:2339-2346).
 *     out.bin : int32 EdgeM, EDGENODE_LIST rows 1..EdgeM (6 ints each),
 *               NodeM, NEIGHNODE_LIST rows (10 ints), NEIGHTRI_LIST rows
 *               (10 ints), then float32 LMAX, Plength, kp, Theta0,
 *               SINETHETA0, COSINETHETA0 (EdgeM each); then the nucleus:
 *               int32 EdgeM_N, EDGENODE_LIST_N rows (2 ints), NodeM_N,
 *               double NODECOORD_N of cell 1 (3 per node), float32 LMAX_N,
 *               Plength_N, kp_N (EdgeM_N each).
 *
 *   ref_harness parents <workdir> <out.bin>
 *     Runs LBM_PARENTDOMAIN_SETUP (:515) and LBM_PARENTBC_SETUP (:1015) in
 *     <workdir>, which must hold InputData/DAparentBoundary.csv,
 *     InputData/PCVparent1Boundary.csv and InputData/PCVparent2Boundary.csv
 *     and an empty LBMGridDataStructure/ (the setup writes its domain files
 *     there).
 *     out.bin, for parents 1..3 in turn: int32 im, jm, km, B_index_p
 *     (im*jm*km, i outer, k inner), DOMMAX_p, BCMAX_p, BCLIST_p (BCMAX*4),
 *     BCNLIST_p (BCMAX*3), vBCNLIST_p (BCMAX*3).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define R_IM 249
#define R_JM 547
#define R_KM 723
#define R_NODEM 504
#define R_EDGEM 1506
#define R_TRIM 1004
#define R_NLISTM 9

extern int B_index[R_IM + 1][R_JM + 1][R_KM + 1];
extern int DOMLIST[8500000][3], DOMMAX;
extern int BCLIST[1200000][4], BCNLIST[1200000][3], BCMAX;
extern float Xlattice[3][19 + 1];
extern int iter, iter_local;

extern int EDGENODE_LIST[R_EDGEM + 1][6 + 1];
extern int NEIGHNODE_LIST[R_NODEM + 1][R_NLISTM + 1];
extern int NEIGHTRI_LIST[R_NODEM + 1][R_NLISTM + 1];
extern float LMAX[R_EDGEM + 1], Plength[R_EDGEM + 1], kp[R_EDGEM + 1],
    Theta0[R_EDGEM + 1], SINETHETA0[R_EDGEM + 1], COSINETHETA0[R_EDGEM + 1];

/* the nucleus network, ZebraflowWT.c:120-122, :237-240 */
#define R_CAPSM 3000
#define R_NODEM_N 121
#define R_EDGEM_N 544
extern int EDGENODE_LIST_N[R_EDGEM_N + 1][2 + 1];
extern double NODECOORD_N[3][R_CAPSM + 1][R_NODEM_N + 1];
extern float LMAX_N[R_EDGEM_N + 1], Plength_N[R_EDGEM_N + 1], kp_N[R_EDGEM_N + 1];

/* the three parent (feeder) lattices, ZebraflowWT.c:91-112, :188-230 */
#define P1_IM 69
#define P1_JM 56
#define P1_KM 121
#define P2_IM 64
#define P2_JM 60
#define P2_KM 121
#define P3_IM 42
#define P3_JM 27
#define P3_KM 121
extern int B_index_1[P1_IM + 1][P1_JM + 1][P1_KM + 1];
extern int B_index_2[P2_IM + 1][P2_JM + 1][P2_KM + 1];
extern int B_index_3[P3_IM + 1][P3_JM + 1][P3_KM + 1];
extern int BCLIST_1[60000][4], BCNLIST_1[60000][3], BCMAX_1, DOMMAX_1, vBCNLIST_1[60000][3];
extern int BCLIST_2[60000][4], BCNLIST_2[60000][3], BCMAX_2, DOMMAX_2, vBCNLIST_2[60000][3];
extern int BCLIST_3[60000][4], BCNLIST_3[60000][3], BCMAX_3, DOMMAX_3, vBCNLIST_3[60000][3];
void LBM_PARENTDOMAIN_SETUP(void);
void LBM_PARENTBC_SETUP(void);

void LBM_DOM_SETUP(void);
void LBM_BC_SETUP(void);
void CGSM_ZFISHRBC_DATA_GEN(void);
void CGSM_FILAMENT_PARAMETERS(void);
void CGSM_EQUILIBRIUM_METRICS(void);

/* The lattice table main() fills at ZebraflowWT.c:2471-2504. */
static const int E[19][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
    {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0},
    {1, 0, 1},  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {0, 1, 1},
    {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 0, 0}};

static void wr(FILE *f, const void *p, size_t sz, size_t n) {
  if (fwrite(p, sz, n, f) != n) { perror("fwrite"); exit(2); }
}

static int run_lists(const char *in, const char *out) {
  FILE *f = fopen(in, "rb");
  if (!f) { perror(in); return 2; }
  int dims[3];
  if (fread(dims, 4, 3, f) != 3) return 2;
  const int im = dims[0], jm = dims[1], km = dims[2];
  if (im > R_IM || jm > R_JM || km > R_KM) { fprintf(stderr, "box too big\n"); return 2; }
  int *blk = malloc(sizeof(int) * (size_t)im * jm * km);
  if (fread(blk, 4, (size_t)im * jm * km, f) != (size_t)im * jm * km) return 2;
  fclose(f);
  for (int i = 1; i <= R_IM; i++)
    for (int j = 1; j <= R_JM; j++)
      for (int k = 1; k <= R_KM; k++) B_index[i][j][k] = -1;
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++)
        B_index[i][j][k] = blk[((size_t)(i - 1) * jm + (j - 1)) * km + (k - 1)];
  for (int c = 1; c <= 19; c++)
    for (int d = 0; d < 3; d++) Xlattice[d][c] = (float)E[c - 1][d];
  iter = 0;
  LBM_DOM_SETUP();
  LBM_BC_SETUP();
  FILE *o = fopen(out, "wb");
  if (!o) { perror(out); return 2; }
  wr(o, &DOMMAX, 4, 1);
  wr(o, &DOMLIST[1][0], 4, (size_t)3 * DOMMAX);
  wr(o, &BCMAX, 4, 1);
  wr(o, &BCLIST[1][0], 4, (size_t)4 * BCMAX);
  wr(o, &BCNLIST[1][0], 4, (size_t)3 * BCMAX);
  fclose(o);
  free(blk);
  return 0;
}

static int run_mesh(const char *dir, const char *out) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return 2;
  char outabs[4096];
  if (out[0] == '/') snprintf(outabs, sizeof outabs, "%s", out);
  else snprintf(outabs, sizeof outabs, "%s/%s", cwd, out);
  if (chdir(dir) != 0) { perror(dir); return 2; }
  CGSM_ZFISHRBC_DATA_GEN();
  /* main(): iter_local runs 1..51 around the two parameter routines
   * (ZebraflowWT.c:2338-2346). */
  iter_local = 0;
  do {
    iter_local += 1;
    CGSM_FILAMENT_PARAMETERS();
    CGSM_EQUILIBRIUM_METRICS();
  } while (iter_local <= 50);
  FILE *o = fopen(outabs, "wb");
  if (!o) { perror(outabs); return 2; }
  int em = R_EDGEM, nm = R_NODEM;
  wr(o, &em, 4, 1);
  for (int e = 1; e <= R_EDGEM; e++) wr(o, &EDGENODE_LIST[e][1], 4, 6);
  wr(o, &nm, 4, 1);
  for (int n = 1; n <= R_NODEM; n++) wr(o, &NEIGHNODE_LIST[n][0], 4, 10);
  for (int n = 1; n <= R_NODEM; n++) wr(o, &NEIGHTRI_LIST[n][0], 4, 10);
  wr(o, &LMAX[1], 4, R_EDGEM);
  wr(o, &Plength[1], 4, R_EDGEM);
  wr(o, &kp[1], 4, R_EDGEM);
  wr(o, &Theta0[1], 4, R_EDGEM);
  wr(o, &SINETHETA0[1], 4, R_EDGEM);
  wr(o, &COSINETHETA0[1], 4, R_EDGEM);
  /* the nucleus network of the same set-up: edges, the node coordinates as
   * read (cell 1), and its equilibrium parameters (0.1 Es, :2003-2050, and
   * :2203-2221) */
  int en = R_EDGEM_N, nn_n = R_NODEM_N;
  wr(o, &en, 4, 1);
  for (int e = 1; e <= R_EDGEM_N; e++) wr(o, &EDGENODE_LIST_N[e][1], 4, 2);
  wr(o, &nn_n, 4, 1);
  for (int n = 1; n <= R_NODEM_N; n++)
    for (int a = 0; a < 3; a++) wr(o, &NODECOORD_N[a][1][n], 8, 1);
  wr(o, &LMAX_N[1], 4, R_EDGEM_N);
  wr(o, &Plength_N[1], 4, R_EDGEM_N);
  wr(o, &kp_N[1], 4, R_EDGEM_N);
  fclose(o);
  return 0;
}

static void wr_parent(FILE *o, int im, int jm, int km, const int *b, int dommax, int bcmax,
                      const int (*bcl)[4], const int (*bcn)[3], const int (*vbc)[3]) {
  const int dims[3] = {im, jm, km};
  wr(o, dims, 4, 3);
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      wr(o, &b[((size_t)i * (jm + 1) + j) * (km + 1) + 1], 4, km);
  wr(o, &dommax, 4, 1);
  wr(o, &bcmax, 4, 1);
  wr(o, &bcl[1][0], 4, (size_t)4 * bcmax);
  wr(o, &bcn[1][0], 4, (size_t)3 * bcmax);
  wr(o, &vbc[1][0], 4, (size_t)3 * bcmax);
}

static int run_parents(const char *dir, const char *out) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return 2;
  char outabs[4096];
  if (out[0] == '/') snprintf(outabs, sizeof outabs, "%s", out);
  else snprintf(outabs, sizeof outabs, "%s/%s", cwd, out);
  if (chdir(dir) != 0) { perror(dir); return 2; }
  for (int c = 1; c <= 19; c++)
    for (int d = 0; d < 3; d++) Xlattice[d][c] = (float)E[c - 1][d];
  iter = 0;
  LBM_PARENTDOMAIN_SETUP();
  LBM_PARENTBC_SETUP();
  FILE *o = fopen(outabs, "wb");
  if (!o) { perror(outabs); return 2; }
  wr_parent(o, P1_IM, P1_JM, P1_KM, &B_index_1[0][0][0], DOMMAX_1, BCMAX_1, BCLIST_1, BCNLIST_1,
            vBCNLIST_1);
  wr_parent(o, P2_IM, P2_JM, P2_KM, &B_index_2[0][0][0], DOMMAX_2, BCMAX_2, BCLIST_2, BCNLIST_2,
            vBCNLIST_2);
  wr_parent(o, P3_IM, P3_JM, P3_KM, &B_index_3[0][0][0], DOMMAX_3, BCMAX_3, BCLIST_3, BCNLIST_3,
            vBCNLIST_3);
  fclose(o);
  return 0;
}

int main(int argc, char **argv) {
  if (argc == 4 && strcmp(argv[1], "lists") == 0) return run_lists(argv[2], argv[3]);
  if (argc == 4 && strcmp(argv[1], "mesh") == 0) return run_mesh(argv[2], argv[3]);
  if (argc == 4 && strcmp(argv[1], "parents") == 0) return run_parents(argv[2], argv[3]);
  fprintf(stderr, "usage: ref_harness lists <in.bin> <out.bin> | mesh <dir> <out.bin> | parents <dir> <out.bin>\n");
  return 1;
}
