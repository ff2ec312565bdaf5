/*
 * zf_oracle_lists.c -- index-list builders of the oracle (TEST INFRASTRUCTURE).
 * Restates LBM_DOM_SETUP (ZebraflowWT.c:1481-1502), LBM_BC_SETUP
 * (ZebraflowWT.c:1506-1589) and the pull-neighbour lookup of the stream step
 * (ZebraflowWT.c:6113-6125).  See zf_oracle.h for conventions.
 */
#include "zf_oracle.h"

#include <stddef.h>

/* Xlattice, ZebraflowWT.c:2471-2504: odd c and c+1 are opposites, 19 = rest. */
const int ZFO_E[19][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
    {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0},
    {1, 0, 1},  {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {0, 1, 1},
    {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 0, 0}};

/* Wcoeff, ZebraflowWT.c:2329-2333 (We1_3D, We2_3D, We3_3D at :75-77). */
const double ZFO_W[19] = {
    1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
    1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 3};

static inline size_t bidx(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

/* Lexicographic scan, i outer / k inner, over B >= 0 (ZebraflowWT.c:1485-1497). */
int zfo_dom_setup(const int32_t *b, int im, int jm, int km, int32_t *domid,
                  int32_t *domlist, int cap) {
  int n = 0;
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++) {
        size_t q = bidx(i, j, k, jm, km);
        domid[q] = -1;
        if (b[q] >= 0) {
          if (n >= cap) return -1;
          domlist[3 * n + 0] = i;
          domlist[3 * n + 1] = j;
          domlist[3 * n + 2] = k;
          n++;
          domid[q] = n;
        }
      }
  return n;
}

/* Neighbour of (i,j,k) in direction c, honouring periodic axes.  Returns 0
 * when outside a non-periodic box (the reference's `continue`,
 * ZebraflowWT.c:1541). */
static int step_to(int i, int j, int k, int c, int sign, int im, int jm,
                   int km, const int per[3], int *oi, int *oj, int *ok) {
  int ii = i + sign * ZFO_E[c][0];
  int jj = j + sign * ZFO_E[c][1];
  int kk = k + sign * ZFO_E[c][2];
  if (per[0]) { if (ii < 1) ii += im; else if (ii > im) ii -= im; }
  if (per[1]) { if (jj < 1) jj += jm; else if (jj > jm) jj -= jm; }
  if (per[2]) { if (kk < 1) kk += km; else if (kk > km) kk -= km; }
  if (ii > im || ii < 1 || jj > jm || jj < 1 || kk > km || kk < 1) return 0;
  *oi = ii; *oj = jj; *ok = kk;
  return 1;
}

/* BC nodes are B==0 or B>=1000 (ZebraflowWT.c:1527).  Neighbour choice in
 * three tiers, each scanning c = 1..18 in order: interior fluid 0<B<1000
 * (:1535-1547), else inlet/outlet B>1000 (:1549-1561), else wall B==0
 * (:1563-1574).  Note tier 2 tests B>1000, so a B==1000 neighbour is never
 * chosen there; this is kept. */
int zfo_bc_setup(const int32_t *b, int im, int jm, int km, const int per[3],
                 int32_t *bclist, int32_t *bcnlist, int cap, int *n_fail) {
  int n = 0, fails = 0;
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++) {
        int bv = b[bidx(i, j, k, jm, km)];
        if (!(bv >= 1000 || bv == 0)) continue;
        if (n >= cap) return -1;
        bclist[4 * n + 0] = i;
        bclist[4 * n + 1] = j;
        bclist[4 * n + 2] = k;
        bclist[4 * n + 3] = bv;
        int found = 0, fi = 0, fj = 0, fk = 0;
        for (int tier = 0; tier < 3 && !found; tier++) {
          for (int c = 0; c < 18; c++) {
            int ii, jj, kk;
            if (!step_to(i, j, k, c, +1, im, jm, km, per, &ii, &jj, &kk))
              continue;
            int nb = b[bidx(ii, jj, kk, jm, km)];
            int ok = (tier == 0) ? (nb > 0 && nb < 1000)
                     : (tier == 1) ? (nb > 1000)
                                   : (nb == 0);
            if (ok) { fi = ii; fj = jj; fk = kk; found = 1; break; }
          }
        }
        if (!found) fails++;
        bcnlist[3 * n + 0] = fi;
        bcnlist[3 * n + 1] = fj;
        bcnlist[3 * n + 2] = fk;
        n++;
      }
  if (n_fail) *n_fail = fails;
  return n;
}

/* Parent (feeder) lattice p, LBM_PARENTBC_SETUP ZebraflowWT.c:1015-1350
 * (parent 1 :1028-1137; parents 2 and 3 repeat it with their own codes and
 * overlap plane).  BC list: B == 0 or B >= 10 (:1045), lexicographic.
 * Neighbour of a B == 0 or B == 20+p node (:1053-1100): the first of
 * c = 1..18 inside the box with 0 < B < 10, else with B > 10, else with
 * B == 0.  A B == 10+p node (:1103-1110) and the velocity source (vBCNLIST,
 * :1112-1122) of a B == 20+p node are the main-domain node over the parent
 * node's column at the overlap plane k_over: the reference maps the parent
 * coordinate (X_p at k_over) to the main index with
 * (int)((X - X0)/DX + 1.01); the parent origins sit on the main lattice
 * (ZebraflowWT.c:85-112), so that is (i + off[0], j + off[1], k_over +
 * off[2]).  vbcn is (0,0,0) for the other nodes; a node of any other code
 * >= 10, or with no neighbour, gets (0,0,0) and is counted in *n_fail (the
 * reference leaves stale values there).  Pinned by reference goldens on the
 * three parent lattices the reference ships. */
int zfo_parent_bc_setup(const int32_t *b, int im, int jm, int km, int p, int k_over,
                        const int off[3], int32_t *bclist, int32_t *bcnlist, int32_t *vbcn,
                        int cap, int *n_fail) {
  const int per[3] = {0, 0, 0};
  int n = 0, fails = 0;
  for (int i = 1; i <= im; i++)
    for (int j = 1; j <= jm; j++)
      for (int k = 1; k <= km; k++) {
        const int bv = b[bidx(i, j, k, jm, km)];
        if (!(bv >= 10 || bv == 0)) continue;
        if (n >= cap) return -1;
        bclist[4 * n + 0] = i;
        bclist[4 * n + 1] = j;
        bclist[4 * n + 2] = k;
        bclist[4 * n + 3] = bv;
        int found = 0, fi = 0, fj = 0, fk = 0, vi = 0, vj = 0, vk = 0;
        if (bv == 0 || bv == 20 + p) {
          for (int tier = 0; tier < 3 && !found; tier++) {
            for (int c = 0; c < 18; c++) {
              int ii, jj, kk;
              if (!step_to(i, j, k, c, +1, im, jm, km, per, &ii, &jj, &kk)) continue;
              const int nb = b[bidx(ii, jj, kk, jm, km)];
              const int ok = (tier == 0) ? (nb > 0 && nb < 10) : (tier == 1) ? (nb > 10) : (nb == 0);
              if (ok) { fi = ii; fj = jj; fk = kk; found = 1; break; }
            }
          }
          if (bv == 20 + p) { vi = i + off[0]; vj = j + off[1]; vk = k_over + off[2]; }
        } else if (bv == 10 + p) {
          fi = i + off[0]; fj = j + off[1]; fk = k_over + off[2];
          found = 1;
        }
        if (!found) fails++;
        bcnlist[3 * n + 0] = fi;
        bcnlist[3 * n + 1] = fj;
        bcnlist[3 * n + 2] = fk;
        vbcn[3 * n + 0] = vi;
        vbcn[3 * n + 1] = vj;
        vbcn[3 * n + 2] = vk;
        n++;
      }
  if (n_fail) *n_fail = fails;
  return n;
}

/* Pull source of direction c is (i,j,k) - e_c: the reference steps along the
 * opposite direction cc (ZebraflowWT.c:6114-6119). */
void zfo_neighbours(const int32_t *domid, const int32_t *domlist, int dommax,
                    int im, int jm, int km, const int per[3], int32_t *nbr) {
  for (int l = 1; l <= dommax; l++) {
    int i = domlist[3 * (l - 1)], j = domlist[3 * (l - 1) + 1],
        k = domlist[3 * (l - 1) + 2];
    for (int c = 0; c < 18; c++) {
      int ii, jj, kk, v = -1;
      if (step_to(i, j, k, c, -1, im, jm, km, per, &ii, &jj, &kk))
        v = domid[bidx(ii, jj, kk, jm, km)];
      nbr[(size_t)c * dommax + (l - 1)] = v;
    }
  }
}
