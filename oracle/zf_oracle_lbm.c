/*
 * zf_oracle_lbm.c -- LBM step of the oracle (TEST INFRASTRUCTURE).
 * See zf_oracle_lbm.inc for the step itself and zf_oracle.h for conventions.
 */
#include <math.h>
#include <stddef.h>
#include <omp.h>

#include "zf_oracle.h"

static inline size_t bidx(int i, int j, int k, int jm, int km) {
  return ((size_t)(i - 1) * jm + (size_t)(j - 1)) * km + (size_t)(k - 1);
}

/* feq with C_LBM = 1: EQ_A = 3, EQ_B = 4.5, EQ_C = -1.5 (ZebraflowWT.c:56-61,
 * formula :6139). */
double zfo_feq(int c, double rho, double u, double v, double w) {
  const double cu = u * ZFO_E[c][0] + v * ZFO_E[c][1] + w * ZFO_E[c][2];
  return rho * ZFO_W[c] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (u * u + v * v + w * w));
}

/* Cardiac pressure waveforms of BC codes 1001-1009, ZebraflowWT.c:6318-6421,
 * with the reference's types: simtime, cycletime, taucycle and pressure are
 * float (:294), the constants double, so every expression is evaluated in
 * double and rounded to float at each assignment, and the sines are sinf.
 * Each code is (phase shift, base, amplitude, shape); code 1001 has no shift
 * (its fmod of a non-negative time needs no wrap, which the shared code
 * below then never takes). */
float zfo_bc_pressure(int code, float simtime) {
  const double P = 3.14159265359, T = 0.36;
  double shift = 0.2 * T, base, amp;
  int shape_a = 1; /* 1: sin^2 rise to 0.23 then cosine fall; 0: sin rise to 0.3 */
  switch (code) {
    case 1001: shift = 0.0; base = 54.0; amp = 82.0; break;
    case 1002: shift = 0.06 * T; base = 48.0; amp = 68.0; break;
    case 1003: base = 11.0; amp = 23.0; shape_a = 0; break;
    case 1004: base = 18.0; amp = 27.0; shape_a = 0; break;
    case 1005: base = 23.0; amp = 27.0; shape_a = 0; break;
    case 1006: case 1008: base = 29.0; amp = 66.0; break;
    case 1007: case 1009: base = 28.5; amp = 60.0; break;
    default: return 0.0f;
  }
  float cyc = (float)fmod(simtime - shift, T);
  if (cyc < 0) cyc = (float)(cyc + T);
  const float tc = (float)(cyc / T);
  float pr;
  if (shape_a) {
    if (tc <= 0.23) {
      const float s = sinf((float)(P * tc / 0.46));
      pr = (float)(120.0 + base + amp * s * s);
    } else {
      pr = (float)(120.0 + base +
                   amp * (1.0 + sinf((float)(2 * P * (tc + 0.54) / 1.54 - P / 2.0))) / 2.0);
    }
  } else {
    if (tc <= 0.3) pr = (float)(120.0 + base + amp * sinf((float)(P * tc / 0.6)));
    else pr = (float)(120.0 + base + amp * (1.0 + sinf((float)(2 * P * (tc + 0.4) / 1.4 - P / 2.0))) / 2.0);
  }
  return pr;
}

double zfo_wtime(void) { return omp_get_wtime(); }
int zfo_max_threads(void) { return omp_get_max_threads(); }

#define REAL double
#define FN(x) x##_f64
#include "zf_oracle_lbm.inc"
#undef REAL
#undef FN

#define REAL float
#define FN(x) x##_f32
#include "zf_oracle_lbm.inc"
#undef REAL
#undef FN
