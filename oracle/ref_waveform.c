/*
 * ref_waveform.c -- evaluates the reference's own cardiac pressure
 * waveforms of BC codes 1001-1009 (ZebraflowWT.c:6318-6421) to write golden
 * vectors (TEST INFRASTRUCTURE; built and run only in the development
 * container, where the reference source exists).
 *
 * The waveforms sit inline in main()'s boundary loop and cannot be called,
 * so oracle/Makefile extracts, at build time and from the reference file
 * where it lies, its #define constants (ZebraflowWT.c:34-74) into
 * _ref/ref_defines.inc and the code-by-code branch chain through the D_bc
 * line (:6318-6421) into _ref/ref_waveform.inc; both stay in the
 * git-ignored _ref/.  This file declares the loop's variables with the
 * reference's types (float, ZebraflowWT.c:294) and evaluates the chain.
 *
 *   ref_waveform <out.bin> iter...
 *   out.bin: for each iter, for code 1001..1009: int32 iter, int32 code,
 *            float32 pressure, float32 D_bc
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "_ref/ref_defines.inc"

static void eval(int it, int code, float *p_out, float *d_out) {
  int BCLIST[2][4] = {{0}};
  const int l = 1;
  float pressure = 0.f, simtime, cycletime, taucycle, D_bc;
  int iter = it;
  BCLIST[l][3] = code;
  simtime = iter * T_SCALE; /* ZebraflowWT.c:6303 */
#include "_ref/ref_waveform.inc"
  (void)cycletime;
  (void)taucycle;
  *p_out = pressure;
  *d_out = D_bc;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s out.bin iter...\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "wb");
  if (!f) return 1;
  for (int a = 2; a < argc; a++) {
    const int it = atoi(argv[a]);
    for (int code = 1001; code <= 1009; code++) {
      float p, d;
      eval(it, code, &p, &d);
      const int hdr[2] = {it, code};
      fwrite(hdr, sizeof hdr, 1, f);
      fwrite(&p, sizeof p, 1, f);
      fwrite(&d, sizeof d, 1, f);
    }
  }
  fclose(f);
  return 0;
}
