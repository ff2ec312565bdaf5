"""The C binding shown in INTEGRATION.md compiles as one translation unit
against include/zf.h and links against libzfgpu.so (no GPU needed: nothing
runs).  The reference's own globals that the snippets read (B_index, F_OLD,
the mesh and wall arrays, loop counters) are declared here with the
reference's names and types (ZebraflowWT.c:81-122, :176-277), as they are
when the snippets are pasted into ZebraflowWT.c."""
import os
import re
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = r"""
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#define IM 24
#define JM 20
#define JM_ 0
#define KM 20
#define KM_1 121
#define NodeM 642
#define TriM 1280
#define NodeM_N 121
#define EdgeM_N 544
#define CapsM 8
#define NSTEPS 100
#define OUTPUTSTEP 10
int B_index[IM + 1][JM + 1][KM + 1];
float F_OLD[19 + 1][IM * JM * KM + 1];
int iter, simfailure, caps = 8, ncells = 1;
int32_t tri_1based[3 * TriM];
double rest_xyz[3 * NodeM], xyz[3 * NodeM];
int WallNodeM = 3, WallTriM = 1;
double wall_xyz[9], wall_normal[3];
int32_t wall_tri[3] = {1, 2, 3};
int32_t pim[3], pjm[3], pkm[3];
int32_t *parent_b[3];
double *parent_f[3];
double D1[IM * JM * KM], U1[IM * JM * KM], V1[IM * JM * KM], W1[IM * JM * KM];
double vel[3 * NodeM];
int EDGENODE_LIST_N[EdgeM_N + 1][3];
double NODECOORD_N[3][CapsM + 1][NodeM_N + 1];
int32_t status[8];
double area[8], vol[8];
double dJ_dx0[3 * NodeM], dJ_df0[19 * IM * JM * KM], dJ_dF[3 * IM * JM * KM];
"""

EPILOGUE = r"""
int main(void) {
  (void)zf_bind_setup; (void)zf_bind_nuclei; (void)zf_bind_loop; (void)zf_bind_parents;
  (void)zf_bind_parent_steps; (void)zf_bind_gradient; (void)zf_bind_parent_cells;
  (void)zf_bind_linked_step;
  return 0;
}
"""


def test_integration_snippets_compile_and_link(tmp_path):
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, flags=re.S)
    assert len(blocks) >= 7
    src = tmp_path / "binding.c"
    src.write_text(PRELUDE + "\n".join(blocks) + EPILOGUE)
    exe = tmp_path / "binding"
    pkg = os.path.join(REPO, "zebraflow_amd")
    r = subprocess.run(["gcc", "-std=gnu11", "-Wall", "-Werror", "-Wno-unused-function",
                        "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe),
                        "-L", pkg, "-lzfgpu", "-Wl,-rpath," + pkg, "-lm"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
