"""Write golden vectors for the membrane mesh topology (a7) and equilibrium
parameters (a8) from the reference's own CGSM_ZFISHRBC_DATA_GEN
(ZebraflowWT.c:1592-1925) and the 51-pass CGSM_FILAMENT_PARAMETERS /
CGSM_EQUILIBRIUM_METRICS loop (ZebraflowWT.c:2338-2346), run on the
zebrafish RBC mesh the reference ships (ZFISHRBCNODES/TRIANGLES.dat.gz).
Development container only:

    make -C oracle ref && python tests/golden/make_mesh_goldens.py

Output mesh_zfish.npz: input nodes and triangles, the reference's
EDGENODE_LIST (1-based n1 n2 n3 n4 t1 t2), NEIGHNODE_LIST, NEIGHTRI_LIST and
the float32 LMAX, Plength, kp, Theta0, SINETHETA0, COSINETHETA0; and the
nucleus network (NucleusNodes/Edges.dat): its node coordinates as read,
EDGENODE_LIST_N (1-based) and the float32 LMAX_N, Plength_N, kp_N.
"""
import gzip
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
REF = "/root/reference"
FILES = ["ZFISHRBCNODES.dat", "ZFISHRBCTRIANGLES.dat", "NucleusNodes.dat", "NucleusEdges.dat"]


def read_table(path, cols, dtype):
    with open(path) as f:
        f.readline()
        vals = f.read().split()
    return np.array(vals, dtype=dtype).reshape(-1, cols)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -C oracle ref")
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "InputData"))
        os.makedirs(os.path.join(td, "CGSMDataStructure"))
        for name in FILES:
            with gzip.open(os.path.join(REF, name + ".gz"), "rb") as src, \
                    open(os.path.join(td, "InputData", name), "wb") as dst:
                shutil.copyfileobj(src, dst)
        out = os.path.join(td, "mesh.bin")
        subprocess.run([HARNESS, "mesh", td, out], check=True, stdout=subprocess.DEVNULL)
        raw = np.fromfile(out, np.int32)
        nodes = read_table(os.path.join(td, "InputData", FILES[0]), 3, np.float64)
        tris = read_table(os.path.join(td, "InputData", FILES[1]), 3, np.int32)
    em = raw[0]
    o = 1
    edge = raw[o:o + 6 * em].reshape(em, 6); o += 6 * em
    nm = raw[o]; o += 1
    nbr_node = raw[o:o + 10 * nm].reshape(nm, 10); o += 10 * nm
    nbr_tri = raw[o:o + 10 * nm].reshape(nm, 10); o += 10 * nm
    fl = raw[o:o + 6 * em].view(np.float32).reshape(6, em); o += 6 * em
    enn = raw[o]; o += 1
    nedge = raw[o:o + 2 * enn].reshape(enn, 2); o += 2 * enn
    nnn = raw[o]; o += 1
    nxyz = raw[o:o + 6 * nnn].view(np.float64).reshape(nnn, 3); o += 6 * nnn
    nfl = raw[o:o + 3 * enn].view(np.float32).reshape(3, enn)
    np.savez_compressed(os.path.join(HERE, "mesh_zfish.npz"), nodes=nodes[:nm], tris=tris,
                        edge=edge, nbr_node=nbr_node, nbr_tri=nbr_tri, lmax=fl[0], plength=fl[1],
                        kp=fl[2], theta0=fl[3], sin0=fl[4], cos0=fl[5], nuc_nodes=nxyz,
                        nuc_edge=nedge, nuc_lmax=nfl[0], nuc_plength=nfl[1], nuc_kp=nfl[2])
    print(f"zebrafish mesh: {nm} nodes, {len(tris)} triangles, {em} edges; "
          f"nucleus {nnn} nodes, {enn} edges")


if __name__ == "__main__":
    main()
