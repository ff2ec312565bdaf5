"""Write golden vectors for the parent (feeder) lattices' lists from the
reference's own LBM_PARENTDOMAIN_SETUP (ZebraflowWT.c:515-1013) and
LBM_PARENTBC_SETUP (ZebraflowWT.c:1015-1350), run on the three parent
boundary files the reference ships (DAparentBoundary.csv.gz,
PCVparent1Boundary.csv.gz, PCVparent2Boundary.csv.gz).  Development
container only:

    make -C oracle ref && python tests/golden/make_parent_goldens.py

Output parents_ref.npz, for p = 1, 2, 3: b{p} (the parent's B_index box as
the reference fills it: 0 wall, 1 fluid, 10+p the boundary copied from the
main domain, 20+p the upstream boundary), dommax{p}, bclist{p} (BCMAX x 4),
bcnlist{p} (BCMAX x 3) and vbcn{p} (BCMAX x 3, the main-domain node whose
velocity an upstream node takes; zero elsewhere).
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
FILES = ["DAparentBoundary.csv", "PCVparent1Boundary.csv", "PCVparent2Boundary.csv"]


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -C oracle ref")
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(os.path.join(td, "InputData"))
        os.makedirs(os.path.join(td, "LBMGridDataStructure"))  # the setup writes its domain files there
        for name in FILES:
            with gzip.open(os.path.join(REF, name + ".gz"), "rb") as src, \
                    open(os.path.join(td, "InputData", name), "wb") as dst:
                shutil.copyfileobj(src, dst)
        out = os.path.join(td, "parents.bin")
        subprocess.run([HARNESS, "parents", td, out], check=True, stdout=subprocess.DEVNULL)
        raw = np.fromfile(out, np.int32)
    res = {}
    o = 0
    for p in (1, 2, 3):
        im, jm, km = raw[o:o + 3]; o += 3
        res[f"b{p}"] = raw[o:o + im * jm * km].reshape(im, jm, km).astype(np.int16); o += im * jm * km
        res[f"dommax{p}"] = np.int32(raw[o]); o += 1
        bcmax = raw[o]; o += 1
        res[f"bclist{p}"] = raw[o:o + 4 * bcmax].reshape(bcmax, 4); o += 4 * bcmax
        res[f"bcnlist{p}"] = raw[o:o + 3 * bcmax].reshape(bcmax, 3); o += 3 * bcmax
        res[f"vbcn{p}"] = raw[o:o + 3 * bcmax].reshape(bcmax, 3); o += 3 * bcmax
        print(f"parent {p}: box {im}x{jm}x{km}, DOMMAX {res[f'dommax{p}']}, BCMAX {bcmax}, "
              f"codes {sorted(set(res[f'bclist{p}'][:, 3].tolist()))}")
    assert o == raw.size
    np.savez_compressed(os.path.join(HERE, "parents_ref.npz"), **res)


if __name__ == "__main__":
    main()
