"""Write golden vectors for the cardiac pressure waveforms of BC codes
1001-1009 (ZebraflowWT.c:6318-6421).

The expected values come from the reference's own branch chain, evaluated
with its float variables by ``oracle/_ref/ref_waveform`` (built by
``oracle/Makefile`` target ``ref`` from the reference file where it lies).
Run in the development container only:

    make -C oracle ref && python tests/golden/make_waveform_goldens.py

``waveforms_ref.npz`` holds, per sampled iteration and code, the float32
pressure and D_bc = pressure/LBM_PRESSURE_SCALE + 1 the reference computes.
The iterations sample more than one 0.36 s cycle (720,000 steps of
T_SCALE = 5e-7 s) evenly and add the branch joins (tau_cycle = 0.23 or
0.3), the cycle wraps of every phase shift (0, 0.06 T, 0.2 T) with their
neighbours, and late iterations where the float time loses resolution.
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(REPO, "oracle", "_ref", "ref_waveform")


def iterations():
    its = set(range(0, 800000, 1999))
    # branch joins and cycle wraps (in steps of 5e-7 s) and their neighbours
    for centre in (165600, 208800, 309600, 360000, 144000, 43200, 720000, 763200, 864000,
                   885600, 1080000):
        its.update(centre + d for d in (-2, -1, 0, 1, 2))
    its.update([1234567, 2345678, 4999999, 16777215, 16777217, 33554431])
    return sorted(its)


def main():
    its = iterations()
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "w.bin")
        subprocess.run([EXE, out, *map(str, its)], check=True)
        raw = np.fromfile(out, np.int32).reshape(-1, 4)
    it, code = raw[:, 0], raw[:, 1]
    vals = raw[:, 2:].copy().view(np.float32)
    np.savez_compressed(os.path.join(HERE, "waveforms_ref.npz"), iter=it, code=code,
                        pressure=vals[:, 0], d_bc=vals[:, 1])
    print(f"{len(its)} iterations x 9 codes")


if __name__ == "__main__":
    main()
