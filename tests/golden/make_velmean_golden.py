"""Writes the VELMEAN post-processing fixtures.

velmean_profiles-2.dat: a synthetic profiles-2.dat, seeded, in the format of
the statistics writer (src/output.cu:325-333: an "%8.8d" step line, then one
row of nine "% 7.7e" columns per wall-normal index), on the block schedule
istart, istart + step, ..., iend.

velmean_VELMEAN.dat: what the reference's post-processing script
examples/channel/ProcessProfiles.py:1-79 writes from that file, run here on a
params module of this fixture's schedule (NZ, ISTART, IEND, STEP,
IGNORE_BEFORE below) in a scratch directory.

The script reads /root/reference, so it runs in the build container only;
the two output files are committed data.  Run: python tests/golden/make_velmean_golden.py
"""
import os
import runpy
import shutil
import sys
import tempfile

import numpy as np

REF_SCRIPT = "/root/reference/examples/channel/ProcessProfiles.py"
HERE = os.path.dirname(os.path.abspath(__file__))
NZ, ISTART, IEND, STEP, IGNORE_BEFORE = 8, 1, 20001, 5000, 5001
ROW = "% 7.7e % 7.7e % 7.7e   % 7.7e % 7.7e % 7.7e   % 7.7e % 7.7e % 7.7e \n"


def synthetic_profiles(path):
    """Mean velocities ~0.1, shear stresses ~1e-6 of either sign, variances
    ~1e-4 (positive), one block per scheduled step."""
    rng = np.random.default_rng(20261203)
    scale = np.array([0.1, 0.01, 0.1, 1e-6, 1e-6, 1e-6, 1e-4, 1e-4, 1e-4])
    with open(path, "w") as f:
        for step in range(ISTART, IEND + STEP, STEP):
            f.write("%8.8d\n" % step)
            for _ in range(NZ):
                row = rng.standard_normal(9) * scale
                row[6:] = np.abs(row[6:])
                f.write(ROW % tuple(row))


def main():
    with tempfile.TemporaryDirectory() as d:
        prof = os.path.join(d, "profiles-2.dat")
        synthetic_profiles(prof)
        with open(os.path.join(d, "params.py"), "w") as f:
            f.write(f"istart={ISTART}\niend={IEND}\nstep={STEP}\nignore_before={IGNORE_BEFORE}\nnz={NZ}\n")
        cwd = os.getcwd()
        sys.path.insert(0, d)
        sys.modules.pop("params", None)
        os.chdir(d)
        try:
            runpy.run_path(REF_SCRIPT, run_name="__main__")
        finally:
            os.chdir(cwd)
            sys.path.remove(d)
            sys.modules.pop("params", None)
        shutil.copy(prof, os.path.join(HERE, "velmean_profiles-2.dat"))
        shutil.copy(os.path.join(d, "VELMEAN.dat"), os.path.join(HERE, "velmean_VELMEAN.dat"))
    print("wrote velmean_profiles-2.dat and velmean_VELMEAN.dat")


if __name__ == "__main__":
    main()
