"""Write the small data fixtures the tests use (run in a checkout next to the reference tree).

std_problem_4_field{1,2}.npz: the muMAG standard problem 4 consensus curves
shipped with the reference (matlab/examples/Micromagnetism/
mumag_micromag_Std_problem_4/Published_solutions_field{1,2}.mat: mean_avg and
err, <m_i>/Ms on 1000 points over 0..1 ns), read with scipy.io.loadmat (MAT v5,
data only) and stored as plain arrays.

comsol_prism_normH.npz: the COMSOL 5.4 line profiles of |H| around one rotated
prism magnet shipped with the reference
(documentation/examples_FEM_validation/Validation_prism/
Validation_prism_normH_{x,y,z}.txt: coordinate along the line and the field
norm, 8 header lines), for the tile of python/examples/magnetostatics/
validation.py:9-20.  Stored as (coordinate, normH) pairs per axis.
"""

import os
import sys

import numpy as np
import scipy.io as sio

REF = os.environ.get("MAGTENSE_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(REF, "matlab", "examples", "Micromagnetism", "mumag_micromag_Std_problem_4")
    for k in (1, 2):
        d = sio.loadmat(os.path.join(src, f"Published_solutions_field{k}.mat"))
        np.savez_compressed(os.path.join(HERE, f"std_problem_4_field{k}.npz"),
                            t=np.linspace(0.0, 1e-9, 1000), mean_avg=d["mean_avg"][0], err=d["err"][0])
        print("wrote field", k, d["mean_avg"].shape)
    comsol = {}
    vdir = os.path.join(REF, "documentation", "examples_FEM_validation", "Validation_prism")
    for ax in "xyz":
        with open(os.path.join(vdir, f"Validation_prism_normH_{ax}.txt")) as fh:
            rows = [ln.split() for ln in fh.readlines()[8:] if ln.strip()]
        comsol[ax] = np.asarray(rows, dtype=np.float64)
        print("COMSOL line", ax, comsol[ax].shape)
    np.savez_compressed(os.path.join(HERE, "comsol_prism_normH.npz"), **comsol)


if __name__ == "__main__":
    sys.exit(main())
