"""Write the small data fixtures the tests use (run in a checkout next to the reference tree).

std_problem_4_field{1,2}.npz: the muMAG standard problem 4 consensus curves
shipped with the reference (matlab/examples/Micromagnetism/
mumag_micromag_Std_problem_4/Published_solutions_field{1,2}.mat: mean_avg and
err, <m_i>/Ms on 1000 points over 0..1 ns), read with scipy.io.loadmat (MAT v5,
data only) and stored as plain arrays.
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


if __name__ == "__main__":
    sys.exit(main())
