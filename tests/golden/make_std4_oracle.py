"""Oracle fixture for the fine-grid muMAG standard problem 4 check (about
11 minutes on 8 cores; run once, from the repository root):

    python tests/golden/make_std4_oracle.py 1 128 32

Runs the FP64 oracle (FP32 demag semantics) exactly as the reference example
python/examples/micromagnetism/std_problem_4.py:8-46 does, at the given
resolution, and stores the averaged magnetization curve of the dynamic run
with the times the oracle returned.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

MU0 = 4 * np.pi * 1e-7


def main():
    field = int(sys.argv[1])
    res = [int(sys.argv[2]), int(sys.argv[3]), 1]
    L = [500e-9, 125e-9, 3e-9]
    n = res[0] * res[1]
    nth = 2000
    ht = np.linspace(0, 100e-9, nth)
    hext = np.zeros((nth, 4))
    hext[:, 0] = ht
    hext[:, 1:] = np.where(ht < 1e-9, 1e-9 - ht, 0)[:, None] * (np.array([1, 1, 1]) / MU0)
    q = O.Problem(res, L, alpha=4.42e3, gamma=0.0, hext=hext, t=np.linspace(0, 100e-9, 200),
                  m0=np.full(3 * n, 1 / np.sqrt(3)), return_fields=False)
    _, y, _, _ = q.run()
    ms = y[0, -1].reshape(-1)
    h_nist = np.array([-24.6, 4.3, 0]) if field == 1 else np.array([-35.5, -6.3, 0])
    ht = np.linspace(0, 1e-9, nth)
    hext = np.zeros((nth, 4))
    hext[:, 0] = ht
    hext[:, 1:] = h_nist / 1000 / MU0
    t = np.linspace(0, 1e-9, 200)
    q = O.Problem(res, L, alpha=4.42e3, gamma=2.21e5, hext=hext, t=t, m0=ms, return_fields=False)
    t_out, y, _, st = q.run()
    out = os.path.join(HERE, f"std_problem_4_oracle_{res[0]}x{res[1]}_field{field}.npz")
    # t: the times the oracle returned (t_got < t_want where the work limit fired), t_want: the output grid
    np.savez_compressed(out, t=t_out[0], t_want=t, mean=y[0].mean(axis=2))
    print("wrote", out, st)


if __name__ == "__main__":
    main()
