"""Whole-pipeline checks on muMAG standard problem 4, run exactly as the
reference example python/examples/micromagnetism/std_problem_4.py:8-46 does:
s-state relaxation (100 ns, gamma = 0) from m = 1/sqrt(3), then 1 ns of
dynamics in the NIST field.

* Parity: the GPU run equals the FP64 oracle's run of the same problem at the
  example's resolution (64 x 16 x 1): averaged magnetization curves within
  2e-3 (they agree to ~1e-8 for field 1).
* Physics: against the muMAG consensus curves shipped with the reference
  (fixtures in tests/golden, from Published_solutions_field{1,2}.mat).  The
  reference's own comparison is visual (Standard_problem_4.m:185-205,
  Benchmark_results.m:79-83: the integral of |<m_i> - <m_i>_NIST| over time).
  At the example's 7.8 nm cells the discretisation dominates: the oracle
  itself deviates by (0.033, 0.078, 0.017) for field 1 and (0.072, 0.170,
  0.045) for field 2 (time-averaged), so the bars there are 0.1 and 0.2.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MU0 = 4 * np.pi * 1e-7
GRID_L = [500e-9, 125e-9, 3e-9]


def _nist(field):
    return np.array([-24.6, 4.3, 0]) if field == 1 else np.array([-35.5, -6.3, 0])


def _run_std4(field, res=(64, 16, 1), precision="fp32"):
    from magtense_amd.micromag import MicromagProblem

    p_ini = MicromagProblem(list(res), GRID_L, m0=1 / np.sqrt(3), alpha=4.42e3, precision=precision)
    h_ext = np.array([1, 1, 1]) / MU0

    def h_ini(t):
        return np.expand_dims(np.where(t < 1e-09, 1e-09 - t, 0), axis=1) * h_ext

    _, M_ini, *_ = p_ini.run_simulation(100e-9, 200, h_ini, 2000)
    m_s = np.squeeze(M_ini, axis=2)[-1]
    p_dym = MicromagProblem(list(res), GRID_L, m0=m_s, alpha=4.42e3, gamma=2.21e5, precision=precision)
    h_nist = _nist(field)

    def h_dym(t):
        return np.expand_dims(t > -1, axis=1) * (h_nist / 1000 / MU0)

    t, M, *_ = p_dym.run_simulation(1e-9, 200, h_dym, 2000)
    return t, np.squeeze(M, axis=2).mean(axis=1), m_s, p_dym.last_stats[0]


def _oracle_std4(oracle, field, res=(64, 16, 1)):
    """The same two runs with the FP64 oracle (FP32 demag semantics)."""
    n = int(np.prod(res))
    nth = 2000
    ht = np.linspace(0, 100e-9, nth)
    hext = np.zeros((nth, 4))
    hext[:, 0] = ht
    hext[:, 1:] = np.where(ht < 1e-9, 1e-9 - ht, 0)[:, None] * (np.array([1, 1, 1]) / MU0)
    q = oracle.Problem(list(res), GRID_L, alpha=4.42e3, gamma=0.0, hext=hext, t=np.linspace(0, 100e-9, 200),
                       m0=np.full(3 * n, 1 / np.sqrt(3)), return_fields=False)
    _, y, _, _ = q.run()
    ms = y[0, -1].reshape(-1)
    ht = np.linspace(0, 1e-9, nth)
    hext = np.zeros((nth, 4))
    hext[:, 0] = ht
    hext[:, 1:] = _nist(field) / 1000 / MU0
    t = np.linspace(0, 1e-9, 200)
    q = oracle.Problem(list(res), GRID_L, alpha=4.42e3, gamma=2.21e5, hext=hext, t=t, m0=ms, return_fields=False)
    _, y, _, st = q.run()
    return t, y[0].mean(axis=2), st


def _deviation(field, t, mavg):
    ref = np.load(os.path.join(HERE, "golden", f"std_problem_4_field{field}.npz"))
    ref_i = np.stack([np.interp(t, ref["t"], ref["mean_avg"][:, k]) for k in range(3)], axis=1)
    return np.trapezoid(np.abs(mavg - ref_i), t, axis=0) / 1e-9, np.abs(mavg - ref_i).max(axis=0)


@pytest.mark.parametrize("field,bar", [(1, 0.1), (2, 0.2)])
def test_std_problem_4_matches_oracle_and_consensus(oracle, field, bar):
    t, mavg, m_s, stats = _run_std4(field)
    assert np.mean(m_s[:, 0]) > 0.9  # the relaxed s-state is mostly along +x
    assert np.abs(np.linalg.norm(mavg, axis=1)).max() <= 1.0 + 1e-9
    to, mo, st = _oracle_std4(oracle, field)
    diff = np.abs(mavg - mo).max()
    diff_avg = (np.trapezoid(np.abs(mavg - mo), t, axis=0) / 1e-9).max()
    dev, dmax = _deviation(field, t, mavg)
    print(f"field {field}: GPU vs oracle max |d<m>| {diff:.2e}, time-averaged {diff_avg:.2e}; time-averaged "
          f"deviation from muMAG {dev}, max {dmax}; gpu stats {stats}; oracle stats {st}")
    assert np.allclose(t, to)
    # When the 5000-evaluation work limit fires (flag 3, Q8) the run stores y at
    # t_got < t_want for one output time; GPU and oracle reach the limit a few
    # evaluations apart, so that output sample may differ.  Elsewhere the curves
    # agree to the integration tolerance.
    limit_hit = stats["flag3"] > 0 or st["flag3"] > 0
    assert diff < (2e-2 if limit_hit else 2e-3), diff
    assert diff_avg < 5e-4, diff_avg
    assert np.all(dev < bar), dev


def test_std_problem_4_fine_grid():
    """128 x 32 x 1 (3.9 nm cells), field 1: the GPU curve against the oracle's
    curve for the same problem (fixture from tests/golden/make_std4_oracle.py)
    and against the muMAG consensus, where the finer grid brings the
    time-averaged deviation down to the consensus spread (oracle: 0.0069,
    0.018, 0.0040)."""
    t, mavg, m_s, stats = _run_std4(1, res=(128, 32, 1))
    orc = np.load(os.path.join(HERE, "golden", "std_problem_4_oracle_128x32_field1.npz"))
    assert np.allclose(t, orc["t"])
    diff = np.abs(mavg - orc["mean"]).max()
    diff_avg = (np.trapezoid(np.abs(mavg - orc["mean"]), t, axis=0) / 1e-9).max()
    dev, dmax = _deviation(1, t, mavg)
    print(f"128x32 field 1: GPU vs oracle max |d<m>| {diff:.2e}, time-averaged {diff_avg:.2e}; deviation from "
          f"muMAG {dev}, max {dmax}; stats {stats}")
    # both runs reach the 5000-evaluation work limit (the oracle twice, flag 3):
    # the output samples stored at t_got < t_want may differ (see above)
    assert diff < (5e-2 if stats["flag3"] > 0 else 2e-3), diff
    assert diff_avg < 5e-4, diff_avg
    assert np.all(dev < 0.03), dev
