"""Whole-pipeline checks on muMAG standard problem 4, run exactly as the
reference example python/examples/micromagnetism/std_problem_4.py:8-46 does
(``from magtense.micromag import MicromagProblem``, the example's body with
show=False): s-state relaxation (100 ns, gamma = 0) from m = 1/sqrt(3), then
1 ns of dynamics in the NIST field.

* Parity: the GPU run against the FP64 oracle's run of the same problem at
  the example's resolution (64 x 16 x 1), in the reference's FP32 demag
  arithmetic and with the split-BF16 product: the averaged magnetization
  (the quantity the problem reports) within 1e-4 at every stored sample and
  every cell within 1e-4 after 1 ns, except the samples that RKSUITE's
  5000-evaluation work limit stored early (returned t_out < t_want, flag 3,
  Q8).  GPU and oracle integrate with different roundings, reach that limit
  a few evaluations apart and store y at different t_got there; the
  trajectory itself is not interrupted, so the samples after it agree again.
  The reversal amplifies rounding differences in single cells mid-run (FP32
  GPU vs oracle up to 1.2e-4), so the per-cell bar over all samples is 1e-3.
* Physics: against the muMAG consensus curves shipped with the reference
  (fixtures in tests/golden, from Published_solutions_field{1,2}.mat).  The
  reference's own comparison is visual (Standard_problem_4.m:185-205,
  Benchmark_results.m:79-83: the integral of |<m_i> - <m_i>_NIST| over time).
  At the example's 7.8 nm cells the discretisation dominates: the oracle
  itself deviates by (0.033, 0.078, 0.017) for field 1 and (0.072, 0.170,
  0.045) for field 2 (time-averaged), so the bars there are 0.1 and 0.2.
"""

import functools
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MU0 = 4 * np.pi * 1e-7
GRID_L = [500e-9, 125e-9, 3e-9]


def _nist(field):
    return np.array([-24.6, 4.3, 0]) if field == 1 else np.array([-35.5, -6.3, 0])


def std_prob_4(res=[64, 16, 1], NIST_field=1, cuda=False, precision=False):
    """Body of std_problem_4.py:8-46 (show=False) plus the additive
    ``precision`` keyword; returns the dynamic run's outputs."""
    from magtense.micromag import MicromagProblem

    mu0 = 4 * np.pi * 1e-7
    grid_L = [500e-9, 125e-9, 3e-9]

    ### Magnetization to s-state
    problem_ini = MicromagProblem(res, grid_L, m0=1 / np.sqrt(3), alpha=4.42e3, cuda=cuda, precision=precision)
    h_ext = np.array([1, 1, 1]) / mu0

    def h_ext_fct(t):
        return np.expand_dims(np.where(t < 1e-09, 1e-09 - t, 0), axis=1) * h_ext

    _, M_out, _, _, _, _, _ = problem_ini.run_simulation(100e-9, 200, h_ext_fct, 2000)
    M_sq_ini = np.squeeze(M_out, axis=2)

    ### Time-dependent solver
    problem_dym = MicromagProblem(res, grid_L, m0=M_sq_ini[-1], alpha=4.42e3, gamma=2.21e5, cuda=cuda,
                                  precision=precision)

    # Two applied external fields of std problem 4
    if NIST_field == 1:
        h_ext_nist = np.array([-24.6, 4.3, 0])
    elif NIST_field == 2:
        h_ext_nist = np.array([-35.5, -6.3, 0])
    else:
        raise NotImplementedError()

    def h_ext_fct(t):
        return np.expand_dims(t > -1, axis=1) * (h_ext_nist / 1000 / mu0)

    t_dym, M_out, _, _, _, _, _ = problem_dym.run_simulation(1e-9, 200, h_ext_fct, 2000)
    return t_dym, np.squeeze(M_out, axis=2), M_sq_ini[-1], problem_dym.last_stats[0]


@functools.lru_cache(maxsize=None)
def _oracle_std4(field, res=(64, 16, 1)):
    """The same two runs with the FP64 oracle (FP32 demag semantics):
    (t_out, M (nt, N, 3), stats)."""
    from oracle import oracle

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
    t_out, y, _, st = q.run()
    return t_out[0], np.transpose(y[0], (0, 2, 1)), st


def _deviation(field, t, mavg):
    ref = np.load(os.path.join(HERE, "golden", f"std_problem_4_field{field}.npz"))
    ref_i = np.stack([np.interp(t, ref["t"], ref["mean_avg"][:, k]) for k in range(3)], axis=1)
    return np.trapezoid(np.abs(mavg - ref_i), t, axis=0) / 1e-9, np.abs(mavg - ref_i).max(axis=0)


def _early(t_out, t_want):
    """Samples stored before their output time (work-limit returns)."""
    return t_out < t_want * (1 - 1e-12)


@pytest.mark.parametrize("field,bar,precision", [(1, 0.1, "fp32"), (2, 0.2, "fp32"), (1, 0.1, "bf16x2")])
def test_std_problem_4_matches_oracle_and_consensus(oracle, field, bar, precision):
    t, M, m_s, stats = std_prob_4([64, 16, 1], field, precision=precision)
    assert np.mean(m_s[:, 0]) > 0.9  # the relaxed s-state is mostly along +x
    mavg = M.mean(axis=1)
    assert np.abs(np.linalg.norm(mavg, axis=1)).max() <= 1.0 + 1e-9
    to, Mo, st = _oracle_std4(field)
    t_want = np.linspace(0, 1e-9, 200)
    early = _early(t, t_want) | _early(to, t_want)
    d_avg = np.abs(mavg - Mo.mean(axis=1)).max(axis=1)
    d_cell = np.abs(M - Mo).max(axis=(1, 2))
    dev, dmax = _deviation(field, t, mavg)
    print(f"field {field} {precision}: GPU vs oracle max |d<m>| {d_avg[~early].max():.2e}, per cell "
          f"{d_cell[~early].max():.2e} over {int((~early).sum())} samples ({int(early.sum())} stored early: "
          f"{d_avg[early]}); final |d<m>| {d_avg[-1]:.2e}, per cell {d_cell[-1]:.2e}; deviation from muMAG {dev}, "
          f"max {dmax}; gpu stats {stats}; oracle stats {st}")
    # an early sample exists only where a run reported a work-limit return
    assert int(early.sum()) <= stats["flag3"] + st["flag3"]
    assert d_avg[~early].max() < 1e-4, d_avg[~early].max()
    assert not early[-1] and d_cell[-1] < 1e-4, d_cell[-1]
    assert d_cell[~early].max() < 1e-3, d_cell[~early].max()
    assert np.all(dev < bar), dev


def test_std_problem_4_fine_grid():
    """128 x 32 x 1 (3.9 nm cells), field 1: the GPU curve against the oracle's
    curve for the same problem (fixture from tests/golden/make_std4_oracle.py,
    which stores the averaged magnetization and the oracle's t_out) and
    against the muMAG consensus, where the finer grid brings the time-averaged
    deviation down to the consensus spread (oracle: 0.0069, 0.018, 0.0040)."""
    t, M, m_s, stats = std_prob_4([128, 32, 1], 1)
    mavg = M.mean(axis=1)
    orc = np.load(os.path.join(HERE, "golden", "std_problem_4_oracle_128x32_field1.npz"))
    t_want = np.linspace(0, 1e-9, 200)
    early = _early(t, t_want) | _early(orc["t"], t_want)
    d_avg = np.abs(mavg - orc["mean"]).max(axis=1)
    dev, dmax = _deviation(1, t, mavg)
    print(f"128x32 field 1: GPU vs oracle max |d<m>| {d_avg[~early].max():.2e} over {int((~early).sum())} samples "
          f"({int(early.sum())} stored early: {d_avg[early]}); deviation from muMAG {dev}, max {dmax}; "
          f"stats {stats}")
    assert d_avg[~early].max() < 1e-4, d_avg[~early].max()
    assert np.all(dev < 0.03), dev
