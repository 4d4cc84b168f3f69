"""Pins of the oracle's table lookup, interp1_MagTense / locate
(util.f90:39-85), which drives the applied field H_ext(t) of the dynamic
solver (LandauLifshitzEquationSolver.f90:485-514) and the damping table
alpha(t) (:281-294).

Independent references: numpy's np.interp and np.searchsorted on non-uniform
tables, and the reference's own special cases read from util.f90 -- a value
left of the table gives 0 (ind_x == 0 returns yval = 0), x == x(1) takes the
first interval, x == x(n) the last one (so it returns y(n) exactly).  Right of
the table the reference reads y(n+1) past the array end; the oracle (and the
GPU solver) clamp to y(n) instead, a documented deviation.
"""

import numpy as np
import pytest

MU0 = 4e-7 * np.pi


def _table(rng, n):
    x = np.sort(rng.uniform(-3e-9, 7e-9, size=n))
    x[0], x[-1] = -3e-9, 7e-9
    y = rng.normal(size=n) * 1e5
    return x, y


@pytest.mark.parametrize("n", [2, 3, 17, 200])
def test_interp1_matches_numpy_inside_the_table(oracle, n):
    rng = np.random.default_rng(n)
    x, y = _table(rng, n)
    q = rng.uniform(x[0], x[-1], size=400)
    scale = np.abs(y).max()
    for v in q:
        got, j = oracle.interp1(x, y, v)
        assert abs(got - np.interp(v, x, y)) <= 1e-13 * scale, (v, got)
        # locate's bracket: x(j) <= v < x(j+1) in 1-based indices
        assert j == np.searchsorted(x, v, side="right"), (v, j)


def test_interp1_knots_and_endpoints(oracle):
    rng = np.random.default_rng(5)
    x, y = _table(rng, 12)
    n = len(x)
    for k in range(1, n - 1):  # interior knots: the interval starting there, weight 0
        got, j = oracle.interp1(x, y, x[k])
        assert j == k + 1 and got == pytest.approx(y[k], rel=0, abs=1e-12 * abs(y).max())
    got, j = oracle.interp1(x, y, x[0])
    assert j == 1 and got == pytest.approx(y[0], rel=1e-15)
    got, j = oracle.interp1(x, y, x[-1])  # x == x(n): j = n - 1, weight exactly 1
    assert j == n - 1 and got == y[-1]


def test_interp1_outside_the_table(oracle):
    rng = np.random.default_rng(9)
    x, y = _table(rng, 7)
    for v in (x[0] - 1e-12, x[0] - 1.0, -np.inf):
        got, j = oracle.interp1(x, y, v)
        assert j == 0 and got == 0.0  # left of the table: yval stays 0 (np.interp would give y[0])
    got, j = oracle.interp1(x, y, x[-1] + 1e-12)
    assert j == len(x) and got == y[-1]  # clamped (the reference reads past the end)


def test_applied_field_ramp_through_dmdt(oracle):
    """H_ext(t) of a non-uniform ramp table, read back through dm/dt's stored
    (negated) applied field at times inside, on and left of the table."""
    rng = np.random.default_rng(3)
    res, L = [4, 3, 1], [40e-9, 30e-9, 3e-9]
    n = 12
    ht = np.sort(rng.uniform(0, 1e-9, size=9))
    ht[0] = 1e-11
    tab = np.zeros((len(ht), 4))
    tab[:, 0] = ht
    tab[:, 1:] = rng.normal(size=(len(ht), 3)) * 2e4
    m = rng.normal(size=(3, n))
    m /= np.linalg.norm(m, axis=0)
    q = oracle.Problem(res, L, gamma=2.21e5, alpha=4.42e3, hext=tab)
    times = list(rng.uniform(ht[0], ht[-1], size=20)) + list(ht) + [0.0, 5e-12]
    for t in times:
        _, H = q.dmdt(t, m.reshape(-1))
        hh = H[1].reshape(3, n)
        want = np.array([np.interp(t, ht, tab[:, 1 + c]) for c in range(3)]) if t >= ht[0] else np.zeros(3)
        assert np.allclose(hh, -want[:, None], rtol=1e-13, atol=1e-13 * np.abs(tab[:, 1:]).max()), t


def test_damping_table_through_dmdt(oracle):
    """alpha = 0 selects the alpha(t) table (LandauLifshitzEquationSolver.f90:281-294):
    with gamma = 0, dm/dt = -alpha(t) m x (m x H) scales linearly with the
    interpolated damping, so dm/dt at t equals the constant-alpha dm/dt times
    interp(t) / alpha_ref."""
    res, L = [4, 3, 1], [40e-9, 30e-9, 3e-9]
    n = 12
    rng = np.random.default_rng(4)
    m = rng.normal(size=(3, n))
    m /= np.linalg.norm(m, axis=0)
    ta = np.array([0.0, 2e-10, 3e-10, 9e-10])
    av = np.array([1e3, 5e3, 2e3, 8e3])
    hx = np.array([[0.0, 3e4, -1e4, 2e3], [1e-9, 3e4, -1e4, 2e3]])
    ref = oracle.Problem(res, L, gamma=0.0, alpha=1e3, hext=hx)
    tab = oracle.Problem(res, L, gamma=0.0, alpha=0.0, alphat=np.stack([ta, av], axis=1), hext=hx)
    for t in (1e-10, 2e-10, 2.5e-10, 6e-10, 9e-10):
        d_ref, _ = ref.dmdt(t, m.reshape(-1))
        d_tab, _ = tab.dmdt(t, m.reshape(-1))
        a = np.interp(t, ta, av)
        assert np.allclose(d_tab, d_ref * a / 1e3, rtol=1e-12, atol=1e-12 * np.abs(d_ref).max() * a / 1e3), t
