"""Oracle checks for the unstructured-prism micromagnetic grid
(gridTypeUnstructuredPrisms; LandauLifshitzEquationSolver.f90:1044-1133 tensor,
:1530-1544 loaded exchange operator).

* A uniform grid given as unstructured prisms (its cell centres and edges, and
  the uniform exchange operator loaded as a 1-based CSR) reproduces the
  uniform-grid right-hand side: the two paths share no tensor or exchange code.
* Receiver averaging (N_ave): every sub-point lies inside its cell, so the
  averaged self term of any prism has trace -1 (up to rounding); for a cube
  the sub-point lattice is symmetric under axis permutations and reflections,
  so the averaged diagonal is -1/3 and the off-diagonal entries vanish at every
  N_ave.  Far pairs: the tensor entries are harmonic outside the source and the
  sub-point lattice of a cube has isotropic second moments, so the average
  differs from the centre value only at O((a/r)^4).
"""

import numpy as np
import pytest

MU0 = 4 * np.pi * 1e-7


def _unit(rng, n):
    m = rng.normal(size=(n, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def uniform_as_unstructured(oracle, res, L):
    pts, d = oracle.setup_grid(res, L)
    n = pts.shape[0]
    rowptr, col, val = oracle.exchange_csr(res, L)
    rows_start = rowptr[:-1] + 1
    rows_end = rowptr[1:] + 1
    abc = np.tile(d, (n, 1))
    return pts, abc, (rows_start, rows_end, col + 1, val)


@pytest.mark.parametrize("res", [[8, 4, 1], [5, 3, 2]])
def test_uniform_grid_as_unstructured_prisms(oracle, res):
    L = [40e-9, 12e-9, 3e-9]
    pts, abc, exch = uniform_as_unstructured(oracle, res, L)
    n = pts.shape[0]
    rng = np.random.default_rng(11)
    m = _unit(rng, n)
    soa = np.concatenate([m[:, 0], m[:, 1], m[:, 2]])
    Ms = rng.uniform(7e5, 9e5, n)
    hext = np.array([[0.0, 1e4, -2e4, 5e3]])
    kw = dict(Ms=Ms, gamma=2.21e5, alpha=4.42e3, hext=hext, demag_fp64=True)
    uni = oracle.Problem(res, L, **kw)
    uns = oracle.Problem([n, 1, 1], L, grid_pts=pts, grid_abc=abc, exch=exch, **kw)
    d_u, H_u = uni.dmdt(0.0, soa)
    d_s, H_s = uns.dmdt(0.0, soa)
    for a, b in zip(H_u, H_s):
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(a).max(), 1.0)
    assert np.abs(d_u - d_s).max() <= 1e-12 * np.abs(d_u).max()


def _avg_self(oracle, abc, nave):
    pts = np.zeros((1, 3))
    exch = (np.array([1]), np.array([2]), np.array([1]), np.array([0.0]))
    q = oracle.Problem([1, 1, 1], [1, 1, 1], grid_pts=pts, grid_abc=np.array([abc]), exch=exch,
                       n_ave=nave, demag_fp64=True)
    m = np.array([1.0, 0.0, 0.0])
    _, H = q.dmdt(0.0, m)
    return H  # H_dem = -Ms * K m (negated field), SoA


def test_averaged_self_term(oracle):
    Ms = 8e5
    for n_a in (1, 2, 4, 8):
        # cube: K m for m = x gives the first column of the averaged tensor
        H = _avg_self(oracle, [5e-9, 5e-9, 5e-9], (n_a, n_a, n_a))[2]
        col = -H / Ms
        assert abs(col[0] + 1 / 3) < 1e-14
        assert abs(col[1]) < 1e-14 and abs(col[2]) < 1e-14
    # unequal sub-point counts break the cube's permutation symmetry: the
    # diagonal entries then differ from -1/3 (the averaging is not a no-op)
    H = _avg_self(oracle, [5e-9, 5e-9, 5e-9], (1, 1, 3))[2]
    assert abs(-H[0] / Ms + 1 / 3) > 1e-3
    # a flat prism: the trace of the averaged tensor is -1 at every N_ave
    for nave in ((1, 1, 1), (3, 2, 4), (6, 6, 2)):
        tr = 0.0
        for k, m in enumerate(np.eye(3)):
            pts = np.zeros((1, 3))
            exch = (np.array([1]), np.array([2]), np.array([1]), np.array([0.0]))
            q = oracle.Problem([1, 1, 1], [1, 1, 1], grid_pts=pts, grid_abc=np.array([[6e-9, 3e-9, 1e-9]]), exch=exch,
                               n_ave=nave, demag_fp64=True)
            _, H = q.dmdt(0.0, m)
            tr += -H[2][k] / Ms
        assert abs(tr + 1.0) < 1e-12, (nave, tr)


def test_averaged_far_pair(oracle):
    """Two cubes 20 edges apart: the receiver-averaged tensor stays within
    O((a/r)^4) = 6e-6 of the centre value (and differs from it), and equals it
    at N_ave = 1."""
    a = 2e-9
    pts = np.array([[0.0, 0.0, 0.0], [40e-9, 0.0, 0.0]])
    abc = np.full((2, 3), a)
    exch = (np.array([1, 1]), np.array([1, 1]), np.array([1]), np.array([0.0]))
    vals = {}
    for nave in ((1, 1, 1), (4, 4, 4)):
        q = oracle.Problem([2, 1, 1], [1, 1, 1], grid_pts=pts, grid_abc=abc, exch=exch, n_ave=nave, demag_fp64=True,
                           Ms=1.0)
        m = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.0])  # cell 0 along x, cell 1 zero
        _, H = q.dmdt(0.0, m)
        vals[nave] = -H[2][1]  # K_xx(receiver 1, source 0)
    Kc = oracle.tensor_points(pts[:1], abc[:1], pts[1:])[0, 0, 0]
    assert vals[(1, 1, 1)] == pytest.approx(Kc, rel=1e-12)
    assert 1e-9 * abs(Kc) < abs(vals[(4, 4, 4)] - Kc) < 1e-5 * abs(Kc)


def test_unstructured_tensor_entry_matches_uniform(oracle):
    """or_tensor_unstructured at N_ave = 1 on a uniform grid's cells equals the
    uniform-grid assembly (the evaluation point (p - a/2) + a/2 may differ from
    p in the last bit)."""
    res, L = [6, 3, 2], [30e-9, 9e-9, 4e-9]
    pts, d = oracle.setup_grid(res, L)
    Ku = oracle.tensor_uniform(res, L)
    Ks = oracle.tensor_unstructured(pts, np.tile(d, (pts.shape[0], 1)))
    assert np.abs(Ks - Ku).max() <= 1e-12 * np.abs(Ku).max()
    assert np.mean(Ks == Ku) > 0.5
