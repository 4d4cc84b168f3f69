"""CPU pins of the FP64 oracle (no GPU).

Every check here is independent of the oracle's own formulas: closed forms,
invariants, symmetries, brute-force quadrature and known-answer ODE results,
chosen so that a dropped term, a wrong sign or a transposed operand fails.
"""

import numpy as np
import pytest
from scipy import integrate


# ---------------------------------------------------------------------------
# prism tensor (TileNComponents.f90:629-724)
# ---------------------------------------------------------------------------
def K6to33(k):
    return np.array([[k[0], k[1], k[2]], [k[1], k[3], k[4]], [k[2], k[4], k[5]]])


def test_cube_self_term(oracle):
    k = oracle.prism_k6([1.0, 1.0, 1.0], [0.0, 0.0, 0.0])
    assert np.allclose(K6to33(k), -np.eye(3) / 3, atol=1e-15)


def test_centre_closed_form_noncube(oracle):
    # N_zz at the centre of an a x b x c prism = (2/pi) atan(ab / (c sqrt(a^2+b^2+c^2)))
    # (solid angle of the two z faces); K = -N.
    for a, b, c in [(2.0, 1.0, 0.5), (5.0, 5.0, 2.0), (1.0, 3.0, 7.0)]:
        k = oracle.prism_k6([a, b, c], [0, 0, 0])
        d = np.sqrt(a * a + b * b + c * c)
        nxx = 2 / np.pi * np.arctan(b * c / (a * d))
        nyy = 2 / np.pi * np.arctan(a * c / (b * d))
        nzz = 2 / np.pi * np.arctan(a * b / (c * d))
        assert np.allclose([k[0], k[3], k[5]], [-nxx, -nyy, -nzz], rtol=1e-13)
        assert abs(nxx + nyy + nzz - 1) < 1e-13


def test_trace_inside_and_outside(oracle):
    rng = np.random.default_rng(0)
    for _ in range(50):
        abc = rng.uniform(0.5, 3.0, 3)
        inside = rng.uniform(-0.45, 0.45, 3) * abc
        outside = inside + np.sign(rng.normal(size=3)) * abc * rng.uniform(0.6, 3.0, 3)
        ki = oracle.prism_k6(abc, inside)
        ko = oracle.prism_k6(abc, outside)
        assert abs(ki[0] + ki[3] + ki[5] + 1.0) < 1e-12  # tr N = 1 inside
        assert abs(ko[0] + ko[3] + ko[5]) < 1e-12         # tr N = 0 outside


def test_inversion_and_permutation_symmetry(oracle):
    rng = np.random.default_rng(1)
    for _ in range(20):
        abc = rng.uniform(0.5, 3.0, 3)
        r = rng.uniform(-4, 4, 3)
        k = K6to33(oracle.prism_k6(abc, r))
        assert np.allclose(k, K6to33(oracle.prism_k6(abc, -r)), rtol=1e-12, atol=1e-15)
        # mirror y -> -y flips the xy and yz components only
        km = K6to33(oracle.prism_k6(abc, r * np.array([1, -1, 1])))
        S = np.diag([1, -1, 1])
        assert np.allclose(km, S @ k @ S, rtol=1e-11, atol=1e-14)
        # axis permutation (x,y,z) -> (y,z,x) with edges permuted alike
        P = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]])
        kp = K6to33(oracle.prism_k6(P @ abc, P @ r))
        assert np.allclose(kp, P @ k @ P.T, rtol=1e-11, atol=1e-14)


def test_far_field_dipole_limit(oracle):
    abc = np.array([2.0, 1.5, 1.0])
    V = abc.prod()
    rng = np.random.default_rng(2)
    for _ in range(10):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = 60.0 * d
        k = K6to33(oracle.prism_k6(abc, r))
        dip = V / (4 * np.pi * 60.0**3) * (3 * np.outer(d, d) - np.eye(3))
        assert np.abs(k - dip).max() <= 2e-3 * np.abs(dip).max()


def _quadrature_K(abc, r):
    """K_ab = (1/4pi) sum_{faces normal to b} n_b * int (r - r')_a / |r - r'|^3 dS' (surface charges)."""
    a2 = np.asarray(abc) / 2
    K = np.zeros((3, 3))
    for bdir in range(3):
        u, v = [d for d in range(3) if d != bdir]
        for sgn in (-1.0, 1.0):
            for adir in range(3):
                def f(y, x):
                    rp = np.zeros(3)
                    rp[bdir] = sgn * a2[bdir]
                    rp[u] = x
                    rp[v] = y
                    d = r - rp
                    return d[adir] / np.linalg.norm(d) ** 3

                val, _ = integrate.dblquad(f, -a2[u], a2[u], -a2[v], a2[v], epsabs=1e-13, epsrel=1e-11)
                K[adir, bdir] += sgn * val / (4 * np.pi)
    return K


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_surface_quadrature(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    abc = rng.uniform(0.8, 2.0, 3)
    r = rng.uniform(-2.5, 2.5, 3)
    r = np.where(np.abs(r) < abc / 2 + 0.2, r + np.sign(r) * (abc / 2 + 0.3), r)  # keep off the surface
    k = K6to33(oracle.prism_k6(abc, r))
    kq = _quadrature_K(abc, r)
    assert np.abs(k - kq).max() <= 1e-8 * max(np.abs(kq).max(), 1e-3)


def test_guard_points_are_finite_and_continuous(oracle):
    abc = np.array([2.0, 1.0, 0.5])
    # on a face plane outside the face and on the axes: the guarded value is
    # finite and continuous with its neighbourhood
    for r in ([1.0, 0.7, 0.1], [3.0, 0.0, 0.0], [0.0, 0.0, 4.0]):
        r = np.array(r)
        k = oracle.prism_k6(abc, r)
        assert np.all(np.isfinite(k))
        kn = oracle.prism_k6(abc, r * (1 + 1e-7))
        assert np.abs(k - kn).max() < 1e-3
    # on edge lines (log-singular off-diagonals) the limit rule still gives finite values
    for r in ([1.0, 0.5, 0.3], [1.0, 0.5, 0.25], [1.0, 0.5, 0.0]):
        assert np.all(np.isfinite(oracle.prism_k6(abc, np.array(r))))
    # on the face itself the field jumps by the surface charge (N_xx by 1);
    # the guard returns the mean of the two one-sided limits
    r = np.array([1.0, 0.2, 0.1])
    k = oracle.prism_k6(abc, r)
    kin = oracle.prism_k6(abc, r * np.array([1 - 1e-9, 1, 1]))
    kout = oracle.prism_k6(abc, r * np.array([1 + 1e-9, 1, 1]))
    assert abs((kout[0] - kin[0]) - 1.0) < 1e-6
    assert abs(k[0] - 0.5 * (kin[0] + kout[0])) < 1e-3


def test_uniform_tensor_layout_and_symmetry(oracle):
    res, L = [4, 3, 2], [8.0, 6.0, 2.0]
    K = oracle.tensor_uniform(res, L)
    pts, d = oracle.setup_grid(res, L)
    n = pts.shape[0]
    # receiver-major: K[c, i, j] = tensor at receiver i due to source j
    i, j = 5, 17
    k = oracle.prism_k6(d, pts[i] - pts[j])
    assert np.allclose(K[:, i, j], k)
    for c in range(6):
        assert np.allclose(K[c], K[c].T, rtol=1e-10, atol=1e-14)  # identical prisms: K_ij = K_ji
    # diagonal self term of a 2x2x1 cell
    assert np.allclose(K[:, 0, 0], oracle.prism_k6(d, [0, 0, 0]))
    assert n == 24


def test_setup_grid_reference_formula(oracle):
    pts, d = oracle.setup_grid([4, 2, 1], [8.0, 2.0, 3.0])
    assert np.allclose(d, [2.0, 1.0, 3.0])
    assert np.allclose(pts[:4, 0], [-3, -1, 1, 3]) and np.allclose(pts[:, 2], 0.0)
    assert np.allclose(pts[4:, 1], 0.5) and np.allclose(pts[:4, 1], -0.5)


# ---------------------------------------------------------------------------
# exchange operator (LandauLifshitzEquationSolver.f90:1194-1522)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("res", [[5, 4, 1], [3, 3, 3], [2, 2, 2]])
def test_exchange_operator_structure(oracle, res):
    L = [res[0] * 1.0, res[1] * 2.0, res[2] * 0.5]
    rowptr, col, val = oracle.exchange_csr(res, L)
    n = int(np.prod(res))
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    assert np.allclose(A.sum(axis=1), 0.0)          # Neumann rows: constants in the kernel
    assert np.allclose(A, A.T)                       # symmetric
    assert np.all(np.linalg.eigvalsh(A) <= 1e-9)     # negative semi-definite
    # a linear field along x has zero Laplacian in the interior
    pts, _ = oracle.setup_grid(res, L)
    lin = A @ pts[:, 0]
    nx = res[0]
    interior = [r for r in range(n) if 0 < r % nx < nx - 1]
    assert np.allclose(lin[interior], 0.0, atol=1e-12)
    if res[2] == 1:  # no z coupling for a single layer
        assert all(abs(c - r) < res[0] * res[1] for r in range(n) for c in col[rowptr[r]:rowptr[r + 1]])


def test_exchange_needs_two_cells():
    from oracle import oracle as O

    with pytest.raises(ValueError):
        O.exchange_csr([1, 4, 1], [1.0, 1.0, 1.0])


# ---------------------------------------------------------------------------
# dm/dt (LandauLifshitzEquationSolver.f90:220-273)
# ---------------------------------------------------------------------------
def _prob(oracle, res=(6, 4, 2), gamma=2.21e5, alpha=4.42e3, K0=0.0, u=None, h=(1e4, -2e4, 5e3)):
    L = [res[0] * 5e-9, res[1] * 5e-9, res[2] * 3e-9]
    hext = np.array([[0.0, *h], [1.0, *h]])
    return oracle.Problem(list(res), L, gamma=gamma, alpha=alpha, K0=K0, u_ea=u, hext=hext)


def _rand_m(n, seed):
    m = np.random.default_rng(seed).normal(size=(3, n))
    return m / np.linalg.norm(m, axis=0, keepdims=True)


def test_dmdt_is_orthogonal_to_m(oracle):
    q = _prob(oracle)
    m = _rand_m(q.n, 0)
    d, _ = q.dmdt(0.0, m.reshape(-1))
    d = d.reshape(3, -1)
    dot = np.abs((m * d).sum(axis=0))
    assert dot.max() <= 1e-10 * np.abs(d).max()


def test_damping_only_decreases_energy(oracle):
    # gamma = 0: dm/dt = -alpha m x (m x H') with H' = -H, so dm/dt . H >= 0
    n_u = 48
    u = np.tile([0.0, 0.0, 1.0], (n_u, 1))
    q = _prob(oracle, gamma=0.0, K0=3e4, u=u)
    m = _rand_m(q.n, 1)
    d, H = q.dmdt(0.0, m.reshape(-1))
    Hphys = -(H[0] + H[1] + H[2] + H[3])
    assert (d * Hphys).reshape(3, -1).sum(axis=0).min() >= -1e-9 * np.abs(d * Hphys).max()


def test_precession_only_conserves_energy_density(oracle):
    q = _prob(oracle, alpha=0.0)
    q2 = oracle.Problem(q.res, q.L, gamma=2.21e5, alpha=0.0, hext=np.array([[0.0, 1e4, -2e4, 5e3], [1.0, 1e4, -2e4, 5e3]]),
                        alphat=np.array([[0.0, 0.0]]))
    m = _rand_m(q2.n, 2)
    d, H = q2.dmdt(0.0, m.reshape(-1))
    Hphys = -(H[0] + H[1] + H[2] + H[3])
    # pure precession: dm/dt is orthogonal to the local field
    assert np.abs((d * Hphys).reshape(3, -1).sum(axis=0)).max() <= 1e-9 * np.abs(d).max() * np.abs(Hphys).max()


def test_field_terms_closed_forms(oracle):
    u = np.tile([1.0, 0.0, 0.0], (48, 1))
    q = _prob(oracle, K0=2e4, u=u, h=(3e3, 0.0, -1e3))
    m = np.zeros((3, q.n))
    m[0] = 1.0  # uniform along x
    _, H = q.dmdt(0.0, m.reshape(-1))
    Hexc, Hext, Hdem, Hani = (h.reshape(3, -1) for h in H)
    assert np.abs(Hexc).max() < 1e-6                                   # uniform m: no exchange
    assert np.allclose(Hext[0], -3e3) and np.allclose(Hext[2], 1e3)    # stored negated
    mu0 = 4e-7 * np.pi
    assert np.allclose(Hani[0], -2 * 2e4 / (mu0 * 8e5))                # -2 K0/(mu0 Ms) (u.m) u
    assert np.allclose(Hani[1:], 0.0)
    # demag of a uniformly magnetized film: mostly along -x physically -> stored positive-ish mean
    assert Hdem[0].mean() > 0


# ---------------------------------------------------------------------------
# RKSUITE 'M' (RKSuitef90.f90)
# ---------------------------------------------------------------------------
def test_bs45_order_conditions(oracle):
    a, b, c, e, r = oracle.rk_tables()
    A = a[:8, :8]
    cc = c[:8]
    w5 = A[7, :7]                     # 5th-order weights (FSAL row)
    w4 = b                            # 4th-order weights (stages 1..8)
    assert np.allclose(A.sum(axis=1), cc, atol=1e-15)
    for k in range(1, 6):             # bushy-tree conditions sum w c^(k-1) = 1/k
        assert abs((w5 * cc[:7] ** (k - 1)).sum() - 1.0 / k) < 1e-14
    for k in range(1, 5):
        assert abs((w4 * cc ** (k - 1)).sum() - 1.0 / k) < 1e-14
    assert abs(w5 @ A[:7, :7] @ cc[:7] - 1 / 6) < 1e-14
    assert abs(w5 @ A[:7, :7] @ (cc[:7] ** 2) - 1 / 12) < 1e-14
    assert abs(w5 @ A[:7, :7] @ A[:7, :7] @ cc[:7] - 1 / 24) < 1e-14
    # e = 5th-order minus the first 4th-order formula: zero on constants
    assert abs(e.sum()) < 1e-15
    # the interpolant at sigma = 1 reproduces the step (r rows sum to the FSAL weights)
    assert np.allclose(c[8:11], [0.5, 5 / 6, 1 / 9])


def test_rk_exponential_decay_accuracy(oracle):
    t = np.linspace(0.0, 2.0, 21)
    to, y, st = oracle.rk_integrate(lambda t, y: -y, t, [1.0], tol=1e-8, thres=1e-10)
    assert np.abs(y[:, 0] - np.exp(-t)).max() < 2e-8
    assert np.allclose(to, t)
    assert st["flag3"] == 0


def test_rk_convergence_with_tolerance(oracle):
    def f(t, y):
        return np.array([y[1], -y[0]])

    t = np.linspace(0, 10, 51)
    errs = []
    for tol in (1e-4, 1e-6, 1e-8):
        _, y, _ = oracle.rk_integrate(f, t, [1.0, 0.0], tol=tol, thres=1e-10)
        errs.append(np.abs(y[:, 0] - np.cos(t)).max())
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-6


def test_rk_interpolant_dense_output(oracle):
    # y' = cos t: outputs between steps come from the degree-6 interpolant
    t = np.linspace(0, 6, 301)
    _, y, st = oracle.rk_integrate(lambda t, y: np.array([np.cos(t)]), t, [0.0], tol=1e-9, thres=1e-10)
    assert np.abs(y[:, 0] - np.sin(t)).max() < 1e-8
    assert st["steps"] < 300  # many outputs per step


def test_rk_work_limit_returns_early_output(oracle):
    # >5000 RHS calls: step_integrate returns flag 3 and the driver stores y at
    # t_got < t_want (RKSuitef90.f90:1374-1389, ODE_Solvers.f90:202).
    w = 400.0

    def f(t, y):
        return np.array([w * y[1], -w * y[0]])

    t = np.linspace(0, 20, 41)
    to, y, st = oracle.rk_integrate(f, t, [1.0, 0.0], tol=1e-9, thres=1e-10)
    assert st["flag3"] >= 1
    early = np.where(np.abs(to - t) > 0)[0]
    assert len(early) >= 1
    k = early[0]
    assert to[k] < t[k]
    assert abs(y[k, 0] - np.cos(w * to[k])) < 1e-5


def test_rk_rejects_bad_tolerance(oracle):
    with pytest.raises(RuntimeError):
        oracle.rk_integrate(lambda t, y: -y, np.linspace(0, 1, 3), [1.0], tol=0.05)
