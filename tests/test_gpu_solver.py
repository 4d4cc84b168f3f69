"""GPU parity of the batched solver (drop-in Python API -> C ABI) against the oracle.

The oracle integrates the same problem with the serial RKSUITE 'M' restatement
in FP64 (FP32 demag semantics).  The GPU path differs only in FP32 summation
order and ulp-level libm differences, so trajectories agree to the integration
tolerance; the bar is 1e-4 absolute on unit magnetization components.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _random_m(n, seed):
    rng = np.random.default_rng(seed)
    m = rng.normal(size=(n, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def _oracle_run(oracle, res, L, m0, hfield, t_end, nt, nt_h, **kw):
    t = np.linspace(0, t_end, nt)
    hext = np.zeros((nt_h, 4))
    hext[:, 0] = np.linspace(0, t_end, nt_h)
    hext[:, 1:] = hfield
    q = oracle.Problem(res, L, hext=hext, t=t, m0=np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]]), **kw)
    t_out, y, H, st = q.run()
    return t_out[0], np.transpose(y[0], (0, 2, 1)), [np.transpose(h[0], (0, 2, 1)) for h in H], st


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_run_simulation_matches_oracle(oracle, precision):
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    m0 = _random_m(64, 1)
    hfield = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0
    p = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6, precision=precision)
    t_out, M, pts, Hexc, Hext, Hdem, Hani = p.run_simulation(1e-10, 11, lambda t: np.ones((len(t), 1)) * hfield, 20)
    tr, yr, Hr, st = _oracle_run(oracle, res, L, m0, hfield, 1e-10, 11, 20, alpha=4.42e3, gamma=2.21e5, tol=1e-6)
    assert M.shape == (11, 64, 1, 3) and M.flags.f_contiguous
    assert np.allclose(t_out, tr)
    err = np.abs(M[:, :, 0, :] - yr).max()
    print(precision, "max traj err", err, "gpu stats", p.last_stats[0], "oracle", st)
    assert err < 1e-4
    # Returned fields (negated, as the reference stores them) are linear in the
    # stored m, so the trajectory difference bounds theirs: exchange by
    # 2 Jfact ||A||_inf max|dm| (A the Neumann Laplacian), demag by
    # Ms max_i sum_b sum_j |K_ab(i, j)| max|dm| plus the product's own rounding
    # (FP32: 1e-6, split-BF16: 1e-5 of the largest field).
    ms, a0 = 8e5, 1.3e-11
    rowptr, _, val = oracle.exchange_csr(res, L)
    a_inf = max(np.abs(val[rowptr[i]:rowptr[i + 1]]).sum() for i in range(64))
    bar_exc = 2 * a0 / (MU0 * ms) * a_inf * err + 1e-12 * np.abs(Hr[0]).max()
    K = oracle.tensor_uniform(res, L)
    comp = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    k_inf = max(sum(np.abs(K[comp[a][b]]).sum(axis=1) for b in range(3)).max() for a in range(3))
    bar_dem = ms * k_inf * err + (1e-5 if precision == "bf16x2" else 1e-6) * np.abs(Hr[2]).max()
    d_exc = np.abs(Hexc[:, :, 0, :] - Hr[0]).max()
    d_dem = np.abs(Hdem[:, :, 0, :] - Hr[2]).max()
    print(f"fields: exchange {d_exc:.3e} (bar {bar_exc:.3e}), demag {d_dem:.3e} (bar {bar_dem:.3e})")
    assert d_exc <= bar_exc
    assert np.allclose(Hext[:, :, 0, :], Hr[1])
    assert d_dem <= bar_dem
    # step sequence: same algorithm, ulp-level differences only
    if precision == "fp32":
        assert abs(p.last_stats[0]["steps"] - st["steps"]) <= max(2, 0.05 * st["steps"])


def test_batch_equals_individual_runs(oracle):
    from magtense.micromag import MicromagProblem

    res, L = [8, 4, 1], [40e-9, 20e-9, 3e-9]
    n = 32
    B = 3
    m0 = np.stack([_random_m(n, 10 + b) for b in range(B)])
    fields = np.array([[-20e3, 5e3, 0.0], [10e3, 0.0, 2e3], [0.0, -15e3, 0.0]])
    p = MicromagProblem(res, L, alpha=4.42e3, gamma=2.21e5, tol=1e-6)
    t_b, M_b = p.run_simulation_batch(5e-11, 6, np.repeat(fields[:, None, :], 10, axis=1), 10, m0)
    for b in range(B):
        tr, yr, _, _ = _oracle_run(oracle, res, L, m0[b], fields[b], 5e-11, 6, 10, alpha=4.42e3, gamma=2.21e5,
                                   tol=1e-6)
        assert np.abs(M_b[b] - yr).max() < 1e-4


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_batch_beyond_one_demag_group(oracle, precision):
    """130 states: every tick runs three demag passes (64 + 64 + 2 columns).  Each
    state's column of the product, its LL kernel row and its step control are
    independent of the others, so a state in the batch reproduces its own
    one-state run bit for bit, on both sides of the split; in FP32 it also
    follows the oracle run (split-BF16 against the oracle: the single-run tests
    above, from starts whose trajectories do not amplify the 4e-6 field error)."""
    from magtense.micromag import MicromagProblem

    res, L = [8, 4, 1], [40e-9, 20e-9, 3e-9]
    n, B = 32, 130
    m0 = np.stack([_random_m(n, 200 + b) for b in range(B)])
    ang = np.linspace(0, 2 * np.pi, B, endpoint=False)
    fields = np.stack([2e4 * np.cos(ang), 2e4 * np.sin(ang), np.zeros(B)], axis=1)
    rows = np.repeat(fields[:, None, :], 10, axis=1)
    p = MicromagProblem(res, L, alpha=4.42e3, gamma=2.21e5, tol=1e-6, precision=precision)
    _, M_b = p.run_simulation_batch(5e-11, 6, rows, 10, m0)
    assert M_b.shape == (B, 6, n, 3)
    for b in (0, 63, 64, 128, 129):
        _, M_1 = p.run_simulation_batch(5e-11, 6, rows[b:b + 1], 10, m0[b:b + 1])
        assert np.array_equal(M_b[b], M_1[0]), (precision, b, np.abs(M_b[b] - M_1[0]).max())
        if precision == "fp32":
            _, yr, _, _ = _oracle_run(oracle, res, L, m0[b], fields[b], 5e-11, 6, 10, alpha=4.42e3,
                                      gamma=2.21e5, tol=1e-6)
            err = np.abs(M_b[b] - yr).max()
            print(f"state {b}: max |m - oracle| {err:.2e}")
            assert err < 1e-4, (b, err)


def test_explicit_solver_rows(oracle):
    """solver='explicit': each field row relaxes from the previous row's final state."""
    from magtense.micromag import MicromagProblem

    res, L = [8, 4, 1], [40e-9, 20e-9, 3e-9]
    m0 = _random_m(32, 4)
    p = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=0.0, tol=1e-5, solver="explicit")

    def fh(t):
        return np.stack([np.linspace(-5e4, 5e4, len(t)), np.zeros(len(t)), np.zeros(len(t))], axis=1)

    t_out, M, pts, *_ = p.run_simulation(2e-11, 4, fh, 3)
    assert M.shape == (4, 32, 3, 3)
    t = np.linspace(0, 2e-11, 4)
    hext = np.zeros((3, 4))
    hext[:, 0] = np.linspace(0, 2e-11, 3)
    hext[:, 1:] = fh(hext[:, 0])
    q = oracle.Problem(res, L, alpha=4.42e3, gamma=0.0, tol=1e-5, solver=1, hext=hext, t=t,
                       m0=np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]]))
    _, y, _, _ = q.run()
    ref = np.transpose(y, (1, 3, 0, 2))  # (nt, N, n_h, 3)
    assert np.abs(M - ref).max() < 1e-4


def test_convergence_early_exit_matches_oracle(oracle):
    """t_conv / conv_tol (MagTense_ODE_RKSuite, ODE_Solvers.f90:180-233): at
    every t_conv time the largest change since the previous one is compared
    with conv_tol; below it the state is stored once more and the run ends.
    A damping-only relaxation exits at the same output sample as the oracle,
    with the same trajectory up to it and zero rows after it."""
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    n = 64
    rng = np.random.default_rng(8)
    m0 = np.tile([0.95, 0.3, 0.0], (n, 1)) + 0.05 * rng.normal(size=(n, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    h = np.array([200.0, 0.0, 0.0]) / 1000 / 4e-7 / np.pi
    t_end, nt = 2e-9, 40
    t = np.linspace(0, t_end, nt)
    # the largest change per output interval falls by ~6 % per interval (oracle:
    # 0.010 at sample 13, 0.009 at 14): conv_tol 0.0095 ends the run mid-way,
    # 5e-4 away from the neighbouring values
    p = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=0.0, tol=1e-6, nt_conv=nt, conv_tol=0.0095)
    p.t_conv[:] = t
    _, M, *_ = p.run_simulation(t_end, nt, lambda tt: np.ones((len(tt), 1)) * h, 10)
    gpu = np.squeeze(M, axis=2)
    hext = np.zeros((10, 4))
    hext[:, 0] = np.linspace(0, t_end, 10)
    hext[:, 1:] = h
    soa = np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]])
    q = oracle.Problem(res, L, alpha=4.42e3, gamma=0.0, tol=1e-6, hext=hext, t=t, m0=soa, t_conv=t, conv_tol=0.0095,
                       return_fields=False)
    _, y, _, _ = q.run()
    orc = np.transpose(y[0], (0, 2, 1))
    stored_g = np.nonzero(np.abs(gpu).max(axis=(1, 2)) > 0)[0]
    stored_o = np.nonzero(np.abs(orc).max(axis=(1, 2)) > 0)[0]
    print(f"early exit: GPU stored {len(stored_g)} of {nt} samples, oracle {len(stored_o)}; "
          f"max |dm| {np.abs(gpu - orc).max():.2e}")
    assert len(stored_g) < nt  # the run ended early
    assert np.array_equal(stored_g, stored_o)
    assert np.abs(gpu - orc).max() < 1e-4


def test_time_dependent_alpha_matches_oracle(oracle):
    """alpha = 0 selects the damping table alphat (t, alpha(t)) interpolated at
    every right-hand side (LandauLifshitzEquationSolver.f90:281-294, Q12): a
    damping ramp from 0 to 8.84e3 over 0.2 ns, GPU against the oracle."""
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    # a smooth start (from random unit vectors the exchange-driven dynamics
    # amplify rounding differences towards the 1e-4 bar)
    rng = np.random.default_rng(12)
    m0 = np.tile([0.3, 0.9, 0.3], (64, 1)) + 0.05 * rng.normal(size=(64, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    hfield = np.array([30.0, -10.0, 0.0]) / 1000 / MU0
    t_alpha = np.linspace(0, 2e-10, 9)

    def ramp(t):
        return 8.84e3 * np.asarray(t) / 2e-10

    p = MicromagProblem(res, L, m0=m0, alpha=0.0, gamma=2.21e5, tol=1e-6, t_alpha=t_alpha, alpha_fct=ramp)
    assert p.alphat.shape == (9, 2) and np.allclose(p.alphat[:, 1], ramp(t_alpha))
    _, M, *_ = p.run_simulation(2e-10, 11, lambda t: np.ones((len(t), 1)) * hfield, 10)
    tr, yr, _, st = _oracle_run(oracle, res, L, m0, hfield, 2e-10, 11, 10, alpha=0.0, gamma=2.21e5, tol=1e-6,
                                alphat=np.stack([t_alpha, ramp(t_alpha)], axis=1))
    err = np.abs(M[:, :, 0, :] - yr).max()
    # the same run with a constant damping differs: the table is used
    p2 = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6)
    _, M2, *_ = p2.run_simulation(2e-10, 11, lambda t: np.ones((len(t), 1)) * hfield, 10)
    print(f"alpha(t): GPU vs oracle {err:.2e}; vs constant alpha {np.abs(M - M2).max():.2e}; gpu {p.last_stats[0]}; "
          f"oracle {st}")
    assert err < 1e-4
    assert np.abs(M - M2).max() > 1e-2


@pytest.mark.parametrize("res", [[32, 8, 1], [9, 7, 2]])
def test_stage_argument_in_demag_launch_same_bits(monkeypatch, res):
    """Small FP32 dense problems form each stage argument inside the demag + LL
    launch (every workgroup evaluates the whole combination into LDS and writes
    its own cells of it); MT_STAGE_FUSED=0 runs the vector kernel first.  Same
    arithmetic in the same order: trajectories and returned fields agree bit
    for bit, for one state and for a batch of three (the 9x7x2 grid takes the
    scalar row loads: 126 cells is not a multiple of four)."""
    from magtense.micromag import MicromagProblem

    n = int(np.prod(res))
    L = [res[0] * 15.625e-9, res[1] * 15.625e-9, res[2] * 3e-9]
    m0 = _random_m(n, 61)
    hfield = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0
    B = 3
    rng = np.random.default_rng(62)
    mb = rng.normal(size=(B, n, 3))
    mb /= np.linalg.norm(mb, axis=2, keepdims=True)
    rows = np.repeat((np.arange(B)[:, None] * np.array([1e4, -5e3, 0.0]))[:, None, :], 4, axis=1)
    out = {}
    for sf in ("1", "0"):
        monkeypatch.setenv("MT_STAGE_FUSED", sf)
        kw = dict(alpha=4.42e3, gamma=2.21e5, tol=1e-6, precision="fp32")
        p = MicromagProblem(res, L, m0=m0, **kw)
        single = p.run_simulation(2e-11, 6, lambda t: np.ones((len(t), 1)) * hfield, 10)
        q = MicromagProblem(res, L, **kw)
        batch = q.run_simulation_batch(2e-11, 6, rows, 4, mb, return_fields=True)
        out[sf] = (single, batch, p.last_stats[0]["f_calls"])
    (s1, b1, f1), (s0, b0, f0) = out["1"], out["0"]
    print(f"res={res}: {f1} right-hand sides")
    assert f1 == f0
    for a, b in zip(s1, s0):
        assert np.array_equal(a, b)
    for a, b in zip(b1, b0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("toeplitz", [-1, 1])
def test_ll_in_near_band_kernel_same_bits(monkeypatch, toeplitz):
    """On uniform grids the split product's near-band kernel evaluates the LL
    right-hand side of its cell (MT_NEAR_LLG=0: the LL kernel runs after it and
    reads the correction from its own partial-sum slot).  The demag sum is
    formed in the same order either way (the GEMM's slots, then the
    correction), so trajectories and returned fields agree bit for bit, over
    the dense tensor and over the translation-invariant table."""
    from magtense.micromag import MicromagProblem

    res, L = [64, 16, 2], [128e-9, 32e-9, 3e-9]
    n, B = 2048, 6
    rng = np.random.default_rng(71)
    m0 = rng.normal(size=(B, n, 3))
    m0 /= np.linalg.norm(m0, axis=2, keepdims=True)
    fields = np.stack([np.linspace(-3e4, 3e4, B), 5e3 * np.ones(B), np.zeros(B)], axis=1)
    rows = np.repeat(fields[:, None, :], 4, axis=1)
    out = {}
    for v in ("1", "0"):
        monkeypatch.setenv("MT_NEAR_LLG", v)
        p = MicromagProblem(res, L, alpha=4.42e3, gamma=2.21e5, tol=1e-6, precision="bf16x2")
        p._toeplitz = toeplitz
        out[v] = p.run_simulation_batch(2e-11, 5, rows, 4, m0, return_fields=True)
    for a, b in zip(out["1"], out["0"]):
        assert np.array_equal(a, b)
