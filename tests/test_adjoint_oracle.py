"""CPU pins of the FP64 oracle's discrete RK4 adjoint: central finite differences
of the integrated loss (FP64 demag, so differences are not masked by rounding)."""

import numpy as np


def _setup(oracle, seed=0):
    res, L = [4, 3, 2], [20e-9, 15e-9, 6e-9]
    n = 24
    rng = np.random.default_rng(seed)
    ms = rng.uniform(7e5, 9e5, size=n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    m0 = rng.normal(size=(3, n))
    m0 /= np.linalg.norm(m0, axis=0, keepdims=True)
    target = rng.normal(size=(3, n))
    target /= np.linalg.norm(target, axis=0, keepdims=True)
    h = np.array([2.0e4, -1.0e4, 5.0e3])
    return res, L, n, ms, u, m0.reshape(-1), target.reshape(-1), h


def _loss(mT, target, n):
    return mT[2 * n:].mean() + np.linalg.norm(mT - target)


def _dloss(mT, target, n):
    g = (mT - target) / np.linalg.norm(mT - target)
    g[2 * n:] += 1.0 / n
    return g


def test_oracle_adjoint_matches_finite_differences(oracle):
    res, L, n, ms, u, m0, target, h = _setup(oracle)
    dt, nsteps = 2e-14, 60

    def run(ms_, h_, m0_, grad=False):
        q = oracle.Problem(res, L, Ms=ms_, K0=3e4, u_ea=u, gamma=2.21e5, alpha=4.42e3, demag_fp64=True)
        mT, *_ = q.rk4_adjoint(dt, nsteps, h_, m0_)
        if not grad:
            return _loss(mT, target, n)
        return q.rk4_adjoint(dt, nsteps, h_, m0_, _dloss(mT, target, n))

    mT, gm0, gh, gms = run(ms, h, m0, grad=True)
    # applied field (3 values), central differences h = 1e-4 relative
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-4 * np.abs(h).max()
        fd = (run(ms, h + e, m0) - run(ms, h - e, m0)) / (2 * e[k])
        assert abs(fd - gh[k]) <= 1e-6 * np.abs(gh).max() + 1e-5 * abs(fd), (k, fd, gh[k])
    # Ms: a few components and one random direction
    rng = np.random.default_rng(9)
    for i in (0, 7, 23):
        e = np.zeros(n)
        e[i] = 1e-4 * ms[i]
        fd = (run(ms + e, h, m0) - run(ms - e, h, m0)) / (2 * e[i])
        assert abs(fd - gms[i]) <= 1e-4 * np.abs(gms).max() + 1e-4 * abs(fd), (i, fd, gms[i])
    v = rng.normal(size=n) * 1e-4 * ms
    fd = (run(ms + v, h, m0) - run(ms - v, h, m0)) / 2
    assert abs(fd - gms @ v) <= 1e-4 * abs(fd) + 1e-9
    # initial magnetization: random direction
    w = rng.normal(size=3 * n) * 1e-5
    fd = (run(ms, h, m0 + w) - run(ms, h, m0 - w)) / 2
    assert abs(fd - gm0 @ w) <= 1e-4 * abs(fd) + 1e-12
