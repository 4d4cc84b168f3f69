"""FP64 oracle gradients of the differentiation configuration (run once, from
the repository root; about 10 minutes per precision on 8 cores):

    python tests/golden/make_c5_adjoint.py fp32
    python tests/golden/make_c5_adjoint.py fp64

128 x 32 x 1 film (N = 4096, 500 x 125 x 3 nm), per-cell Ms uniform in
[7e5, 9e5] A/m, smooth near-uniform +x start, applied field (-24.6, 4.3, 0)
mT, 2000 fixed RK4 steps of 0.1 ps; loss = mean m_z(T) + |m(T) - m_target|
with seeded random unit target vectors.  The oracle's discrete adjoint
(serial RK4 restatement, FP64 states, demag in the reference's FP32
arithmetic -- FP32 tensor and FP32 m, FP64-exact sums -- or all-FP64) gives
dL/dH_ext (3) and dL/dMs (4096); both are stored with m(T) and the inputs'
seed.  The oracle's adjoint is pinned by central differences
(tests/test_adjoint_oracle.py).
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

RES, L = [128, 32, 1], [500e-9, 125e-9, 3e-9]
N = 4096
SEED = 2024
DT, NSTEPS = 1e-13, 2000
MU0 = 4e-7 * np.pi


def c5_inputs():
    """(ms, m0 (N, 3), target (N, 3), h) exactly as the GPU test draws them."""
    rng = np.random.default_rng(SEED)
    ms = rng.uniform(7e5, 9e5, size=N)
    nx, ny, nz = RES
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ix = np.tile(ix.reshape(-1), nz)
    iy = np.tile(iy.reshape(-1), nz)
    m0 = np.stack([np.ones(ix.size), 0.2 * np.sin(2 * np.pi * (ix + 0.5) / nx),
                   0.1 * np.cos(2 * np.pi * (iy + 0.5) / ny)], axis=1)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    t = rng.normal(size=(N, 3))
    target = t / np.linalg.norm(t, axis=1, keepdims=True)
    h = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0
    return ms, m0, target, h


def main():
    precision = sys.argv[1]
    nsteps = int(sys.argv[2]) if len(sys.argv) > 2 else NSTEPS
    ms, m0, target, h = c5_inputs()
    soa = lambda a: np.concatenate([a[:, 0], a[:, 1], a[:, 2]])  # noqa: E731
    q = O.Problem(RES, L, Ms=ms, gamma=2.21e5, alpha=0.02 * 2.21e5, demag_fp64=precision == "fp64")
    t0 = time.time()
    mT, *_ = q.rk4_adjoint(DT, nsteps, h, soa(m0))
    mT = mT.reshape(3, N).T
    d = (mT - target) / np.linalg.norm(mT - target)
    d[:, 2] += 1.0 / N
    _, gm0, gh, gms = q.rk4_adjoint(DT, nsteps, h, soa(m0), soa(d))
    out = os.path.join(HERE, f"c5_adjoint_oracle_{precision}.npz")
    if nsteps == NSTEPS:
        np.savez_compressed(out, grad_h=gh, grad_ms=gms, mT=mT, seed=SEED, nsteps=nsteps, dt=DT)
        print("wrote", out)
    print(f"{precision}: {nsteps} steps in {time.time() - t0:.1f} s; grad h {gh}; |grad Ms| {np.linalg.norm(gms):.6e}")


if __name__ == "__main__":
    main()
