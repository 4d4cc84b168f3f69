"""Unstructured-prism micromagnetic grids on the GPU (grid_type
'unstructuredPrisms': LandauLifshitzEquationSolver.f90:1044-1133 tensor with
N_ave receiver averaging, :1530-1544 loaded exchange operator), through the
drop-in MicromagProblem.

* A uniform 32 x 8 film given as unstructured prisms (its centres, edges and
  the uniform exchange operator as a CSR) follows the uniform-grid run.
* A graded film (column widths 4.0-5.2 nm, N = 48) with a finite-volume
  exchange operator, N_ave = (1, 1, 1) and (2, 2, 1), from a smooth seeded
  state: the FP32 tensor the solver assembles is within 2 ulp of the FP64
  oracle's averaged tensor, and the trajectory matches the oracle's run of the
  same problem within 1e-4; the split-BF16 product follows the FP32 one within
  1e-4.  (Entries that the sub-point average cancels to ~0 are held to 1e-12
  of the largest entry instead of 2 ulp.)
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU0 = 4 * np.pi * 1e-7


def _unit(rng, n):
    m = rng.normal(size=(n, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def graded_film(ncol=12, nrow=4):
    w = 4e-9 * (1 + 0.1 * (np.arange(ncol) % 4))
    xe = np.concatenate([[0.0], np.cumsum(w)])
    xc = (xe[:-1] + xe[1:]) / 2
    dy, dz = 5e-9, 3e-9
    n = ncol * nrow
    pts = np.zeros((n, 3))
    abc = np.zeros((n, 3))
    rows, cols, vals = [], [], []
    for j in range(nrow):
        for i in range(ncol):
            r = i + j * ncol
            pts[r] = (xc[i], (j + 0.5) * dy, 0.0)
            abc[r] = (w[i], dy, dz)
            entries = {}
            for di in (-1, 1):  # finite-volume second difference, Neumann where the neighbour is missing
                k = i + di
                if 0 <= k < ncol:
                    c = 1.0 / (0.5 * (w[i] + w[k]) * w[i])
                    entries[r + di] = entries.get(r + di, 0.0) + c
                    entries[r] = entries.get(r, 0.0) - c
            for dj in (-1, 1):
                k = j + dj
                if 0 <= k < nrow:
                    c = 1.0 / dy ** 2
                    entries[r + dj * ncol] = entries.get(r + dj * ncol, 0.0) + c
                    entries[r] = entries.get(r, 0.0) - c
            for cidx in sorted(entries):
                rows.append(r)
                cols.append(cidx)
                vals.append(entries[cidx])
    import scipy.sparse as sp

    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return pts, abc, A


def _exch_tuple(A):
    A = A.tocsr()
    A.sort_indices()
    return (A.indptr[:-1] + 1, A.indptr[1:] + 1, A.indices + 1, A.data)


def test_uniform_film_as_unstructured_prisms(oracle):
    """Fields at the start: exchange equal to rounding (same operator, summed in
    CSR order instead of the stencil), demag within 1e-6 (the evaluation point
    (p - a/2) + a/2 moves the FP32 tensor entries by at most an ulp).  From a
    smooth state the trajectories agree to 1e-5 (from random unit vectors the
    exchange-driven dynamics amplify last-bit differences through the step
    sequence)."""
    import scipy.sparse as sp
    from magtense.micromag import MicromagProblem

    res, L = [32, 8, 1], [500e-9, 125e-9, 3e-9]
    n = 256
    pts, d = oracle.setup_grid(res, L)
    rowptr, col, val = oracle.exchange_csr(res, L)
    A = sp.csr_matrix((val, col, rowptr), shape=(n, n))
    h = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0

    def hf(t):
        return np.ones((len(t), 1)) * h

    def both(m0, t_end):
        pu = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6)
        out_u = pu.run_simulation(t_end, 21, hf, 10)
        ps = MicromagProblem([n, 1, 1], L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6,
                             grid_type="unstructuredPrisms")
        ps.grid_pts[:] = pts
        ps.grid_abc[:] = d
        ps.set_exchange_matrix(A)
        out_s = ps.run_simulation(t_end, 21, hf, 10)
        return out_u, out_s, pu.last_stats[0], ps.last_stats[0]

    m_rand = _unit(np.random.default_rng(5), n)
    (_, _, pts_u, Hj_u, _, Hd_u, _), (_, _, pts_s, Hj_s, _, Hd_s, _), _, _ = both(m_rand, 2e-12)
    assert np.array_equal(pts_s, pts_u)
    e_j = np.abs(Hj_u[0] - Hj_s[0]).max() / np.abs(Hj_u[0]).max()
    e_d = np.abs(Hd_u[0] - Hd_s[0]).max() / np.abs(Hd_u[0]).max()
    x = pts[:, 0] / L[0]
    m_smooth = np.stack([np.cos(0.3 * np.sin(2 * np.pi * x)), np.sin(0.3 * np.sin(2 * np.pi * x)),
                         0.05 * np.ones(n)], axis=1)
    m_smooth /= np.linalg.norm(m_smooth, axis=1, keepdims=True)
    (_, Mu, *_), (_, Ms_, *_), su, ss = both(m_smooth, 2e-10)
    diff = np.abs(Mu - Ms_).max()
    print(f"uniform vs unstructured-as-uniform: H_exc {e_j:.2e}, H_dem {e_d:.2e} at t = 0; smooth start max |dm| "
          f"{diff:.2e}; steps {su['steps']} / {ss['steps']}")
    assert e_j < 1e-12 and e_d < 1e-6
    assert diff < 1e-5


@pytest.mark.parametrize("nave", [(1, 1, 1), (2, 2, 1)])
def test_graded_film_matches_oracle(oracle, tmp_path, nave):
    from magtense import micromag
    from magtense.micromag import MicromagProblem

    pts, abc, A = graded_film()
    n = pts.shape[0]
    # a smooth, seeded start; the cells keep the run below RKSUITE's
    # 5000-evaluation work limit (flag 3, where runs part by construction)
    rng = np.random.default_rng(17)
    x = pts[:, 0] / pts[:, 0].max()
    ph = 0.4 * np.sin(2 * np.pi * x + rng.uniform(0, 2 * np.pi))
    m0 = np.stack([np.cos(ph), np.sin(ph), 0.1 + 0.02 * rng.normal(size=n)], axis=1)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    h = np.array([20.0, 5.0, 0.0]) / 1000 / MU0
    t_end, nt = 1e-10, 21

    def run(precision, fname="t"):
        p = MicromagProblem([n, 1, 1], [1, 1, 1], m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6,
                            grid_type="unstructuredPrisms", N_ave=list(nave), precision=precision, filename=fname)
        p.grid_pts[:] = pts
        p.grid_abc[:] = abc
        p.set_exchange_matrix(A)
        t, M, *_ = p.run_simulation(t_end, nt, lambda tt: np.ones((len(tt), 1)) * h, 10)
        return t, np.squeeze(M, axis=2), p

    path = str(tmp_path / "K_graded.bin")
    t, M32, p32 = run("fp32", path)
    _, Mbf, _ = run("bf16x2")
    _, Mmf, _ = run("fp32_mfma")  # the dense tensor as FP32 MFMA fragments (grid-agnostic)
    # the tensor the solver assembled (written to the tensor file) vs the oracle's FP64 average
    K = micromag.load_demag_tensor(path, n).reshape(6, n, n)
    ref = oracle.tensor_unstructured(pts, abc, nave)
    r32 = ref.astype(np.float32)
    # 2 FP32 ulp, with a floor of 1e-12 of the largest entry for the entries
    # that the sub-point average cancels to ~0 (FP64 rounding of the sum)
    err = np.abs(K.astype(np.float64) - ref)
    bar = 2 * np.spacing(np.abs(r32)).astype(np.float64) + 1e-12 * np.abs(ref).max()
    ulps = np.abs(K.astype(np.float64) - r32) / np.spacing(np.abs(r32))
    assert np.all(err <= bar), (err / bar).max()
    # the oracle's run of the same problem
    hext = np.zeros((10, 4))
    hext[:, 0] = np.linspace(0, t_end, 10)
    hext[:, 1:] = h
    soa = np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]])
    q = oracle.Problem([n, 1, 1], [1, 1, 1], alpha=4.42e3, gamma=2.21e5, tol=1e-6, hext=hext,
                       t=np.linspace(0, t_end, nt), m0=soa, return_fields=False, grid_pts=pts, grid_abc=abc,
                       n_ave=nave, exch=_exch_tuple(A))
    _, y, _, st = q.run()
    orc = np.transpose(y[0], (0, 2, 1))
    d_or = np.abs(M32 - orc).max()
    d_bf = np.abs(Mbf - M32).max()
    d_mf = np.abs(Mmf - M32).max()
    print(f"graded film N_ave {nave}: tensor {np.mean(ulps <= 2):.4f} within 2 ulp, max err/bar {(err / bar).max():.2f}; GPU vs oracle {d_or:.2e}; bf16x2 vs fp32 "
          f"{d_bf:.2e}, fp32_mfma vs fp32 {d_mf:.2e}; gpu {p32.last_stats[0]}; oracle {st}")
    assert p32.last_stats[0]["flag3"] == 0 and st["flag3"] == 0
    assert d_or < 1e-4 and d_bf < 1e-4
    assert d_mf < 1e-5  # FP32 products and sums in another order than the per-state rows
    assert np.abs(np.linalg.norm(M32, axis=2) - 1).max() < 1e-4
