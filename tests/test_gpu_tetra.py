"""Tetrahedra on the GPU (tileType 5: getN_tensor_tetrahedron,
TileNComponents.f90:944-961 over getN_Triangle, TileTriangle.f90:26-286).

* Magnetostatic tiles: tetrahedron tiles (offset, rotation, eight-fold
  symmetry, mixed with prisms) through magstatics.get_demag_tensor and
  get_H_field against the oracle, 1e-13 absolute (the tensor entries are sums
  of O(1) face terms).
* Tetrahedral micromagnetic grid (LandauLifshitzEquationSolver.f90:988-1041):
  3 x 2 x 1 boxes of 5 x 4 x 3 nm, each split into six Kuhn tetrahedra (36
  elements, evaluation points at the centroids), a face-neighbour exchange
  operator: the FP32 tensor the solver assembles is within 2 ulp of the
  oracle's (floor 1e-12 for entries that cancel), the trajectory matches the
  oracle's run within 1e-4, and split-BF16 follows FP32 within 1e-4.
"""

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU0 = 4 * np.pi * 1e-7
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_tetra import box_mesh  # noqa: E402


def test_tetra_tiles_match_oracle(oracle):
    from magtense import magstatics as ms

    rng = np.random.default_rng(21)
    nt = 4
    t = ms.Tiles(nt)
    t.tile_type = [5, 5, 2, 5]
    v = np.array([[0, 0, 0], [4e-9, 0, 0], [0, 3e-9, 0], [0, 0, 2e-9]], dtype=float)
    verts = np.zeros((nt, 3, 4))
    for i in range(nt):
        verts[i] = (v + rng.uniform(-0.5e-9, 0.5e-9, (4, 3))).T
    t.vertices = verts
    t.size = [3e-9, 2e-9, 1e-9]
    t.offset = rng.uniform(-3e-9, 3e-9, (nt, 3))
    t.rot = rng.uniform(-0.5, 0.5, (nt, 3))
    t.M = rng.normal(size=(nt, 3)) * 1e5
    t.use_sym[1] = 1  # read-only attributes: written in place
    t.sym_op[:] = [1.0, -1.0, 1.0]
    pts = rng.uniform(-10e-9, 10e-9, (50, 3))
    N = ms.get_demag_tensor(t, pts)
    ref = oracle.tiles_tensor(t.size, t.offset, pts, rot=t.rot, tile_type=t.tile_type, use_sym=t.use_sym,
                              sym_ops=t.sym_op, vertices=t.vertices)
    err = np.abs(N - ref).max()
    H = ms.get_H_field(t, pts)
    Href = oracle.tiles_field(t.size, t.offset, t.M, pts, rot=t.rot, tile_type=t.tile_type, use_sym=t.use_sym,
                              sym_ops=t.sym_op, vertices=t.vertices)
    eh = np.abs(H - Href).max() / np.abs(Href).max()
    H2 = ms.get_H_field(t, pts, N)  # stored tensor
    print(f"tetra tiles: tensor max |d| {err:.2e}; field rel {eh:.2e}")
    assert err <= 1e-13
    assert eh <= 1e-12
    assert np.abs(H2 - H).max() <= 1e-12 * np.abs(H).max()


def _face_laplacian(elems, w):
    import scipy.sparse as sp

    n = len(elems)
    faces = {}
    rows, cols = [], []
    for e in range(n):
        for k in range(4):
            f = tuple(sorted(np.delete(elems[e], k)))
            faces.setdefault(f, []).append(e)
    for es in faces.values():
        if len(es) == 2:
            rows += [es[0], es[1]]
            cols += [es[1], es[0]]
    A = sp.csr_matrix((np.full(len(rows), w), (rows, cols)), shape=(n, n))
    A = A - sp.diags(np.asarray(A.sum(axis=1)).ravel())
    return A.tocsr()


def test_tetra_grid_matches_oracle(oracle, tmp_path):
    from magtense import micromag
    from magtense.micromag import MicromagProblem

    nodes, elems, cent, box, a = box_mesh()
    n = len(elems)
    A = _face_laplacian(elems, 1.0 / (4e-9) ** 2)
    rng = np.random.default_rng(31)
    x = cent[:, 0] / cent[:, 0].max()
    ph = 0.4 * np.sin(2 * np.pi * x + rng.uniform(0, 2 * np.pi))
    m0 = np.stack([np.cos(ph), np.sin(ph), 0.1 + 0.02 * rng.normal(size=n)], axis=1)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    h = np.array([20.0, 5.0, 0.0]) / 1000 / MU0
    t_end, nt = 1e-10, 21

    def run(precision, fname="t"):
        p = MicromagProblem([n, 1, 1], [1, 1, 1], m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6,
                            grid_type="tetrahedron", grid_nnod=len(nodes), precision=precision, filename=fname)
        p.grid_pts[:] = cent
        p.grid_nod[:] = nodes
        p.grid_ele = elems.T.astype(np.float64)
        p.set_exchange_matrix(A)
        t, M, pts_out, *_ = p.run_simulation(t_end, nt, lambda tt: np.ones((len(tt), 1)) * h, 10)
        return t, np.squeeze(M, axis=2), p, pts_out

    path = str(tmp_path / "K_tetra.bin")
    t, M32, p32, pts_out = run("fp32", path)
    _, Mbf, _, _ = run("bf16x2")
    _, Mmf, _, _ = run("fp32_mfma")  # the dense tensor as FP32 MFMA fragments
    assert np.array_equal(pts_out, cent)
    K = micromag.load_demag_tensor(path, n)
    ref = np.zeros((6, n, n))
    for j in range(n):
        vj = nodes[elems[j] - 1]
        for i in range(n):
            Nij = oracle.tetra_n(vj, cent[i])
            ref[:, i, j] = (Nij[0, 0], Nij[0, 1], Nij[0, 2], Nij[1, 1], Nij[1, 2], Nij[2, 2])
    r32 = ref.astype(np.float32)
    err = np.abs(K.astype(np.float64) - ref)
    bar = 2 * np.spacing(np.abs(r32)).astype(np.float64) + 1e-12 * np.abs(ref).max()
    assert np.all(err <= bar), (err / bar).max()
    # the oracle's run of the same problem
    A.sort_indices()
    exch = (A.indptr[:-1] + 1, A.indptr[1:] + 1, A.indices + 1, A.data)
    hext = np.zeros((10, 4))
    hext[:, 0] = np.linspace(0, t_end, 10)
    hext[:, 1:] = h
    soa = np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]])
    q = oracle.Problem([n, 1, 1], [1, 1, 1], alpha=4.42e3, gamma=2.21e5, tol=1e-6, hext=hext,
                       t=np.linspace(0, t_end, nt), m0=soa, return_fields=False, grid_pts=cent, exch=exch,
                       grid_nodes=nodes, grid_elements=elems)
    _, y, _, st = q.run()
    orc = np.transpose(y[0], (0, 2, 1))
    d_or = np.abs(M32 - orc).max()
    d_bf = np.abs(Mbf - M32).max()
    d_mf = np.abs(Mmf - M32).max()
    print(f"tetra grid: tensor max err/bar {(err / bar).max():.2f}; GPU vs oracle {d_or:.2e}; bf16x2 vs fp32 "
          f"{d_bf:.2e}, fp32_mfma vs fp32 {d_mf:.2e}; gpu {p32.last_stats[0]}; oracle {st}")
    assert p32.last_stats[0]["flag3"] == 0 and st["flag3"] == 0
    assert d_or < 1e-4 and d_bf < 1e-4
    assert d_mf < 1e-5  # FP32 products and sums in another order than the per-state rows
