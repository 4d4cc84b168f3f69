"""GPU parity of the threshold-sparse demag path against the oracle.

* CSR build: per-component entry counts equal numpy's count of |K| > t on the
  same FP32 tensor (integer, exact); the fraction threshold equals the
  oracle's FP32 bisection exactly (same arithmetic, integer counts).
* Product: 1e-6 relative (norm-wise) to the FP64 product over the kept
  entries, for 1, 2, 3, 4, 5, 15, 17 and 64 states (every few-state kernel
  width, up to 16 states in groups of four, and the many-state kernel).
* Reference plug-in ABI (icl_*_sparse, 1-based host CSR): same bar.
* Block-sparse forms: same kept entries, tile count exact; split-BF16 tiles
  (precision bf16x2) within 1e-5 and FP32 tiles (FP32 MFMAs) within 1e-6 of
  the FP64 product over the kept entries and of the FP32 CSR product.
* Solver with demag_approx 'threshold' / 'threshold_fraction' (FP32 CSR,
  split-BF16 tiles, and FP32 tiles for batches of five or more states):
  trajectories within 1e-4 of the oracle run with the same threshold.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi
COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _tensor(res, L):
    import torch
    from oracle import oracle as O
    from magtense_amd import _lib

    Lb = _lib.lib()
    pts, d = O.setup_grid(res, L)
    n = pts.shape[0]
    dp = torch.as_tensor(pts).cuda()
    da = torch.as_tensor(d).cuda()
    K32 = torch.empty(6 * n * n, dtype=torch.float32, device="cuda")
    _lib.check(Lb.mt_prism_tensor(dp.data_ptr(), da.data_ptr(), 0, None, dp.data_ptr(), n, n, K32.data_ptr(), None,
                                  None))
    return n, K32


def _masked_product(K, thr, m):
    S, _, n = m.shape
    Km = np.where(np.abs(K) > thr, K, 0).astype(np.float64)
    out = np.zeros((S, 3, n))
    for a in range(3):
        for b in range(3):
            out[:, a] += m[:, b] @ Km[COMP[a][b]].T
    return out


@pytest.mark.parametrize("approx,value", [(2, 2e-3), (4, 0.9), (4, 0.5)])
def test_sparse_build_and_product(oracle, approx, value):
    import torch
    from magtense_amd import _lib

    Lb = _lib.lib()
    res, L = [16, 8, 2], [32e-9, 16e-9, 4e-9]
    n, K32 = _tensor(res, L)
    K = K32.cpu().numpy().reshape(6, n, n)
    h = C.c_void_p()
    _lib.check(Lb.mt_demag_sparse_create(K32.data_ptr(), n, approx, value, C.byref(h)))
    try:
        nnz = (C.c_longlong * 6)()
        thr = C.c_float()
        _lib.check(Lb.mt_demag_sparse_info(h, nnz, C.byref(thr)))
        t = np.float32(value) if approx == 2 else np.float32(oracle.threshold_fraction(K, value))
        assert thr.value == t
        expect = [int(np.count_nonzero(np.abs(K[c]) > t)) for c in range(6)]
        assert list(nnz) == expect
        print(f"approx={approx} value={value}: threshold {thr.value:.3e}, nnz {list(nnz)} of {n * n} each")
        for S in (1, 2, 3, 4, 5, 15, 17, 64):
            rng = np.random.default_rng(S)
            m = rng.normal(size=(S, 3, n)).astype(np.float32)
            ref = -_masked_product(K, t, m.astype(np.float64))
            dm = torch.as_tensor(m).cuda()
            out = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
            _lib.check(Lb.mt_demag_apply_sparse(h, dm.data_ptr(), out.data_ptr(), S, -1.0))
            got = out.cpu().numpy().reshape(S, 3, n)
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert rel < 1e-6, (S, rel)
    finally:
        Lb.mt_demag_sparse_destroy(h)


def test_icl_sparse_plugin_abi(oracle):
    from magtense_amd import _lib

    Lb = _lib.lib()
    res, L = [8, 8, 1], [16e-9, 16e-9, 2e-9]
    n, K32 = _tensor(res, L)
    K = K32.cpu().numpy().reshape(6, n, n)
    thr = np.float32(1e-3)
    keep = []
    for c in range(6):
        D = np.where(np.abs(K[c]) > thr, K[c], 0)
        rows, cols = np.nonzero(D)
        vals = np.ascontiguousarray(D[rows, cols], dtype=np.float32)
        colInds = np.ascontiguousarray(cols + 1, dtype=np.int32)
        rowInds = np.ascontiguousarray(np.concatenate([[1], 1 + np.cumsum(np.bincount(rows, minlength=n))]),
                                       dtype=np.int32)
        keep += [vals, colInds, rowInds]
        nn, nnz, mat = C.c_int(n), C.c_int(vals.size), C.c_int(c + 1)
        Lb.icl_initDemagMatrices_sparse(C.byref(nn), C.byref(nnz), C.byref(mat), vals.ctypes.data,
                                        colInds.ctypes.data, rowInds.ctypes.data)
    rng = np.random.default_rng(9)
    M = rng.normal(size=(3, n)).astype(np.float32)
    H = np.zeros((3, n), dtype=np.float32)
    nn, pref = C.c_int(n), C.c_float(-1.0)
    Lb.icl_MVMult_GetH_sparse(M[0].ctypes.data, M[1].ctypes.data, M[2].ctypes.data, H[0].ctypes.data,
                              H[1].ctypes.data, H[2].ctypes.data, C.byref(nn), C.byref(pref))
    Lb.icl_destroy()
    ref = -_masked_product(K, thr, M[None].astype(np.float64))[0]
    rel = np.linalg.norm(H - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel


@pytest.mark.parametrize("kind", ["bf16x2", "f32"])
@pytest.mark.parametrize("res,grid_nx,approx,value", [([64, 16, 2], 64, 4, 0.9), ([64, 16, 2], 0, 2, 2e-3),
                                                     ([20, 7, 1], 0, 4, 0.5), ([32, 8, 1], 32, 2, 10.0)])
def test_block_sparse_product(oracle, res, grid_nx, approx, value, kind):
    """The block-sparse forms of the threshold tensor: the same kept entries
    (per-component counts exact), fragments only for the tiles that hold one,
    and the batched MFMA product — split-BF16 within 1e-5, FP32 (FP32 MFMAs)
    within 1e-6 (norm-wise) — of the FP64 product over the kept entries and of
    the FP32 CSR product, for 1 to 70 states; rows of a uniform grid grouped
    across grid rows (grid_nx) or consecutively (grid_nx = 0, a ragged 140-cell
    grid); a threshold above every entry keeps nothing and gives zero fields."""
    import torch
    from magtense_amd import _lib

    Lb = _lib.lib()
    L = [res[0] * 2e-9, res[1] * 2e-9, res[2] * 1.5e-9]
    n, K32 = _tensor(res, L)
    K = K32.cpu().numpy().reshape(6, n, n)
    hb, hc = C.c_void_p(), C.c_void_p()
    create = Lb.mt_demag_sparse_create_bf16x2 if kind == "bf16x2" else Lb.mt_demag_sparse_create_f32_tiles
    bar = 1e-5 if kind == "bf16x2" else 1e-6
    _lib.check(create(K32.data_ptr(), n, approx, value, grid_nx, C.byref(hb)))
    _lib.check(Lb.mt_demag_sparse_create(K32.data_ptr(), n, approx, value, C.byref(hc)))
    try:
        nnz_b, nnz_c = (C.c_longlong * 6)(), (C.c_longlong * 6)()
        thr_b, thr_c = C.c_float(), C.c_float()
        _lib.check(Lb.mt_demag_sparse_info(hb, nnz_b, C.byref(thr_b)))
        _lib.check(Lb.mt_demag_sparse_info(hc, nnz_c, C.byref(thr_c)))
        assert thr_b.value == thr_c.value and list(nnz_b) == list(nnz_c)
        t = np.float32(thr_b.value)
        assert list(nnz_b) == [int(np.count_nonzero(np.abs(K[c]) > t)) for c in range(6)]
        ent, tiles, nbytes = C.c_longlong(), C.c_longlong(), C.c_longlong()
        ks = C.c_int()
        _lib.check(Lb.mt_demag_sparse_blocks(hb, C.byref(ent), C.byref(tiles), C.byref(nbytes), C.byref(ks)))
        np32 = (n + 31) // 32 * 32
        kept = np.zeros((np32 // 32, np32 // 16), dtype=bool)
        for c in range(6):
            mk = np.zeros((np32, np32), dtype=bool)
            mk[:n, :n] = np.abs(K[c]) > t
            kept |= mk.reshape(np32 // 32, 32, np32 // 16, 16).any(axis=(1, 3))
        assert tiles.value == int(kept.sum())
        assert nbytes.value == ent.value * 4 * 12 * 1024
        print(f"{kind} res={res} grid_nx={grid_nx} approx={approx} value={value}: kept {sum(nnz_b)} entries "
              f"({sum(nnz_b) / (6 * n * n):.3f} of dense), tiles {tiles.value} of {kept.size}, list entries "
              f"{ent.value}, k-split {ks.value}")
        for S in (1, 5, 64, 70):
            rng = np.random.default_rng(S)
            m = rng.normal(size=(S, 3, n)).astype(np.float32)
            ref = -_masked_product(K, t, m.astype(np.float64))
            dm = torch.as_tensor(m).cuda()
            ob = torch.full((S * 3 * n,), float("nan"), dtype=torch.float32, device="cuda")
            oc = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
            _lib.check(Lb.mt_demag_apply_sparse(hb, dm.data_ptr(), ob.data_ptr(), S, -1.0))
            _lib.check(Lb.mt_demag_apply_sparse(hc, dm.data_ptr(), oc.data_ptr(), S, -1.0))
            gb = ob.cpu().numpy().reshape(S, 3, n)
            gc = oc.cpu().numpy().reshape(S, 3, n)
            if sum(nnz_b) == 0:
                assert np.all(gb == 0) and np.all(gc == 0)
                continue
            nr = np.linalg.norm(ref)
            print(f"  S={S}: tiles vs FP64 {np.linalg.norm(gb - ref) / nr:.2e}, CSR vs FP64 "
                  f"{np.linalg.norm(gc - ref) / nr:.2e}")
            assert np.linalg.norm(gb - ref) / nr < bar, (S, np.linalg.norm(gb - ref) / nr)
            assert np.linalg.norm(gb - gc) / nr < bar
    finally:
        Lb.mt_demag_sparse_destroy(hb)
        Lb.mt_demag_sparse_destroy(hc)


@pytest.mark.parametrize("mode,value,precision", [("threshold", 1e-3, "fp32"), ("threshold_fraction", 0.8, "fp32"),
                                                  ("threshold_fraction", 0.8, "bf16x2")])
def test_solver_threshold_matches_oracle(oracle, mode, value, precision):
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    rng = np.random.default_rng(21)
    m0 = rng.normal(size=(64, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    hfield = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0
    p = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6, demag_approx=mode, dem_thres=value,
                        precision=precision)
    t_out, M, *_ = p.run_simulation(1e-10, 11, lambda t: np.ones((len(t), 1)) * hfield, 20)
    t = np.linspace(0, 1e-10, 11)
    hext = np.zeros((20, 4))
    hext[:, 0] = np.linspace(0, 1e-10, 20)
    hext[:, 1:] = hfield
    q = oracle.Problem(res, L, hext=hext, t=t, m0=np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]]), alpha=4.42e3,
                       gamma=2.21e5, tol=1e-6, demag_approx=2 if mode == "threshold" else 4, dem_thres=value)
    _, y, _, st = q.run()
    yr = np.transpose(y[0], (0, 2, 1))
    err = np.abs(M[:, :, 0, :] - yr).max()
    print(mode, value, precision, "threshold info", p.last_demag_info, "max traj err", err)
    assert err < 1e-4
    assert sum(p.last_demag_info["nnz"]) < 6 * 64 * 64


def test_solver_threshold_fp32_tiles_batch(oracle, monkeypatch):
    """66 FP32 states with a threshold tensor take the FP32 tiles (FP32 MFMAs
    over the stored tiles, one pass per 64-state group; the last two states
    form a group whose descriptors go by value): every state within 1e-4 of its
    oracle run with the same threshold, and within 1e-5 of the same batch run
    on the CSR form (MT_F32_TILE_STATES=0), which differs only in the FP32
    summation order."""
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    n, B = 64, 66
    rng = np.random.default_rng(33)
    m0 = rng.normal(size=(B, n, 3))
    m0 /= np.linalg.norm(m0, axis=2, keepdims=True)
    ang = np.linspace(0, 2 * np.pi, B, endpoint=False)
    fields = np.stack([3e4 * np.cos(ang), 3e4 * np.sin(ang), np.zeros(B)], axis=1)
    rows = np.repeat(fields[:, None, :], 10, axis=1)
    kw = dict(alpha=4.42e3, gamma=2.21e5, tol=1e-6, demag_approx="threshold_fraction", dem_thres=0.8)
    p = MicromagProblem(res, L, **kw)
    _, M_t = p.run_simulation_batch(5e-11, 6, rows, 10, m0)
    assert p.last_demag_info["form"] == "fp32 tiles", p.last_demag_info
    monkeypatch.setenv("MT_F32_TILE_STATES", "0")
    q = MicromagProblem(res, L, **kw)
    _, M_c = q.run_simulation_batch(5e-11, 6, rows, 10, m0)
    assert q.last_demag_info["form"] == "csr", q.last_demag_info
    d = np.abs(M_t - M_c).max()
    print(f"fp32 tiles vs CSR batch: max |dm| {d:.2e}")
    assert d < 1e-5, d
    t = np.linspace(0, 5e-11, 6)
    for b in (0, 37, 64, 65):
        hext = np.zeros((10, 4))
        hext[:, 0] = np.linspace(0, 5e-11, 10)
        hext[:, 1:] = fields[b]
        o = oracle.Problem(res, L, hext=hext, t=t, m0=np.concatenate([m0[b][:, 0], m0[b][:, 1], m0[b][:, 2]]),
                           alpha=4.42e3, gamma=2.21e5, tol=1e-6, demag_approx=4, dem_thres=0.8)
        _, y, _, _ = o.run()
        err = np.abs(M_t[b] - np.transpose(y[0], (0, 2, 1))).max()
        print(f"state {b}: max |m - oracle| {err:.2e}")
        assert err < 1e-4, (b, err)
