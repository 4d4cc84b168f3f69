"""GPU parity of the batched FP32 MFMA demag product (precision 'fp32_mfma').

* Product (mt_f32_mfma_tensor + mt_demag_apply_f32_mfma): within 1e-6
  (norm-wise) of the FP64 product and of the per-state FP32 mat-vec, for 1 to
  70 states on a 64-cell, a ragged 540-cell and a 4096-cell grid.
* Solver: a batch of eight states follows the per-state FP32 path (the
  reference's arithmetic, summed in another order) within 1e-5 and the FP64
  oracle within 1e-4; a single-state run and a threshold run (FP32 tiles)
  follow the oracle within 1e-4.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
MU0 = 4e-7 * np.pi


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


@pytest.mark.parametrize("res", [[8, 8, 1], [20, 9, 3], [64, 32, 2]])
def test_fp32_mfma_product(oracle, res):
    import torch
    from magtense_amd import _lib

    Lb = _lib.lib()
    L = [res[0] * 2e-9, res[1] * 2.5e-9, res[2] * 1.5e-9]
    n, K32 = _tensor(res, L)
    K = K32.cpu().numpy().reshape(6, n, n).astype(np.float64)
    frags = torch.empty(Lb.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda")
    _lib.check(Lb.mt_f32_mfma_tensor(K32.data_ptr(), n, frags.data_ptr()))
    for S in (1, 5, 64, 70):
        rng = np.random.default_rng(S + n)
        m = rng.normal(size=(S, 3, n)).astype(np.float32)
        ref = np.zeros((S, 3, n))
        for a in range(3):
            for b in range(3):
                ref[:, a] -= m[:, b].astype(np.float64) @ K[COMP[a][b]].T
        dm = torch.as_tensor(m).cuda()
        hm = torch.full((S * 3 * n,), float("nan"), dtype=torch.float32, device="cuda")
        hf = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        _lib.check(Lb.mt_demag_apply_f32_mfma(frags.data_ptr(), n, dm.data_ptr(), hm.data_ptr(), S, -1.0))
        _lib.check(Lb.mt_demag_apply_f32(K32.data_ptr(), n, dm.data_ptr(), hf.data_ptr(), S, -1.0))
        gm = hm.cpu().numpy().reshape(S, 3, n)
        gf = hf.cpu().numpy().reshape(S, 3, n)
        nr = np.linalg.norm(ref)
        em, ef = np.linalg.norm(gm - ref) / nr, np.linalg.norm(gf - ref) / nr
        print(f"res={res} S={S}: fp32 MFMA vs FP64 {em:.2e}, per-state FP32 vs FP64 {ef:.2e}")
        assert em < 1e-6, (S, em)
        assert np.linalg.norm(gm - gf) / nr < 1e-6


def _oracle_traj(oracle, res, L, m0, field, t_end, nt, nt_h, **kw):
    t = np.linspace(0, t_end, nt)
    hext = np.zeros((nt_h, 4))
    hext[:, 0] = np.linspace(0, t_end, nt_h)
    hext[:, 1:] = field
    q = oracle.Problem(res, L, hext=hext, t=t, m0=np.concatenate([m0[:, 0], m0[:, 1], m0[:, 2]]), **kw)
    _, y, _, _ = q.run()
    return np.transpose(y[0], (0, 2, 1))


def test_solver_fp32_mfma_batch(oracle):
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    n, B = 64, 8
    rng = np.random.default_rng(44)
    m0 = rng.normal(size=(B, n, 3))
    m0 /= np.linalg.norm(m0, axis=2, keepdims=True)
    ang = np.linspace(0, 2 * np.pi, B, endpoint=False)
    fields = np.stack([3e4 * np.cos(ang), 3e4 * np.sin(ang), np.zeros(B)], axis=1)
    rows = np.repeat(fields[:, None, :], 10, axis=1)
    kw = dict(alpha=4.42e3, gamma=2.21e5, tol=1e-6)
    _, M_m = MicromagProblem(res, L, precision="fp32_mfma", **kw).run_simulation_batch(5e-11, 6, rows, 10, m0)
    _, M_f = MicromagProblem(res, L, precision="fp32", **kw).run_simulation_batch(5e-11, 6, rows, 10, m0)
    d = np.abs(M_m - M_f).max()
    print(f"fp32_mfma vs fp32 batch: max |dm| {d:.2e}")
    assert d < 1e-5, d
    for b in (0, 5):
        err = np.abs(M_m[b] - _oracle_traj(oracle, res, L, m0[b], fields[b], 5e-11, 6, 10, **kw)).max()
        print(f"state {b}: max |m - oracle| {err:.2e}")
        assert err < 1e-4, (b, err)


@pytest.mark.parametrize("mode", [None, "threshold_fraction"])
def test_solver_fp32_mfma_single(oracle, mode):
    from magtense.micromag import MicromagProblem

    res, L = [16, 4, 1], [250e-9, 62.5e-9, 3e-9]
    rng = np.random.default_rng(45)
    m0 = rng.normal(size=(64, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    hfield = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0
    kw = dict(alpha=4.42e3, gamma=2.21e5, tol=1e-6)
    extra = dict(demag_approx=mode, dem_thres=0.8) if mode else {}
    p = MicromagProblem(res, L, m0=m0, precision="fp32_mfma", **kw, **extra)
    _, M, *_ = p.run_simulation(1e-10, 11, lambda t: np.ones((len(t), 1)) * hfield, 20)
    if mode:
        assert p.last_demag_info["form"] == "fp32 tiles", p.last_demag_info
        kw.update(demag_approx=4, dem_thres=0.8)
    err = np.abs(M[:, :, 0, :] - _oracle_traj(oracle, res, L, m0, hfield, 1e-10, 11, 20, **kw)).max()
    print(f"mode={mode}: max |m - oracle| {err:.2e}")
    assert err < 1e-4, err
