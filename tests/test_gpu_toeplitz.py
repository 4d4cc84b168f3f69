"""The translation-invariant split-BF16 table (uniform grids, nx % 32 == 0)
against the dense fragment-tiled tensor and the FP64 product.

Grids with nx % 128 == 0 take the entry-reuse kernel (k-split 1 or 2 here,
and 16 on 128 x 16 x 1, whose k-split does not divide the eight-slot ring
and takes the kernel's second work mapping), the others the four-tile
kernel.  The table holds the same FP32 values split
the same way; the products differ
only by the order in which k-chunks are summed (the table kernel visits them by
cell offset) and by ulp-level FP64 differences in the offsets (bar 1e-6
relative, norm-wise); both stay within 1e-5 of the FP64 product.  The solver
with the table follows the dense-tensor run to the integration tolerance
(1e-4 on unit magnetization, the bar used against the oracle).
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _fp64_product(K, m):
    S, _, n = m.shape
    out = np.zeros((S, 3, n))
    for a in range(3):
        for b in range(3):
            out[:, a] += m[:, b] @ K[COMP[a][b]].T
    return out


@pytest.mark.parametrize("res", [[32, 8, 1], [64, 4, 3], [32, 2, 2], [96, 3, 1], [128, 4, 2], [256, 3, 3], [128, 2, 3],
                                 [128, 16, 1]])
def test_table_matches_dense_split(oracle, res):
    import torch
    from magtense_amd import _lib

    Lb = _lib.lib()
    L = [res[0] * 2e-9, res[1] * 2.5e-9, res[2] * 1.5e-9]
    pts, d = oracle.setup_grid(res, L)
    n = pts.shape[0]
    dp = torch.as_tensor(pts).cuda()
    da = torch.as_tensor(d).cuda()
    K64 = torch.empty(6 * n * n, dtype=torch.float64, device="cuda")
    split = torch.empty(Lb.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda")
    _lib.check(Lb.mt_prism_tensor(dp.data_ptr(), da.data_ptr(), 0, None, dp.data_ptr(), n, n, None, K64.data_ptr(),
                                  split.data_ptr()))
    nbytes = Lb.mt_toeplitz_tensor_bytes(*res)
    assert nbytes == (2 * res[2] - 1) * (2 * res[1] - 1) * (res[0] // 8 - 2) * 12 * 1024
    tab = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    _lib.check(Lb.mt_toeplitz_tensor(*res, *L, tab.data_ptr()))
    K = K64.cpu().numpy().reshape(6, n, n)
    for S in (1, 5, 64):
        rng = np.random.default_rng(S)
        m = rng.normal(size=(S, 3, n)).astype(np.float32)
        dm = torch.as_tensor(m).cuda()
        hd = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        ht = torch.empty_like(hd)
        _lib.check(Lb.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), hd.data_ptr(), S, 1.0))
        _lib.check(Lb.mt_demag_apply_bf16x2_toeplitz(tab.data_ptr(), *res, dm.data_ptr(), ht.data_ptr(), S, 1.0))
        a = hd.cpu().numpy().reshape(S, 3, n).astype(np.float64)
        b = ht.cpu().numpy().reshape(S, 3, n).astype(np.float64)
        ref = _fp64_product(K, m.astype(np.float64))
        nr = np.linalg.norm(ref)
        print(f"res={res} S={S}: table vs dense {np.linalg.norm(a - b) / nr:.2e}, "
              f"table vs fp64 {np.linalg.norm(b - ref) / nr:.2e}, identical {np.array_equal(a, b)}")
        assert np.linalg.norm(a - b) / nr < 1e-6
        assert np.linalg.norm(b - ref) / nr < 1e-5
    with pytest.raises(Exception):
        _lib.check(Lb.mt_toeplitz_tensor(48, 2, 1, 1e-7, 1e-8, 1e-9, tab.data_ptr()))


def test_solver_table_equals_dense_run():
    from magtense.micromag import MicromagProblem

    res, L = [32, 4, 1], [250e-9, 31.25e-9, 3e-9]
    rng = np.random.default_rng(8)
    m0 = rng.normal(size=(128, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    h = np.array([-24.6, 4.3, 0.0]) / 1000 / (4e-7 * np.pi)
    out = {}
    for mode in (0, -1):
        p = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6, precision="bf16x2")
        p._toeplitz = mode
        _, M, *_ = p.run_simulation(5e-11, 6, lambda t: np.ones((len(t), 1)) * h, 10)
        out[mode] = (M, p.last_stats[0])
    err = np.abs(out[0][0] - out[-1][0]).max()
    print("table vs dense solver run: max diff", err, out[0][1], out[-1][1])
    assert err < 1e-4
