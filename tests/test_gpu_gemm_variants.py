"""The alternative split-BF16 GEMM kernels kept for measurement
(MT_GEMM_VARIANT, read at every launch) give the default kernel's product.

* Dense tensor: variant 6 (two-tile kernel, four-stage LDS DMA ring) on the
  whole tensor and on row blocks, n = 540 (the last tile holds 28 rows).
* Table: variant 2 (two-tile four-stage kernel) and variant 4 (register-ring
  kernel) on a 64 x 4 x 3 grid; variant 3 (the four-tile kernel) against the
  default entry-reuse kernel on 128 x 4 x 2 and 256 x 2 x 3 grids.

Each is within 1e-6 relative (norm-wise; the kernels differ in k-split and
k-chunk order, hence in FP32 summation order) of the default kernel and within
1e-5 of the FP64 product, for 1, 5 and 64 states.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _fp64_rows(K, m, r0, r1):
    S = m.shape[0]
    out = np.zeros((S, 3, r1 - r0))
    for a in range(3):
        for b in range(3):
            out[:, a] += m[:, b] @ K[COMP[a][b]][r0:r1].T
    return out


def _tensors(oracle, res, table):
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
    tab = None
    if table:
        tab = torch.empty(Lb.mt_toeplitz_tensor_bytes(*res), dtype=torch.uint8, device="cuda")
        _lib.check(Lb.mt_toeplitz_tensor(*res, *L, tab.data_ptr()))
    return Lb, n, split, tab, K64.cpu().numpy().reshape(6, n, n)


def _product(Lb, K, n, dm, S, r0=0, nr=None, table_res=None):
    import torch
    from magtense_amd import _lib

    nr = n if nr is None else nr
    h = torch.empty(S * 3 * nr, dtype=torch.float32, device="cuda")
    if table_res is not None:
        _lib.check(Lb.mt_demag_apply_bf16x2_toeplitz(K.data_ptr(), *table_res, dm.data_ptr(), h.data_ptr(), S, 1.0))
    elif nr < n:
        off = Lb.mt_split_tensor_row_offset(n, r0)
        assert off >= 0
        _lib.check(Lb.mt_demag_apply_bf16x2_rows(K.data_ptr() + off, n, r0, nr, dm.data_ptr(), h.data_ptr(), S, 1.0))
    else:
        _lib.check(Lb.mt_demag_apply_bf16x2(K.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, 1.0))
    return h.cpu().numpy().reshape(S, 3, nr).astype(np.float64)


def test_dense_lds_ring_variant(oracle, monkeypatch):
    import torch

    Lb, n, split, _, K = _tensors(oracle, [20, 9, 3], table=False)
    for S in (1, 5, 64):
        m = np.random.default_rng(40 + S).normal(size=(S, 3, n)).astype(np.float32)
        dm = torch.as_tensor(m).cuda()
        for r0, nr in ((0, n), (0, 32), (96, 444), (512, 28)):
            monkeypatch.delenv("MT_GEMM_VARIANT", raising=False)
            base = _product(Lb, split, n, dm, S, r0, nr)
            monkeypatch.setenv("MT_GEMM_VARIANT", "6")
            got = _product(Lb, split, n, dm, S, r0, nr)
            ref = _fp64_rows(K, m.astype(np.float64), r0, r0 + nr)
            e_base = np.linalg.norm(got - base) / np.linalg.norm(base)
            e_64 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            print(f"dense variant 6 S={S} rows {r0}+{nr}: vs default {e_base:.2e}, vs FP64 {e_64:.2e}")
            assert e_base < 1e-6 and e_64 < 1e-5, (S, r0, nr, e_base, e_64)


@pytest.mark.parametrize("variant,res", [("2", [64, 4, 3]), ("4", [64, 4, 3]), ("3", [128, 4, 2]), ("3", [256, 2, 3])])
def test_table_variants(oracle, monkeypatch, variant, res):
    import torch

    Lb, n, _, tab, K = _tensors(oracle, res, table=True)
    for S in (1, 5, 64):
        m = np.random.default_rng(60 + S).normal(size=(S, 3, n)).astype(np.float32)
        dm = torch.as_tensor(m).cuda()
        monkeypatch.delenv("MT_GEMM_VARIANT", raising=False)
        base = _product(Lb, tab, n, dm, S, table_res=res)
        monkeypatch.setenv("MT_GEMM_VARIANT", variant)
        got = _product(Lb, tab, n, dm, S, table_res=res)
        ref = _fp64_rows(K, m.astype(np.float64), 0, n)
        e_base = np.linalg.norm(got - base) / np.linalg.norm(base)
        e_64 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"table variant {variant} S={S}: vs default {e_base:.2e}, vs FP64 {e_64:.2e}")
        assert e_base < 1e-6 and e_64 < 1e-5, (S, e_base, e_64)
