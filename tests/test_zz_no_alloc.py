"""The allocation-free split-BF16 product (mt_demag_apply_bf16x2_noalloc,
workspace in the library's file-scope device array) against the allocating
entry point mt_demag_apply_bf16x2 on the same tensor and states: the same
kernels in the same order, so the fields are bit-identical; and within 1e-5
(norm-wise) of the FP64 product.  Two calls in a row, the second with a
smaller input (fewer cells and states, so the array still holds the first
call's larger contents), and a size beyond the array refused before any
launch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _fp64_product(K, m):
    S, _, n = m.shape
    out = np.zeros((S, 3, n))
    for a in range(3):
        for b in range(3):
            out[:, a] -= m[:, b] @ K[COMP[a][b]].T
    return out


def test_noalloc_product_matches_allocating_product(oracle):
    import torch
    from magtense_amd import _lib

    L = _lib.lib()
    for res, S in (([32, 32, 2], 70), ([24, 7, 1], 5)):  # 70 states: two 64-column groups
        dims = [res[0] * 2e-9, res[1] * 2e-9, res[2] * 1.5e-9]
        pts, d = oracle.setup_grid(res, dims)
        n = pts.shape[0]
        assert 0 < L.mt_demag_apply_bf16x2_noalloc_bytes(n, S) <= L.mt_noalloc_scratch_bytes()
        dp = torch.as_tensor(pts).cuda()
        da = torch.as_tensor(d).cuda()
        split = torch.empty(L.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda")
        _lib.check(L.mt_prism_tensor(dp.data_ptr(), da.data_ptr(), 0, None, dp.data_ptr(), n, n, None, None,
                                     split.data_ptr()))
        rng = np.random.default_rng(n + S)
        m = rng.normal(size=(S, 3, n))
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        m32 = m.astype(np.float32)
        dm = torch.as_tensor(m32).cuda()
        h_ref = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        h_na = torch.full((S * 3 * n,), float("nan"), dtype=torch.float32, device="cuda")
        _lib.check(L.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), h_ref.data_ptr(), S, -1.0))
        _lib.check(L.mt_demag_apply_bf16x2_noalloc(split.data_ptr(), n, dm.data_ptr(), h_na.data_ptr(), S, -1.0))
        a = h_ref.cpu().numpy()
        b = h_na.cpu().numpy()
        ref = _fp64_product(oracle.tensor_uniform(res, dims), m32.astype(np.float64))
        rel = np.linalg.norm(b.reshape(S, 3, n) - ref) / np.linalg.norm(ref)
        print(f"noalloc n={n} S={S}: identical {np.array_equal(a, b)}, vs FP64 {rel:.2e}")
        assert np.array_equal(a, b)
        assert rel < 1e-5


def test_noalloc_refuses_sizes_beyond_the_array():
    import torch
    from magtense_amd import _lib

    L = _lib.lib()
    n = 200000
    assert L.mt_demag_apply_bf16x2_noalloc_bytes(n, 64) > L.mt_noalloc_scratch_bytes()
    dummy = torch.zeros(16, dtype=torch.float32, device="cuda")
    # rejected on the host before any launch: the pointers are never read
    rc = L.mt_demag_apply_bf16x2_noalloc(dummy.data_ptr(), n, dummy.data_ptr(), dummy.data_ptr(), 64, -1.0)
    assert rc == -1
    assert b"scratch" in L.mt_last_error()
