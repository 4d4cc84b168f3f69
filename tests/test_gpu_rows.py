"""Row blocks of the batched split-BF16 product (mt_demag_apply_bf16x2_rows):
the per-GPU product of a tensor sharded by receiver rows, fed the fragment rows
of its block.

* Each block equals the same rows of the whole-tensor product within 1e-6
  relative (norm-wise; the k-split, hence the FP32 summation order of the
  partial sums, differs with the block height) and the FP64 product of the
  same rows within 1e-5.
* Eight shards of a 4096-cell tensor reassemble the whole product.
* Ragged cases: n not a multiple of 32, a last block ending inside a tile, a
  one-row block, 1 / 5 / 64 states.
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


def _setup(oracle, res):
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
    return Lb, n, split, K64.cpu().numpy().reshape(6, n, n)


def _apply_rows(Lb, split, n, r0, nr, dm, S):
    import torch
    from magtense_amd import _lib

    off = Lb.mt_split_tensor_row_offset(n, r0)
    assert off >= 0
    h = torch.empty(S * 3 * nr, dtype=torch.float32, device="cuda")
    _lib.check(Lb.mt_demag_apply_bf16x2_rows(split.data_ptr() + off, n, r0, nr, dm.data_ptr(), h.data_ptr(), S, 1.0))
    return h.cpu().numpy().reshape(S, 3, nr).astype(np.float64)


def test_row_blocks_match_whole_product(oracle):
    import torch
    from magtense_amd import _lib

    Lb, n, split, K = _setup(oracle, [20, 9, 3])  # n = 540: last tile holds 28 rows
    for S in (1, 5, 64):
        rng = np.random.default_rng(100 + S)
        m = rng.normal(size=(S, 3, n)).astype(np.float32)
        dm = torch.as_tensor(m).cuda()
        hw = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        _lib.check(Lb.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), hw.data_ptr(), S, 1.0))
        whole = hw.cpu().numpy().reshape(S, 3, n).astype(np.float64)
        for r0, nr in ((0, 540), (0, 32), (32, 100), (256, 1), (512, 28), (96, 444)):
            got = _apply_rows(Lb, split, n, r0, nr, dm, S)
            ref = _fp64_rows(K, m.astype(np.float64), r0, r0 + nr)
            e_whole = np.linalg.norm(got - whole[:, :, r0:r0 + nr]) / np.linalg.norm(whole[:, :, r0:r0 + nr])
            e_64 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert e_whole < 1e-6, (S, r0, nr, e_whole)
            assert e_64 < 1e-5, (S, r0, nr, e_64)
    # blocks that do not fit are rejected before any launch
    for r0, nr in ((0, n + 1), (512, 29), (16, 32), (-32, 32)):
        with pytest.raises(_lib.MagTenseHipError):
            _lib.check(Lb.mt_demag_apply_bf16x2_rows(split.data_ptr(), n, r0, nr, dm.data_ptr(), hw.data_ptr(), 1, 1.0))


def test_eight_shards_reassemble(oracle):
    import torch
    from magtense_amd import _lib

    Lb, n, split, K = _setup(oracle, [64, 32, 2])  # n = 4096, shards of 512 rows
    S = 64
    rng = np.random.default_rng(8)
    m = rng.normal(size=(S, 3, n)).astype(np.float32)
    dm = torch.as_tensor(m).cuda()
    hw = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
    _lib.check(Lb.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), hw.data_ptr(), S, 1.0))
    whole = hw.cpu().numpy().reshape(S, 3, n).astype(np.float64)
    shards = np.concatenate([_apply_rows(Lb, split, n, g * 512, 512, dm, S) for g in range(8)], axis=2)
    e = np.linalg.norm(shards - whole) / np.linalg.norm(whole)
    rows = np.sort(np.random.default_rng(9).choice(n, 256, replace=False))
    ref = np.zeros((S, 3, rows.size))
    m64 = m.astype(np.float64)
    for a in range(3):
        for b in range(3):
            ref[:, a] += m64[:, b] @ K[COMP[a][b]][rows].T
    e64 = np.linalg.norm(shards[:, :, rows] - ref) / np.linalg.norm(ref)
    print(f"8 shards vs whole {e:.2e}, vs FP64 on 256 rows {e64:.2e}")
    assert e < 1e-6 and e64 < 1e-5


@pytest.mark.parametrize("res,row0,nrows", [([64, 32, 2], 1536, 512), ([20, 9, 3], 512, 28), ([20, 9, 3], 96, 100)])
def test_assemble_row_block(oracle, res, row0, nrows):
    """A shard assembles only its receivers' fragment rows (pts = its cells):
    byte-identical to the same rows of the whole split tensor, padded rows zero."""
    import torch
    from magtense_amd import _lib

    Lb, n, split, _ = _setup(oracle, res)
    L = [res[0] * 2e-9, res[1] * 2.5e-9, res[2] * 1.5e-9]
    pts, d = oracle.setup_grid(res, L)
    np32 = (n + 31) // 32 * 32
    nbytes = (nrows + 31) // 32 * 32 * np32 * 24
    blk = torch.full((nbytes,), 0x55, dtype=torch.uint8, device="cuda")
    dp = torch.as_tensor(pts).cuda()
    dsub = torch.as_tensor(np.ascontiguousarray(pts[row0:row0 + nrows])).cuda()
    da = torch.as_tensor(d).cuda()
    _lib.check(Lb.mt_prism_tensor(dp.data_ptr(), da.data_ptr(), 0, None, dsub.data_ptr(), n, nrows, None, None,
                                  blk.data_ptr()))
    off = Lb.mt_split_tensor_row_offset(n, row0)
    whole = split[off:off + nbytes]
    if row0 + nrows == n or nrows % 32 == 0:
        assert torch.equal(blk, whole)
    else:  # the block's last tile: its rows past nrows are zero in the block, cells row0+nrows.. in the whole
        b = blk.cpu().numpy().reshape(-1, 64, 16)
        w = whole.cpu().numpy().reshape(-1, 64, 16)
        lane_row = np.arange(64) % 32
        last = (nrows - 1) // 32
        nJ = np32 // 16
        frag_tile = np.arange(b.shape[0]) // (nJ * 12)
        live = (frag_tile[:, None] < last) | (lane_row[None, :] < nrows - 32 * last)
        assert np.array_equal(b[live], w[live])
        assert not b[~live].any()
