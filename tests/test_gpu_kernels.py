"""GPU parity of the individual kernels against the FP64 oracle (C ABI calls).

Tolerances:
  * tensor (FP64 output): |K_gpu - K_oracle| <= 1e-12 * max|K| (libm ulp
    differences amplified by the far-field cancellation);
  * tensor (FP32 output): within 2 ulp of the oracle's FP64 value rounded;
  * demag FP32 product: 1e-5 relative (norm-wise) to the FP64 product;
  * demag split-BF16: 1e-5 relative (norm-wise) to the FP64 product;
  * right-hand side: 1e-9 relative (FP64 terms; FP32 demag term bounded by
    the FP32 rounding of the product).
"""

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _lib():
    from magtense_amd import _lib

    return _lib.lib(), _lib


def _dev(a, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def grid_tensor_gpu(res, L, want64=False, want_split=False):
    from oracle import oracle as O

    L_, lib = _lib()
    pts, d = O.setup_grid(res, L)
    n = pts.shape[0]
    dp = _dev(pts)
    dabc = _dev(d)
    K32 = torch.empty(6 * n * n, dtype=torch.float32, device="cuda")
    K64 = torch.empty(6 * n * n, dtype=torch.float64, device="cuda") if want64 else None
    split = torch.empty(L_.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda") if want_split else None
    rc = L_.mt_prism_tensor(dp.data_ptr(), dabc.data_ptr(), 0, None, dp.data_ptr(), n, n, K32.data_ptr(),
                            K64.data_ptr() if want64 else None, split.data_ptr() if want_split else None)
    lib.check(rc)
    return n, K32, K64, split


def test_prism_tensor_random_pairs(oracle):
    L_, lib = _lib()
    rng = np.random.default_rng(11)
    ns, npt = 37, 53
    src = rng.uniform(-5, 5, size=(ns, 3))
    abc = rng.uniform(0.5, 2.0, size=(ns, 3))
    pts = rng.uniform(-8, 8, size=(npt, 3))
    # guard cases: points exactly on face planes / edge lines of source 0
    pts[0] = src[0] + np.array([abc[0, 0] / 2, 0.3, 0.2])
    pts[1] = src[0] + np.array([abc[0, 0] / 2, abc[0, 1] / 2, 0.7])
    pts[2] = src[0]  # self point
    pts[3] = src[0] + np.array([0.0, 0.0, 5.0])
    ref = oracle.tensor_points(src, abc, pts)
    out64 = torch.empty(6 * npt * ns, dtype=torch.float64, device="cuda")
    out32 = torch.empty(6 * npt * ns, dtype=torch.float32, device="cuda")
    ds, da, dpt = _dev(src), _dev(abc), _dev(pts)
    lib.check(L_.mt_prism_tensor(ds.data_ptr(), da.data_ptr(), 3, None, dpt.data_ptr(), ns, npt, out32.data_ptr(),
                                 out64.data_ptr(), None))
    g64 = out64.cpu().numpy().reshape(6, npt, ns)
    g32 = out32.cpu().numpy().reshape(6, npt, ns)
    scale = np.abs(ref).max()
    assert np.all(np.isfinite(g64))
    assert np.abs(g64 - ref).max() <= 1e-12 * scale
    r32 = ref.astype(np.float32)
    ulp = np.spacing(np.abs(r32).astype(np.float32))
    assert np.all(np.abs(g32 - r32) <= 2 * ulp + 1e-30)


def test_tensor_uniform_grid_matches_oracle(oracle):
    res, L = [16, 8, 2], [80e-9, 40e-9, 6e-9]
    n, K32, K64, _ = grid_tensor_gpu(res, L, want64=True)
    ref = oracle.tensor_uniform(res, L)
    g = K64.cpu().numpy().reshape(6, n, n)
    assert np.abs(g - ref).max() <= 1e-12
    assert np.array_equal(K32.cpu().numpy().reshape(6, n, n) != 0, ref.astype(np.float32) != 0)


def _demag_ref(K, m):
    # K (6,n,n) float64, m (S,3,n) -> H = -K m  (S,3,n)
    comp = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    S, _, n = m.shape
    H = np.zeros((S, 3, n))
    for a in range(3):
        for b in range(3):
            H[:, a] -= np.einsum("ij,sj->si", K[comp[a][b]], m[:, b])
    return H


@pytest.mark.parametrize("res", [[32, 8, 1], [64, 32, 1], [16, 16, 4]])
def test_demag_fp32_and_bf16x2(oracle, res):
    L = [res[0] * 2e-9, res[1] * 2e-9, res[2] * 1.5e-9]
    L_, lib = _lib()
    n, K32, _, split = grid_tensor_gpu(res, L, want_split=True)
    S = 64
    rng = np.random.default_rng(3)
    m = rng.normal(size=(S, 3, n))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    m32 = m.astype(np.float32)
    Kf = K32.cpu().numpy().reshape(6, n, n).astype(np.float64)
    ref = _demag_ref(Kf, m32.astype(np.float64))
    dm = _dev(m32)
    h = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
    lib.check(L_.mt_demag_apply_f32(K32.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
    hf = h.cpu().numpy().reshape(S, 3, n).astype(np.float64)
    rel = np.linalg.norm(hf - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel
    h2 = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
    lib.check(L_.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), h2.data_ptr(), S, -1.0))
    hb = h2.cpu().numpy().reshape(S, 3, n).astype(np.float64)
    relb = np.linalg.norm(hb - ref) / np.linalg.norm(ref)
    print(f"res={res} fp32 rel={rel:.2e} bf16x2 rel={relb:.2e}")
    assert relb < 1e-5, relb


def test_demag_bf16x2_partial_batch(oracle):
    """Fewer than 32 states (single column block) and a ragged cell count."""
    res, L = [20, 6, 1], [40e-9, 12e-9, 3e-9]
    L_, lib = _lib()
    n, K32, _, split = grid_tensor_gpu(res, L, want_split=True)
    for S in (1, 5, 33):
        rng = np.random.default_rng(S)
        m = rng.normal(size=(S, 3, n)).astype(np.float32)
        ref = _demag_ref(K32.cpu().numpy().reshape(6, n, n).astype(np.float64), m.astype(np.float64))
        dm = _dev(m)
        h = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        lib.check(L_.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
        hb = h.cpu().numpy().reshape(S, 3, n)
        rel = np.linalg.norm(hb - ref) / np.linalg.norm(ref)
        assert rel < 1e-5, (S, rel)


def test_ll_rhs_matches_oracle(oracle):
    res, L = [12, 6, 2], [60e-9, 30e-9, 6e-9]
    L_, lib = _lib()
    n, K32, _, _ = grid_tensor_gpu(res, L)
    rng = np.random.default_rng(5)
    m = rng.normal(size=(3, n))
    m /= np.linalg.norm(m, axis=0, keepdims=True)
    ms = rng.uniform(7e5, 9e5, size=n)
    k0 = rng.uniform(0, 5e4, size=n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    hx = np.array([1.0e4, -2.0e4, 3.0e3])
    q = oracle.Problem(res, L, Ms=ms, K0=k0, u_ea=u, gamma=2.21e5, alpha=4.42e3,
                       hext=np.array([[0.0, *hx], [1.0, *hx]]))
    ref, Href = q.dmdt(0.0, m.reshape(-1))
    dm = _dev(m.reshape(-1))
    out = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    fields = torch.empty(12 * n, dtype=torch.float64, device="cuda")
    msd, k0d, ud, hd = (np.ascontiguousarray(v) for v in (ms, k0, u, hx))
    lib.check(L_.mt_ll_rhs(res[0], res[1], res[2], L[0], L[1], L[2], msd.ctypes.data, k0d.ctypes.data,
                           ud.ctypes.data, 1.3e-11, 2.21e5, 4.42e3, K32.data_ptr(), dm.data_ptr(), hd.ctypes.data, 1,
                           out.data_ptr(), fields.data_ptr()))
    g = out.cpu().numpy()
    F = fields.cpu().numpy().reshape(4, 3 * n)
    for k in (0, 1, 3):
        assert np.allclose(F[k], Href[k], rtol=1e-12, atol=1e-12 * np.abs(Href[k]).max())
    # demag term: FP32 summation order differs from the oracle's correctly rounded sums
    assert np.abs(F[2] - Href[2]).max() <= 1e-5 * np.abs(Href[2]).max()
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()
