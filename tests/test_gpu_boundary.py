"""The drop-in boundary on the GPU: the reference's dense plug-in ABI, the
demag tensor file as solver input, and the error paths a caller can force
from the host.

* icl_initDemagMatrices / icl_MVMult_GetH (MagTenseCudaBlasICLWrapper.cxx:39-47,
  MagTenseCudaBlas.cu:306-407): six column-major host N x N FP32 blocks
  (element (i, j) at i + j*N, the Fortran layout), H = pref * (K M) with the
  symmetric block pattern (Kyx = Kxy, ...) and beta = 0 (H overwritten), pref
  passed by reference.  Checked with non-symmetric blocks (so a transposed
  read fails) against the FP64 product: 1e-6 relative.
* Tensor file (MagTenseMicroMagPyIO.f90:221-271): a run that writes the file
  and a second run that loads it (K_in) give bit-identical trajectories, for
  the FP32 product and for the split-BF16 product on the dense tensor.
* Host-forcible failures: NaN in m0 and tol > 0.01 (RKSuitef90.f90:699) raise
  argument/solver errors; a tensor larger than the device raises
  MT_ERR_ALLOC; a plug-in product of the wrong size records an error and
  leaves H untouched.
"""

import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi
COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]


def _f(a):
    return a.ctypes.data_as(C.c_void_p)


def test_icl_dense_plugin_matches_fp64():
    from magtense_amd import _lib

    Lb = _lib.lib()
    rng = np.random.default_rng(17)
    for n in (96, 333):
        K = rng.normal(size=(6, n, n)).astype(np.float32)  # K[c][i, j]: receiver i, source j
        cols = [np.ascontiguousarray(K[c].T) for c in range(6)]  # column-major: (i, j) at i + j*n
        nn = C.c_int(n)
        Lb.icl_initDemagMatrices(*(_f(c) for c in cols), C.byref(nn))
        try:
            for pref in (-1.0, 2.5):
                M = rng.normal(size=(3, n)).astype(np.float32)
                H = np.full((3, n), 1e30, dtype=np.float32)  # overwritten (beta = 0)
                pr = C.c_float(pref)
                Lb.icl_MVMult_GetH(_f(M[0]), _f(M[1]), _f(M[2]), _f(H[0]), _f(H[1]), _f(H[2]), C.byref(nn),
                                   C.byref(pr))
                ref = np.zeros((3, n))
                for a in range(3):
                    for b in range(3):
                        ref[a] += K[COMP[a][b]].astype(np.float64) @ M[b].astype(np.float64)
                ref *= pref
                rel = np.linalg.norm(H - ref) / np.linalg.norm(ref)
                print(f"icl dense n={n} pref={pref}: rel {rel:.2e}")
                assert rel < 1e-6, rel
            # wrong size: an error is recorded and H is left untouched
            bad = C.c_int(n + 1)
            H2 = np.full((3, n + 1), 7.0, dtype=np.float32)
            M2 = np.ones((3, n + 1), dtype=np.float32)
            pr = C.c_float(1.0)
            Lb.icl_MVMult_GetH(_f(M2[0]), _f(M2[1]), _f(M2[2]), _f(H2[0]), _f(H2[1]), _f(H2[2]), C.byref(bad),
                               C.byref(pr))
            assert np.all(H2 == 7.0)
            assert b"icl_initDemagMatrices" in Lb.mt_last_error()
        finally:
            Lb.icl_destroy()


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
def test_tensor_file_reload_is_bit_identical(tmp_path, precision):
    from magtense.micromag import MicromagProblem, load_demag_tensor

    res, L = [24, 8, 1], [120e-9, 40e-9, 3e-9]
    n = 192
    rng = np.random.default_rng(3)
    m0 = rng.normal(size=(n, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    h = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0
    path = str(tmp_path / f"demag_{precision}.bin")
    runs = []
    for k in range(2):
        p = MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, tol=1e-6, precision=precision,
                            filename=path)
        p._toeplitz = -1  # dense tensor on both runs (the table is never built from a file)
        out = p.run_simulation(2e-11, 5, lambda t: np.ones((len(t), 1)) * h, 10)
        runs.append(out)
        if k == 0:
            assert os.path.getsize(path) == 6 * n * n * 4
            K_file = load_demag_tensor(path, n)
    assert np.array_equal(runs[0][1], runs[1][1])
    for a, b in zip(runs[0][3:], runs[1][3:]):
        assert np.array_equal(a, b)
    # the file holds the tensor of the grid: self term of a 5 x 5 x 3 nm prism, trace -1
    tr = K_file[0].diagonal() + K_file[3].diagonal() + K_file[5].diagonal()
    assert np.allclose(tr, -1.0, atol=1e-6)


def _small_problem(**kw):
    from magtense.micromag import MicromagProblem

    res, L = [8, 4, 1], [40e-9, 20e-9, 3e-9]
    rng = np.random.default_rng(1)
    m0 = rng.normal(size=(32, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    return MicromagProblem(res, L, m0=m0, alpha=4.42e3, gamma=2.21e5, **kw)


def _h(t):
    return np.ones((len(t), 1)) * np.array([1e4, 0.0, 0.0])


def test_nan_m0_is_refused():
    from magtense_amd._lib import MagTenseHipError

    p = _small_problem()
    p.m0[5, 1] = np.nan
    with pytest.raises(MagTenseHipError, match="NaN") as e:
        p.run_simulation(1e-11, 3, _h, 4)
    assert e.value.code == -1


def test_tolerance_above_rksuite_limit_is_refused():
    from magtense_amd._lib import MagTenseHipError

    p = _small_problem(tol=0.05)
    with pytest.raises(MagTenseHipError, match="tolerance") as e:
        p.run_simulation(1e-11, 3, _h, 4)
    assert e.value.code == -3


def test_tensor_larger_than_the_device_is_alloc_error():
    """N = 204800 cells: the FP32 tensor alone is 6 N^2 4 B = 1.0 TB, beyond
    the device's 288 GB; the allocation fails cleanly with MT_ERR_ALLOC and
    the next problem runs normally."""
    from magtense.micromag import MicromagProblem
    from magtense_amd._lib import MagTenseHipError

    p = MicromagProblem([512, 400, 1], [1e-6, 8e-7, 3e-9], m0=1 / np.sqrt(3), alpha=4.42e3)
    with pytest.raises(MagTenseHipError) as e:
        p.run_simulation(1e-12, 2, _h, 2)
    assert e.value.code == -4, str(e.value)
    q = _small_problem(tol=1e-6)
    _, M, *_ = q.run_simulation(1e-11, 3, _h, 4)
    assert np.all(np.isfinite(M))
