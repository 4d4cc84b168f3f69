"""Host-side logic of the drop-in API (no GPU): marshalling mirrors
python/src/magtense/micromag.py:46-268, and the tensor file format
MagTenseMicroMagPyIO.f90:221-271."""

import numpy as np
import pytest

from magtense import micromag
from magtense.micromag import MicromagProblem


def test_constructor_attributes_and_dtypes():
    p = MicromagProblem([4, 3, 2])
    assert p.ntot == 24
    assert p.grid_n.dtype == np.int32 and p.grid_n.flags.f_contiguous
    assert p.grid_L.dtype == np.float64 and list(p.grid_L) == [500e-9, 125e-9, 3e-9]
    assert p.u_ea.shape == (24, 3) and p.u_ea.flags.f_contiguous
    assert p.m0.shape == (24, 3)
    assert np.allclose(np.linalg.norm(p.m0, axis=1), 1.0)  # default: random unit vectors
    assert p.grid_type == 1 and p.solver == 2 and p.prob_mode == 1 and p.dem_appr == 1
    assert p.alphat.shape == (1, 2)
    assert p.conv_tol.shape == (1,) and p.t_conv.shape == (1,)
    assert p.N_load == 1 and p.N_ret == 1
    assert p.tol == 1e-4 and p.thres == 1e-6 and p.A0 == 1.3e-11 and p.Ms == 8e5


def test_enum_setters_match_reference_codes():
    p = MicromagProblem([2, 2, 1], solver="explicit", grid_type="unstructuredPrisms", prob_mode="old",
                        demag_approx="fft_threshold_fraction")
    assert (p.solver, p.grid_type, p.prob_mode, p.dem_appr) == (1, 3, 2, 5)
    with pytest.raises(KeyError):
        p.solver = "implicitly"


def test_m0_setter_forms():
    p = MicromagProblem([2, 2, 1], m0=1 / np.sqrt(3))
    assert np.allclose(p.m0, 1 / np.sqrt(3))
    m = np.arange(12.0).reshape(4, 3)
    p.m0 = m
    assert np.array_equal(p.m0, m)
    with pytest.raises(AssertionError):
        p.m0 = np.zeros((3, 3))


def test_unsupported_modes_raise_before_any_gpu_call():
    p = MicromagProblem([4, 4, 1], demag_approx="fft_thres")
    with pytest.raises(NotImplementedError):
        p._check_supported()
    MicromagProblem([4, 4, 1], demag_approx="threshold", dem_thres=1e-3)._check_supported()
    # the threshold modes run in both precisions (FP32 CSR, split-BF16 tiles)
    MicromagProblem([4, 4, 1], demag_approx="threshold_fraction", dem_thres=0.5, precision="bf16x2")._check_supported()
    p = MicromagProblem([1, 4, 1])
    with pytest.raises(ValueError):
        p._check_supported()


def test_precision_codes():
    from magtense_amd import _lib

    assert MicromagProblem([2, 2, 1])._precision_code() == _lib.PREC_FP32
    assert MicromagProblem([2, 2, 1], precision=True)._precision_code() == _lib.PREC_FP32
    assert MicromagProblem([2, 2, 1], precision="bf16x2")._precision_code() == _lib.PREC_BF16X2
    assert MicromagProblem([2, 2, 1], precision="fp32_mfma")._precision_code() == _lib.PREC_FP32_MFMA
    with pytest.raises(ValueError):
        MicromagProblem([2, 2, 1], precision="fp8")._precision_code()


def test_time_tables_match_reference_construction():
    p = MicromagProblem([2, 2, 1])
    t, h = p._time_tables(1e-9, 5, lambda t: np.stack([t, 2 * t, 3 * t], axis=1), 7)
    assert np.allclose(t, np.linspace(0, 1e-9, 5))
    assert h.shape == (7, 4) and h.flags.f_contiguous
    assert np.allclose(h[:, 0], np.linspace(0, 1e-9, 7)) and np.allclose(h[:, 3], 3 * h[:, 0])


def test_points_follow_setup_grid(oracle):
    p = MicromagProblem([4, 3, 2], [8e-9, 6e-9, 4e-9])
    pts, _ = oracle.setup_grid([4, 3, 2], [8e-9, 6e-9, 4e-9])
    assert np.allclose(p._pts(), pts)


def test_reference_layout_conversion():
    y = np.arange(2 * 5 * 3 * 4, dtype=float).reshape(2, 5, 3, 4)  # (n_h, nt, 3, N)
    M, H = MicromagProblem._to_reference_layout(None, y, None)
    assert M.shape == (5, 4, 2, 3) and M.flags.f_contiguous
    assert M[3, 1, 1, 2] == y[1, 3, 2, 1]
    assert all(h.shape == M.shape for h in H)


def test_tensor_file_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    K = rng.normal(size=(6, 5, 5)).astype(np.float32)
    f = tmp_path / "K.dat"
    micromag.save_demag_tensor(str(f), K)
    raw = np.fromfile(str(f), dtype=np.float32)
    assert raw.size == 150
    # record 2 (Kxy) is column-major: element (i, j) at offset j*n + i
    assert raw[25 + 3 * 5 + 1] == K[1, 1, 3]
    assert np.array_equal(micromag.load_demag_tensor(str(f), 5), K)
    with pytest.raises(ValueError):
        micromag.load_demag_tensor(str(f), 4)


def test_unstructured_prisms_marshalling(oracle):
    """grid_type='unstructuredPrisms': set_exchange_matrix writes the 1-based CSR
    the Fortran loader passes to mkl_sparse_d_create_csr (rows_start / rows_end,
    MagTenseMicroMagPyIO.f90:163-180); exch_val keeps float values; the points
    returned are grid_pts; tetrahedra need their nodes and elements, checked
    before any GPU call."""
    import scipy.sparse as sp

    res, L = [6, 3, 1], [30e-9, 9e-9, 3e-9]
    rowptr, col, val = oracle.exchange_csr(res, L)
    n = 18
    A = sp.csr_matrix((val, col, rowptr), shape=(n, n))
    p = MicromagProblem([n, 1, 1], grid_type="unstructuredPrisms")
    with pytest.raises(ValueError):  # no exchange operator yet
        p._check_supported()
    p.set_exchange_matrix(A)
    assert p.exch_nval == len(val) and p.exch_nrow == n
    assert np.array_equal(p.exch_rows, rowptr[:-1] + 1) and np.array_equal(p.exch_rowe, rowptr[1:] + 1)
    assert np.array_equal(p.exch_col, col + 1)
    assert p.exch_val.dtype == np.float64 and np.array_equal(p.exch_val, val)
    assert np.abs(val).max() > 1e15  # physical values that an int32 array cannot hold
    pts, _ = oracle.setup_grid(res, L)
    p.grid_pts[:] = pts
    p._check_supported()
    assert np.array_equal(p._pts(), pts)
    with pytest.raises(ValueError):
        p.set_exchange_matrix(sp.identity(n + 1))
    p.N_ave = np.array([0, 1, 1], dtype=np.int32)
    with pytest.raises(ValueError):
        p._check_supported()
    # tetrahedra: nodes and elements are required
    pt = MicromagProblem([4, 1, 1], grid_type="tetrahedron", grid_nnod=5)
    pt.set_exchange_matrix(sp.identity(4) * 0.0)
    with pytest.raises(ValueError):
        pt._check_supported()
    pt.grid_nod[:] = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]]) * 1e-9
    pt.grid_ele = np.array([[1, 2, 3, 4], [2, 3, 4, 5], [1, 2, 3, 5], [1, 3, 4, 5]], dtype=np.float64).T
    pt._check_supported()
