"""ctypes binding of libmagtense_hip.so (the C ABI declared in include/magtense_hip.h).

The library is built in-tree by ``magtense_amd/csrc/Makefile`` into
``magtense_amd/_lib/``.  There is no CPU fallback: every entry point here runs
on the MI355X, and loading fails loudly when the library or a GPU is missing.
"""

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmagtense_hip.so")
CSRC = os.path.join(_HERE, "csrc")

PREC_FP32 = 0
PREC_BF16X2 = 1
PREC_FP64 = 2
PREC_FP32_MFMA = 3
DEMAG_DENSE = 1
DEMAG_THRESHOLD = 2
DEMAG_THRESHOLD_FRACTION = 4

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class MtProblem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("n_states", C.c_int),
        ("Ms", _dp), ("K0", _dp), ("u_ea", _dp),
        ("A0", C.c_double), ("gamma", C.c_double), ("alpha0", C.c_double),
        ("nt_alpha", C.c_int), ("alphat", _dp),
        ("solver", C.c_int),
        ("nt_hext", C.c_int), ("hext", _dp), ("hext_per_state", C.c_int),
        ("nt", C.c_int), ("t", _dp),
        ("m0", _dp), ("m0_per_state", C.c_int),
        ("tol", C.c_double), ("thres", C.c_double),
        ("nt_conv", C.c_int), ("t_conv", _dp), ("conv_tol", C.c_double),
        ("precision", C.c_int),
        ("return_fields", C.c_int),
        ("max_steps", C.c_longlong),
        ("K_in", _fp), ("K_out", _fp),
        ("keep_fp32_tensor", C.c_int),
        ("ksplit", C.c_int),
        ("demag_approx", C.c_int),
        ("demag_threshold", C.c_double),
        ("toeplitz", C.c_int),
        ("grid_type", C.c_int),
        ("grid_pts", _dp), ("grid_abc", _dp),
        ("n_ave", C.c_int * 3),
        ("exch_nnz", C.c_int),
        ("exch_rows_start", C.POINTER(C.c_int)), ("exch_rows_end", C.POINTER(C.c_int)),
        ("exch_cols", C.POINTER(C.c_int)), ("exch_vals", _dp),
        ("grid_nnod", C.c_int), ("grid_nodes", _dp), ("grid_elements", C.POINTER(C.c_int)),
    ]


class MtStats(C.Structure):
    _fields_ = [
        ("f_calls", C.c_longlong), ("steps", C.c_longlong), ("bad_steps", C.c_longlong),
        ("flag3", C.c_longlong), ("flag4", C.c_longlong),
        ("h_last", C.c_double), ("t_reached", C.c_double),
        ("evals", C.c_longlong), ("ticks", C.c_longlong), ("demag_passes", C.c_longlong),
        ("demag_ms", C.c_double),
        ("vec_ms", C.c_double), ("llg_ms", C.c_double), ("vec_bytes", C.c_double), ("llg_bytes", C.c_double),
        ("vec_launches", C.c_longlong), ("llg_launches", C.c_longlong),
        ("near_ms", C.c_double),
    ]


class MtTiles(C.Structure):
    _fields_ = [
        ("n_tiles", C.c_int),
        ("size", _dp), ("offset", _dp), ("rot", _dp), ("M", _dp),
        ("tile_type", C.POINTER(C.c_int)), ("use_sym", C.POINTER(C.c_int)), ("sym_ops", _dp),
        ("vertices", _dp),
    ]


# every symbol declared in include/magtense_hip.h
EXPORTS = [
    "mt_last_error", "mt_version", "mt_device_count", "mt_measure_bandwidth", "mt_measure_bandwidth_pattern", "mt_prism_tensor", "mt_split_tensor_bytes",
    "mt_packed_m_bytes", "mt_split_tensor", "mt_f32_mfma_tensor", "mt_demag_apply_f32_mfma", "mt_demag_apply_f32", "mt_demag_apply_bf16x2", "mt_ll_rhs",
    "mt_solver_create", "mt_solver_advance", "mt_solver_run", "mt_solver_advance_total", "mt_solver_steps_total", "mt_solver_results", "mt_solver_state",
    "mt_solver_set_timing", "mt_solver_destroy", "mt_run_micromag", "mt_rk4_create", "mt_rk4_forward",
    "mt_rk4_backward", "mt_rk4_destroy", "mt_tiles_demag_tensor", "mt_tiles_h_field", "mt_demag_sparse_create",
    "mt_demag_sparse_info", "mt_demag_apply_sparse", "mt_demag_sparse_destroy", "mt_solver_demag_info", "mt_solver_demag_form",
    "mt_demag_sparse_create_bf16x2", "mt_demag_sparse_create_f32_tiles", "mt_demag_sparse_blocks", "mt_demag_near_correction",
    "mt_noalloc_scratch_bytes", "mt_demag_apply_bf16x2_noalloc_bytes", "mt_demag_apply_bf16x2_noalloc",
    "mt_toeplitz_tensor_bytes", "mt_toeplitz_tensor", "mt_demag_apply_bf16x2_toeplitz",
    "mt_split_tensor_row_offset", "mt_demag_apply_bf16x2_rows",
    "icl_initDemagMatrices", "icl_MVMult_GetH", "icl_destroy", "icl_initDemagMatrices_sparse", "icl_MVMult_GetH_sparse",
]

_lib = None


class MagTenseHipError(RuntimeError):
    """A non-zero status of the C ABI; ``code`` is the status (MT_ERR_* in
    include/magtense_hip.h: -1 argument, -2 HIP runtime, -3 solver, -4
    device allocation, -5 call order, -6 unsupported input)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def build():
    """Compile the HIP sources for gfx950 in-tree (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", CSRC, "-j8"], check=True)


def load():
    """Load the shared library (no GPU call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MagTenseHipError(
            f"{LIB_PATH} is missing: build it with `make -C {CSRC}` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    lib.mt_last_error.restype = C.c_char_p
    lib.mt_split_tensor_bytes.restype = C.c_longlong
    lib.mt_packed_m_bytes.restype = C.c_longlong
    vp = C.c_void_p
    lib.mt_measure_bandwidth.argtypes = [C.c_longlong, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.mt_measure_bandwidth_pattern.argtypes = [C.c_longlong, C.c_int, C.c_int, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double)]
    lib.mt_prism_tensor.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, vp]
    lib.mt_split_tensor.argtypes = [vp, C.c_int, vp]
    lib.mt_demag_apply_f32.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_demag_apply_bf16x2.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_f32_mfma_tensor.argtypes = [vp, C.c_int, vp]
    lib.mt_demag_apply_f32_mfma.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_ll_rhs.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                                                  vp, vp, vp, C.c_int, vp, vp]
    lib.mt_solver_create.argtypes = [C.POINTER(MtProblem), C.POINTER(vp)]
    lib.mt_solver_advance.argtypes = [vp, C.c_longlong]
    lib.mt_solver_run.argtypes = [vp]
    lib.mt_solver_advance_total.argtypes = [vp, C.c_longlong]
    lib.mt_solver_steps_total.argtypes = [vp]
    lib.mt_solver_steps_total.restype = C.c_longlong
    lib.mt_solver_results.argtypes = [vp, vp, vp, vp, vp]
    lib.mt_solver_state.argtypes = [vp, vp]
    lib.mt_solver_set_timing.argtypes = [vp, C.c_int]
    lib.mt_solver_destroy.argtypes = [vp]
    lib.mt_solver_destroy.restype = None
    lib.mt_run_micromag.argtypes = [C.POINTER(MtProblem), vp, vp, vp, vp]
    lib.mt_rk4_create.argtypes = [C.POINTER(MtProblem), C.c_double, C.c_longlong, C.c_int, C.POINTER(vp)]
    lib.mt_rk4_forward.argtypes = [vp, vp, vp]
    lib.mt_rk4_backward.argtypes = [vp, vp, vp, vp, vp]
    lib.mt_rk4_destroy.argtypes = [vp]
    lib.mt_rk4_destroy.restype = None
    lib.mt_tiles_demag_tensor.argtypes = [C.POINTER(MtTiles), C.c_int, vp, vp]
    lib.mt_tiles_h_field.argtypes = [C.POINTER(MtTiles), C.c_int, vp, vp, vp]
    lib.mt_toeplitz_tensor_bytes.argtypes = [C.c_int] * 3
    lib.mt_toeplitz_tensor_bytes.restype = C.c_longlong
    lib.mt_toeplitz_tensor.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [vp]
    lib.mt_noalloc_scratch_bytes.restype = C.c_longlong
    lib.mt_demag_apply_bf16x2_noalloc_bytes.argtypes = [C.c_int, C.c_int]
    lib.mt_demag_apply_bf16x2_noalloc_bytes.restype = C.c_longlong
    lib.mt_demag_apply_bf16x2_noalloc.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_demag_near_correction.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_demag_apply_bf16x2_toeplitz.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_split_tensor_row_offset.argtypes = [C.c_int, C.c_int]
    lib.mt_split_tensor_row_offset.restype = C.c_longlong
    lib.mt_demag_apply_bf16x2_rows.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.c_float]
    lib.mt_demag_sparse_create.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.POINTER(vp)]
    lib.mt_demag_sparse_info.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_float)]
    lib.mt_demag_sparse_create_bf16x2.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
    lib.mt_demag_sparse_create_f32_tiles.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(vp)]
    lib.mt_demag_sparse_blocks.argtypes = [vp] + [C.POINTER(C.c_longlong)] * 3 + [C.POINTER(C.c_int)]
    lib.mt_demag_apply_sparse.argtypes = [vp, vp, vp, C.c_int, C.c_float]
    lib.mt_demag_sparse_destroy.argtypes = [vp]
    lib.mt_demag_sparse_destroy.restype = None
    lib.mt_solver_demag_info.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_float)]
    lib.mt_solver_demag_form.argtypes = [vp, C.POINTER(C.c_int)]
    lib.icl_initDemagMatrices_sparse.argtypes = [C.POINTER(C.c_int)] * 3 + [vp, vp, vp]
    lib.icl_initDemagMatrices_sparse.restype = None
    lib.icl_MVMult_GetH_sparse.argtypes = [vp] * 6 + [C.POINTER(C.c_int), C.POINTER(C.c_float)]
    lib.icl_MVMult_GetH_sparse.restype = None
    lib.icl_initDemagMatrices.argtypes = [vp] * 6 + [C.POINTER(C.c_int)]
    lib.icl_initDemagMatrices.restype = None
    lib.icl_MVMult_GetH.argtypes = [vp] * 6 + [C.POINTER(C.c_int), C.POINTER(C.c_float)]
    lib.icl_MVMult_GetH.restype = None
    lib.icl_destroy.restype = None
    _lib = lib
    return lib


def lib():
    """The loaded library, after checking that a HIP device is visible."""
    L = load()
    if L.mt_device_count() <= 0:
        raise MagTenseHipError("no HIP device visible: the MI355X backend needs a GPU (there is no CPU fallback)")
    return L


def check(rc):
    if rc != 0:
        msg = load().mt_last_error()
        raise MagTenseHipError(f"libmagtense_hip error {rc}: {msg.decode() if msg else ''}", rc)
