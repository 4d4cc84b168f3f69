"""Drop-in ``MicromagProblem`` running on the MI355X.

Mirrors ``magtense.micromag.MicromagProblem`` (python/src/magtense/micromag.py:8-268):
the same constructor keywords and defaults, the same mutable attributes and
dtypes, and ``run_simulation(t_end, nt, fct_h_ext, nt_h_ext)`` returning the
same 7-tuple ``(t_out, M_out, pts, H_exc, H_ext, H_dem, H_ani)``.  The call
goes to the C ABI of ``libmagtense_hip.so`` (ctypes) instead of the f2py
module ``magtensesource``.

Additive extensions (no reference counterpart):
  * ``precision='fp32' | 'bf16x2' | 'fp32_mfma'`` selects the demag product:
    the plain per-state FP32 mat-vec (reference semantics, the default), the
    split-BF16 batched MFMA GEMM, or the batched FP32 MFMA GEMM (FP32
    products and sums, the reference's arithmetic class, one tensor pass for
    up to 64 states).  With ``demag_approx`` 'threshold' or
    'threshold_fraction' the kept entries are an FP32 CSR product ('fp32',
    up to four states) or a GEMM over the 32 x 16 tiles that hold a kept
    entry (FP32 tiles: 'fp32' from five states and 'fp32_mfma'; split-BF16
    tiles: 'bf16x2').  A boolean ``precision`` (the reference's unused
    ``usePrecision`` flag) keeps the default.
  * ``Ms`` may be an array of per-cell values.
  * ``run_simulation_batch`` integrates several independent states (and
    applied fields) in one call; every right-hand side of the batch shares one
    pass over the demag tensor.
  * ``solver='explicit'`` returns all field rows, shape (nt, N, nt_Hext, 3)
    (the reference's f2py wrapper declares only one row).
  * ``set_exchange_matrix(A)`` fills the exch_* arrays of a non-uniform grid
    from a sparse or dense operator (the Matlab front end's
    ``setExchangeMatrixSparse``).

Behaviour that differs on purpose: ``cuda`` is accepted but ignored (this
backend always runs on the GPU and raises when none is visible); the FFT
``demag_approx`` modes ('fft_thres', 'fft_threshold_fraction') raise
``NotImplementedError`` ('threshold' and
'threshold_fraction' build the CSR tensor on the device); the demag tensor
file (``filename`` longer than two characters) is loaded when it exists and
written otherwise.  ``grid_type='unstructuredPrisms'`` (res = [N, 1, 1],
``grid_pts`` / ``grid_abc`` per cell, ``N_ave`` receiver averaging, the
exchange operator in ``exch_rows`` / ``exch_rowe`` / ``exch_col`` /
``exch_val`` as a 1-based CSR) runs as the reference's Fortran does
(MagTenseMicroMagPyIO.f90:76-85,163-180); ``exch_val`` is float64 here
because the reference's int32 array (micromag.py:127) truncates every
physical exchange coefficient (1/d^2 ~ 1e16 m^-2).  ``grid_type='tetrahedron'``
(res = [N, 1, 1], ``grid_nnod`` nodes in ``grid_nod`` (grid_nnod x 3),
``grid_ele`` (4 x N) 1-based node indices per element, ``grid_pts`` the
evaluation points, the same exchange arrays) assembles the tetrahedron tensor
(LandauLifshitzEquationSolver.f90:988-1041) with the element's vertices in
global coordinates; ``grid_nod`` is read as N x 3 coordinates (the
reference's loader assigns the (grid_nnod, 3) array to a (3, grid_nnod) one,
MagTenseMicroMagPyIO.f90:67-70, which reorders its memory instead of
transposing it).
"""

import ctypes as C
import os
from typing import List, Optional, Union

import numpy as np

from magtense_amd import _lib

MU0 = 4 * np.pi * 1e-7


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class MicromagProblem:
    """Micromagnetic problem on a uniform grid or a grid of unstructured prisms
    (reference: micromag.py:8-146)."""

    def __init__(
        self,
        res: List[int],
        grid_L: List[int] = [500e-9, 125e-9, 3e-9],
        grid_nnod: int = 0,
        grid_type: Optional[str] = "uniform",
        prob_mode: Optional[str] = "new",
        solver: Optional[str] = "dynamic",
        m0: Union[None, int, float, List, np.ndarray] = None,
        A0: float = 1.3e-11,
        Ms: float = 8e5,
        K0: float = 0.0,
        alpha: float = 0.02,
        gamma: float = 0.0,
        max_T0: float = 2.0,
        nt_conv: int = 1,
        conv_tol: float = 1e-4,
        tol: float = 1e-4,
        thres: float = 1e-6,
        setTimeDis: int = 10,
        dem_thres: float = 0.0,
        demag_approx: Optional[str] = None,
        CV: float = 0.0,
        ReturnHall: int = 0,
        exch_nval: int = 1,
        exch_nrow: int = 1,
        filename: str = "t",
        cuda: bool = False,
        cvode: bool = False,
        precision: Union[bool, str] = False,
        n_threads: int = 1,
        N_ave: List[int] = [1, 1, 1],
        t_alpha: np.ndarray = np.zeros(1),
        alpha_fct=lambda t: np.atleast_2d(t).T * 0,
    ) -> None:
        ntot = int(np.prod(res))
        self.ntot = ntot
        self.grid_nnod = grid_nnod
        self.nt_conv = nt_conv
        self.exch_nval = exch_nval
        self.exch_nrow = exch_nrow

        self.grid_n = np.array(res, dtype=np.int32, order="F")
        self.grid_L = np.array(grid_L, dtype=np.float64, order="F")
        self.grid_pts = np.zeros(shape=(ntot, 3), dtype=np.float64, order="F")
        self.grid_ele = np.zeros(shape=(4, ntot), dtype=np.float64, order="F")
        self.grid_nod = np.zeros(shape=(grid_nnod, 3), dtype=np.float64, order="F")
        self.grid_abc = np.zeros(shape=(ntot, 3), dtype=np.float64, order="F")
        self.u_ea = np.zeros(shape=(ntot, 3), dtype=np.float64, order="F")

        self.grid_type = grid_type
        self.prob_mode = prob_mode
        self.solver = solver

        self.m0 = m0

        self.A0 = A0
        self.Ms = Ms
        self.K0 = K0
        self.alpha_mm = alpha
        self.gamma = gamma
        self.max_T0 = max_T0

        self.t_conv = np.zeros(shape=(nt_conv), dtype=np.float64, order="F")
        self.conv_tol = np.array(np.repeat(conv_tol, nt_conv), dtype=np.float64, order="F")
        self.tol = tol
        self.thres = thres
        self.setTimeDis = setTimeDis

        self.dem_thres = dem_thres
        self.dem_appr = demag_approx
        self.CV = CV
        self.ReturnHall = ReturnHall

        self.nt_alpha = len(t_alpha)
        self.alphat = np.zeros(shape=(self.nt_alpha, 2), dtype=np.float64, order="F")
        self.alphat[:, 0] = t_alpha
        self.alphat[:, 1] = np.asarray(alpha_fct(t_alpha)).reshape(-1)

        self.exch_val = np.zeros(shape=(exch_nval), dtype=np.float64, order="F")
        self.exch_rows = np.zeros(shape=(exch_nrow), dtype=np.int32, order="F")
        self.exch_rowe = np.zeros(shape=(exch_nrow), dtype=np.int32, order="F")
        self.exch_col = np.zeros(shape=(exch_nval), dtype=np.int32, order="F")

        self.N_load = len(filename)
        self.N_file_in = filename
        self.N_ret = len(filename)
        self.N_file_out = filename

        self.cuda = int(cuda)
        self.cvode = int(cvode)
        self.precision = precision
        self.n_threads = n_threads
        self.N_ave = np.array(N_ave, dtype=np.int32, order="F")

    # --- attribute setters mirroring micromag.py:148-211 -----------------------
    @property
    def m0(self):
        return self._m0

    @m0.setter
    def m0(self, val):
        self._m0 = np.zeros(shape=(self.ntot, 3), dtype=np.float64, order="F")
        if isinstance(val, type(None)):
            theta = np.pi * np.random.rand(self.ntot)
            phi = 2 * np.pi * np.random.rand(self.ntot)
            self._m0[:, 0] = np.sin(theta) * np.cos(phi)
            self._m0[:, 1] = np.sin(theta) * np.sin(phi)
            self._m0[:, 2] = np.cos(theta)
        elif isinstance(val, (int, float)):
            self._m0[:] = val
        else:
            assert np.asarray(val).shape == (self.ntot, 3)
            self._m0 = np.asarray(val, dtype=np.float64, order="F")

    @property
    def dem_appr(self):
        return self._dem_appr

    @dem_appr.setter
    def dem_appr(self, val=None):
        self._dem_appr = {
            None: 1, "threshold": 2, "fft_thres": 3, "threshold_fraction": 4, "fft_threshold_fraction": 5,
        }[val]

    @property
    def grid_type(self):
        return self._grid_type

    @grid_type.setter
    def grid_type(self, val=None):
        self._grid_type = {None: -1, "uniform": 1, "tetrahedron": 2, "unstructuredPrisms": 3}[val]

    @property
    def prob_mode(self):
        return self._prob_mode

    @prob_mode.setter
    def prob_mode(self, val=None):
        self._prob_mode = {None: -1, "new": 1, "old": 2}[val]

    @property
    def solver(self):
        return self._solver

    @solver.setter
    def solver(self, val=None):
        self._solver = {None: -1, "explicit": 1, "dynamic": 2, "implicit": 3}[val]

    # --- helpers -----------------------------------------------------------------
    def _precision_code(self):
        p = self.precision
        if isinstance(p, str):
            try:
                return {"fp32": _lib.PREC_FP32, "bf16x2": _lib.PREC_BF16X2, "fp32_mfma": _lib.PREC_FP32_MFMA}[p]
            except KeyError:
                raise ValueError("precision must be 'fp32', 'bf16x2' or 'fp32_mfma'") from None
        return _lib.PREC_FP32

    def set_exchange_matrix(self, A):
        """Exchange operator of a non-uniform grid (N x N, the second-derivative
        operator that multiplies -2 A0/(mu0 Ms)) into exch_* as a 1-based CSR."""
        import scipy.sparse as sp

        A = sp.csr_matrix(A)
        if A.shape != (self.ntot, self.ntot):
            raise ValueError(f"the exchange operator must be {self.ntot} x {self.ntot}")
        A.sort_indices()
        indptr, indices, data = A.indptr, A.indices, A.data
        self.exch_nrow = self.ntot
        self.exch_nval = int(len(data))
        self.exch_rows = np.asarray(indptr[:-1] + 1, dtype=np.int32, order="F")
        self.exch_rowe = np.asarray(indptr[1:] + 1, dtype=np.int32, order="F")
        self.exch_col = np.asarray(indices + 1, dtype=np.int32, order="F")
        self.exch_val = np.asarray(data, dtype=np.float64, order="F")

    def _check_supported(self):
        if self._grid_type not in (1, 2, 3):
            raise NotImplementedError("grid_type must be 'uniform', 'tetrahedron' or 'unstructuredPrisms'")
        if self._dem_appr in (3, 5):
            raise NotImplementedError("the FFT-threshold demag approximations are not implemented "
                                      "(demag_approx None, 'threshold' and 'threshold_fraction' are)")
        if self._solver not in (1, 2):
            raise NotImplementedError("solver must be 'explicit' or 'dynamic'")
        if self.CV > 0:
            raise NotImplementedError("demag-field noise (CV > 0) is not implemented")
        if self._grid_type == 1 and (int(self.grid_n[0]) < 2 or int(self.grid_n[1]) < 2):
            raise ValueError("the uniform-grid exchange operator needs nx >= 2 and ny >= 2 "
                             "(LandauLifshitzEquationSolver.f90:1215)")
        if self._grid_type == 2:
            if int(self.grid_nnod) < 4 or np.asarray(self.grid_nod).shape != (int(self.grid_nnod), 3):
                raise ValueError("tetrahedra need grid_nnod >= 4 nodes in grid_nod (grid_nnod x 3)")
            ele = np.asarray(self.grid_ele)
            if ele.shape != (4, self.ntot) or ele.min() < 1 or ele.max() > int(self.grid_nnod):
                raise ValueError("tetrahedra need grid_ele (4 x N, 1-based node indices within grid_nnod)")
        if self._grid_type in (2, 3):
            if len(self.exch_rows) != self.ntot or len(self.exch_rowe) != self.ntot:
                raise ValueError("unstructured prisms need the exchange operator: exch_rows / exch_rowe of length N "
                                 "(set_exchange_matrix fills them)")
            if self._grid_type == 3 and np.any(np.asarray(self.N_ave) < 1):
                raise ValueError("N_ave entries must be >= 1")

    def _tensor_file(self):
        """Path of the demag tensor file when filename is longer than two characters."""
        return self.N_file_in if self.N_load > 2 else None

    def _ms_array(self):
        return _c(np.broadcast_to(np.asarray(self.Ms, dtype=np.float64), (self.ntot,)))

    def _k0_array(self):
        return _c(np.broadcast_to(np.asarray(self.K0, dtype=np.float64), (self.ntot,)))

    def _time_tables(self, t_end, nt, fct_h_ext, nt_h_ext):
        t = np.linspace(0, t_end, nt)
        h_ext = np.zeros(shape=(nt_h_ext, 4), dtype=np.float64, order="F")
        h_ext[:, 0] = np.linspace(0, t_end, nt_h_ext)
        h_ext[:, 1:4] = fct_h_ext(np.linspace(0, t_end, nt_h_ext))
        return t, h_ext

    def _run(self, t, h_ext_rows, m0_rows, return_fields=True, max_steps=0):
        """Common path: h_ext_rows (B, nt_h_ext, 4), m0_rows (B, 3N) SoA."""
        self._check_supported()
        L = _lib.lib()
        B = m0_rows.shape[0]
        n = self.ntot
        nt = len(t)
        n_h = h_ext_rows.shape[1] if self._solver == 1 else 1
        keep = {}
        keep["Ms"] = self._ms_array()
        keep["K0"] = self._k0_array()
        keep["u"] = _c(self.u_ea)
        keep["alphat"] = _c(self.alphat)
        keep["hext"] = _c(h_ext_rows)
        keep["t"] = _c(t)
        keep["m0"] = _c(m0_rows)
        keep["tconv"] = _c(self.t_conv)
        q = _lib.MtProblem()
        q.nx, q.ny, q.nz = (int(v) for v in self.grid_n)
        q.Lx, q.Ly, q.Lz = (float(v) for v in self.grid_L)
        q.n_states = B
        q.Ms = _ptr(keep["Ms"]); q.K0 = _ptr(keep["K0"]); q.u_ea = _ptr(keep["u"])
        q.A0 = float(self.A0); q.gamma = float(self.gamma); q.alpha0 = float(self.alpha_mm)
        q.nt_alpha = self.nt_alpha; q.alphat = _ptr(keep["alphat"])
        q.solver = self._solver
        q.nt_hext = h_ext_rows.shape[1]; q.hext = _ptr(keep["hext"]); q.hext_per_state = 1
        q.nt = nt; q.t = _ptr(keep["t"])
        q.m0 = _ptr(keep["m0"]); q.m0_per_state = 1
        q.tol = float(self.tol); q.thres = float(self.thres)
        q.nt_conv = int(self.nt_conv); q.t_conv = _ptr(keep["tconv"]); q.conv_tol = float(self.conv_tol[0])
        q.precision = self._precision_code()
        q.return_fields = int(bool(return_fields))
        q.max_steps = int(max_steps)
        q.demag_approx = int(self._dem_appr)
        q.toeplitz = int(getattr(self, "_toeplitz", 0))
        q.demag_threshold = float(self.dem_thres)
        q.grid_type = self._grid_type
        if self._grid_type in (2, 3):
            keep["gp"] = _c(np.asarray(self.grid_pts).reshape(n, 3))
            keep["ga"] = _c(np.asarray(self.grid_abc).reshape(n, 3))
            if self._grid_type == 2:
                keep["gn"] = _c(np.asarray(self.grid_nod).reshape(-1, 3))
                keep["ge"] = _c(np.rint(np.asarray(self.grid_ele, dtype=np.float64)).T.reshape(n, 4), np.int32)
                q.grid_nnod = int(keep["gn"].shape[0])
                q.grid_nodes = _ptr(keep["gn"])
                q.grid_elements = keep["ge"].ctypes.data_as(C.POINTER(C.c_int))
            keep["xr"] = _c(self.exch_rows, np.int32)
            keep["xe"] = _c(self.exch_rowe, np.int32)
            keep["xc"] = _c(self.exch_col, np.int32)
            keep["xv"] = _c(self.exch_val)
            q.grid_pts = _ptr(keep["gp"]); q.grid_abc = _ptr(keep["ga"])
            q.n_ave[0], q.n_ave[1], q.n_ave[2] = (int(v) for v in np.asarray(self.N_ave).reshape(3))
            q.exch_nnz = int(keep["xc"].size)
            ip = C.POINTER(C.c_int)
            q.exch_rows_start = keep["xr"].ctypes.data_as(ip)
            q.exch_rows_end = keep["xe"].ctypes.data_as(ip)
            q.exch_cols = keep["xc"].ctypes.data_as(ip)
            q.exch_vals = _ptr(keep["xv"])
        path = self._tensor_file()
        Kout = None
        if path is not None and os.path.exists(path):
            keep["K"] = load_demag_tensor(path, n)
            q.K_in = keep["K"].ctypes.data_as(C.POINTER(C.c_float))
        elif path is not None:
            Kout = np.zeros((6, n, n), dtype=np.float32)
            q.K_out = Kout.ctypes.data_as(C.POINTER(C.c_float))
            q.keep_fp32_tensor = 1
        t_out = np.zeros((B, n_h, nt))
        y_out = np.zeros((B, n_h, nt, 3, n))
        fields = np.zeros((B, n_h, nt, 4, 3, n)) if return_fields else None
        stats = (_lib.MtStats * B)()
        h = C.c_void_p()
        _lib.check(L.mt_solver_create(C.byref(q), C.byref(h)))
        try:
            if getattr(self, "_timing", False):  # HIP-event kernel timing (benchmarks)
                L.mt_solver_set_timing(h, 1)
            if Kout is not None:
                # written straight after assembly, before integrating, as the reference
                # does (LandauLifshitzEquationSolver.f90:1149-1151); creation has
                # synchronised the device-to-host copy of the tensor
                save_demag_tensor(path, Kout)
                Kout = None
            _lib.check(L.mt_solver_run(h))
            _lib.check(L.mt_solver_results(h, t_out.ctypes.data, y_out.ctypes.data,
                                           fields.ctypes.data if fields is not None else None,
                                           C.cast(stats, C.c_void_p)))
            self.last_demag_info = None
            if self._dem_appr in (2, 4):
                nnz = (C.c_longlong * 6)()
                thr = C.c_float()
                _lib.check(L.mt_solver_demag_info(h, nnz, C.byref(thr)))
                form = C.c_int()
                _lib.check(L.mt_solver_demag_form(h, C.byref(form)))
                self.last_demag_info = dict(nnz=[int(v) for v in nnz], threshold=float(thr.value),
                                            form={1: "csr", 2: "bf16x2 tiles", 3: "fp32 tiles"}[form.value])
        finally:
            L.mt_solver_destroy(h)
        self.last_stats = [dict(f_calls=s.f_calls, steps=s.steps, bad_steps=s.bad_steps, flag3=s.flag3,
                                flag4=s.flag4, h_last=s.h_last, evals=s.evals, demag_passes=s.demag_passes,
                                demag_ms=s.demag_ms, llg_ms=s.llg_ms, vec_ms=s.vec_ms,
                                llg_launches=s.llg_launches, vec_launches=s.vec_launches) for s in stats]
        return t_out, y_out, fields

    def _pts(self):
        if self._grid_type in (2, 3):
            return np.asfortranarray(np.asarray(self.grid_pts, dtype=np.float64).reshape(self.ntot, 3))
        nx, ny, nz = (int(v) for v in self.grid_n)
        Lx, Ly, Lz = (float(v) for v in self.grid_L)

        def axis(m, L):
            if m > 1:
                d = L / m
                return np.array([-L / 2 + i * d + d / 2 for i in range(m)])
            return np.zeros(1)

        xs, ys, zs = axis(nx, Lx), axis(ny, Ly), axis(nz, Lz)
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        return np.asfortranarray(np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1))

    @staticmethod
    def _to_reference_layout(t_out, y_out, fields):
        """(n_h, nt, 3, N) -> (nt, N, n_h, 3) F-ordered, like the f2py outputs."""
        M = np.asfortranarray(np.transpose(y_out, (1, 3, 0, 2)))
        if fields is None:
            z = np.zeros_like(M, order="F")
            H = [z, z.copy(), z.copy(), z.copy()]
        else:
            H = [np.asfortranarray(np.transpose(fields[:, :, k], (1, 3, 0, 2))) for k in range(4)]
        return M, H

    # --- public API ------------------------------------------------------------------
    def run_simulation(self, t_end, nt, fct_h_ext, nt_h_ext):
        """Integrate the LL equation (reference: micromag.py:213-268).

        Returns (t_out[nt], M_out[nt,N,n_h,3], pts[N,3], H_exc, H_ext, H_dem, H_ani),
        n_h = 1 for the dynamic solver and nt_h_ext for the explicit solver.
        """
        t, h_ext = self._time_tables(t_end, nt, fct_h_ext, nt_h_ext)
        m0 = np.concatenate((self.m0[:, 0], self.m0[:, 1], self.m0[:, 2]), axis=None)
        t_out, y_out, fields = self._run(t, h_ext[None], m0[None])
        M, H = self._to_reference_layout(t_out[0], y_out[0], None if fields is None else fields[0])
        return (np.asfortranarray(t_out[0, 0 if self._solver == 2 else -1]), M, self._pts(), H[0], H[1], H[2], H[3])

    def run_simulation_batch(self, t_end, nt, h_ext_batch, nt_h_ext, m0_batch, return_fields=False, max_steps=0):
        """Integrate B independent states in one call (additive extension).

        h_ext_batch: callable or list of B callables (like fct_h_ext), or an
            array (B, nt_h_ext, 3) of applied fields in A/m on the linspace(0, t_end, nt_h_ext) grid.
        m0_batch: (B, N, 3) initial magnetizations.
        Returns t_out (B, nt), M_out (B, nt, N, 3) and, when requested, the four
        fields (B, nt, N, 3) each (dynamic solver).
        """
        m0_batch = np.asarray(m0_batch, dtype=np.float64)
        B = m0_batch.shape[0]
        assert m0_batch.shape[1:] == (self.ntot, 3)
        t = np.linspace(0, t_end, nt)
        tg = np.linspace(0, t_end, nt_h_ext)
        rows = np.zeros((B, nt_h_ext, 4))
        rows[:, :, 0] = tg
        if callable(h_ext_batch):
            rows[:, :, 1:4] = np.asarray(h_ext_batch(tg))[None]
        elif isinstance(h_ext_batch, (list, tuple)) and callable(h_ext_batch[0]):
            for b, f in enumerate(h_ext_batch):
                rows[b, :, 1:4] = f(tg)
        else:
            rows[:, :, 1:4] = np.asarray(h_ext_batch, dtype=np.float64).reshape(B, nt_h_ext, 3)
        m0 = np.transpose(m0_batch, (0, 2, 1)).reshape(B, 3 * self.ntot)
        t_out, y_out, fields = self._run(t, rows, m0, return_fields=return_fields, max_steps=max_steps)
        M = np.transpose(y_out[:, 0], (0, 1, 3, 2))
        out = (t_out[:, 0], M)
        if return_fields:
            out = out + tuple(np.transpose(fields[:, 0, :, k], (0, 1, 3, 2)) for k in range(4))
        return out


# --- demag tensor file (MagTenseMicroMagPyIO.f90:221-271) -----------------------------
def _transpose_into(a, out, blk=64):
    """out = a.T for a square plane, copied in 64 x 64 blocks that stay in cache
    (three times the rate of a whole-plane strided copy)."""
    n = a.shape[0]
    for i0 in range(0, n, blk):
        for j0 in range(0, n, blk):
            out[j0:j0 + blk, i0:i0 + blk] = a[i0:i0 + blk, j0:j0 + blk].T
    return out


def save_demag_tensor(path, K):
    """Six direct-access records Kxx,Kxy,Kxz,Kyy,Kyz,Kzz, each N*N FP32 column-major.

    K is (6, N, N) receiver-major (K[c][i][j] = receiver i, source j); the file
    stores column j = source as in the reference.
    """
    K = np.asarray(K, dtype=np.float32)
    buf = np.empty(K.shape[1:], dtype=np.float32)
    with open(path, "wb") as f:
        for c in range(6):
            _transpose_into(K[c], buf).tofile(f)


def load_demag_tensor(path, n):
    raw = np.fromfile(path, dtype=np.float32)
    if raw.size != 6 * n * n:
        raise ValueError(f"{path}: expected 6*{n}*{n} float32 values, found {raw.size}")
    planes = raw.reshape(6, n, n)  # each plane in file order = column-major K(:, j) -> [j][i]
    out = np.empty_like(planes)
    for c in range(6):
        _transpose_into(planes[c], out[c])
    return out
