"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  It checks the product; it is never the
product.  The C++ source (``micromag_oracle.cpp``) restates the reference
Fortran listed in its header; this file only marshals numpy arrays.
"""

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class OrProblem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("u_ea", _dp), ("Ms", _dp), ("K0", _dp),
        ("A0", C.c_double), ("gamma", C.c_double), ("alpha0", C.c_double),
        ("solver", C.c_int),
        ("nt_hext", C.c_int), ("hext", _dp),
        ("nt_alpha", C.c_int), ("alphat", _dp),
        ("nt", C.c_int), ("t", _dp),
        ("m0", _dp),
        ("tol", C.c_double), ("thres", C.c_double),
        ("nt_conv", C.c_int), ("t_conv", _dp), ("conv_tol", C.c_double),
        ("demag_fp64", C.c_int),
        ("K_in", _fp),
        ("return_fields", C.c_int),
        ("demag_approx", C.c_int),
        ("demag_threshold", C.c_double),
        ("grid_type", C.c_int),
        ("grid_pts", _dp), ("grid_abc", _dp),
        ("n_ave", C.c_int * 3),
        ("exch_nnz", C.c_int),
        ("exch_rows_start", _ip), ("exch_rows_end", _ip), ("exch_cols", _ip), ("exch_vals", _dp),
        ("grid_nnod", C.c_int), ("grid_nodes", _dp), ("grid_elements", _ip),
    ]


class OrStats(C.Structure):
    _fields_ = [("f_calls", C.c_longlong), ("steps", C.c_longlong), ("bad_steps", C.c_longlong),
                ("flag3", C.c_longlong), ("flag4", C.c_longlong), ("h_last", C.c_double)]


RHS_CB = C.CFUNCTYPE(None, C.c_double, _dp, _dp, C.c_void_p)


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.or_prism_k6.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        _lib.or_setup_grid.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [_dp, _dp]
        _lib.or_tensor_uniform.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [_dp]
        _lib.or_tensor_points.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp]
        _lib.or_exchange_csr.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [_ip, _ip, _dp]
        _lib.or_dmdt.argtypes = [C.POINTER(OrProblem), C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib.or_run_micromag.argtypes = [C.POINTER(OrProblem), _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(OrStats)]
        _lib.or_rk_integrate.argtypes = [C.c_int, RHS_CB, C.c_void_p, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                         _dp, _dp, C.POINTER(OrStats)]
        _lib.or_rk_tables.argtypes = [_dp] * 5
        _lib.or_demag_rows.argtypes = [C.c_int, C.c_int, _fp, _fp, _dp, _dp]
        _lib.or_rk4_adjoint.argtypes = [C.POINTER(OrProblem), C.c_double, C.c_longlong, _dp, _dp, _dp, _dp, _dp,
                                        _dp, _dp]
        _lib.or_rotation_matrices.argtypes = [_dp, _dp, _dp]
        _lib.or_threshold_fraction.argtypes = [_fp, C.c_longlong, C.c_double]
        _lib.or_threshold_fraction.restype = C.c_float
        _lib.or_interp1.argtypes = [C.c_int, _dp, _dp, C.c_double, _ip]
        _lib.or_interp1.restype = C.c_double
        _lib.or_tiles_tensor.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _ip, _dp, C.c_int, _dp, _dp]
        _lib.or_tiles_field.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _ip, _ip, _dp, C.c_int, _dp, _dp, _dp]
        _lib.or_tiles_tensor_v.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _ip, _dp, _dp, C.c_int, _dp, _dp]
        _lib.or_tiles_field_v.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp, C.c_int, _dp, _dp, _dp]
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def prism_k6(abc, pos):
    """K (= -N_standard) of one prism with full edges abc at offset pos: xx,xy,xz,yy,yz,zz."""
    out = np.zeros(6)
    p = _c64(pos)
    lib().or_prism_k6(float(abc[0]), float(abc[1]), float(abc[2]), _d(p), _d(out))
    return out


def setup_grid(res, L):
    n = int(np.prod(res))
    pts = np.zeros((n, 3))
    d = np.zeros(3)
    lib().or_setup_grid(int(res[0]), int(res[1]), int(res[2]), float(L[0]), float(L[1]), float(L[2]), _d(pts), _d(d))
    return pts, d


def tensor_uniform(res, L):
    """FP64 tensor, shape (6, N, N), receiver-major."""
    n = int(np.prod(res))
    K = np.zeros((6, n, n))
    lib().or_tensor_uniform(int(res[0]), int(res[1]), int(res[2]), float(L[0]), float(L[1]), float(L[2]), _d(K))
    return K


def tensor_points(src, abc, pts, j0=0, j1=None):
    src = _c64(src); abc = _c64(abc); pts = _c64(pts)
    ns, npt = src.shape[0], pts.shape[0]
    j1 = ns if j1 is None else j1
    out = np.zeros((6, npt, ns))
    lib().or_tensor_points(ns, _d(src), _d(abc), npt, _d(pts), int(j0), int(j1), _d(out))
    return out


def tensor_unstructured(pts, abc, n_ave=(1, 1, 1)):
    """Receiver-averaged FP64 tensor of unstructured prisms, shape (6, N, N)."""
    pts = _c64(pts); abc = _c64(abc)
    n = pts.shape[0]
    nave = np.ascontiguousarray(n_ave, dtype=np.int32)
    K = np.zeros((6, n, n))
    rc = lib().or_tensor_unstructured(n, _d(pts), _d(abc), nave.ctypes.data_as(_ip), _d(K))
    if rc:
        raise ValueError("or_tensor_unstructured: invalid arguments")
    return K


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops; returns the previous count."""
    return int(lib().or_set_threads(int(n)))


def tetra_n(vert, r):
    """Tetrahedron tensor (MagTense sign) at r; vert (4, 3) vertices; (3, 3)."""
    v = _c64(np.asarray(vert, dtype=np.float64).reshape(4, 3))
    rr = _c64(np.asarray(r, dtype=np.float64).reshape(3))
    N = np.zeros((3, 3))
    if lib().or_tetra_n(_d(v), _d(rr), _d(N)):
        raise ValueError("degenerate tetrahedron face")
    return N


def exchange_csr(res, L):
    nnz = lib().or_exchange_csr(int(res[0]), int(res[1]), int(res[2]), float(L[0]), float(L[1]), float(L[2]),
                                None, None, None)
    if nnz < 0:
        raise ValueError("exchange operator needs nx >= 2 and ny >= 2")
    n = int(np.prod(res))
    rowptr = np.zeros(n + 1, dtype=np.int32)
    col = np.zeros(nnz, dtype=np.int32)
    val = np.zeros(nnz)
    lib().or_exchange_csr(int(res[0]), int(res[1]), int(res[2]), float(L[0]), float(L[1]), float(L[2]),
                          rowptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip), _d(val))
    return rowptr, col, val


class Problem:
    """Holds numpy buffers alive and fills an OrProblem."""

    def __init__(self, res, L, Ms=8e5, A0=1.3e-11, K0=0.0, gamma=0.0, alpha=0.02, u_ea=None, solver=2,
                 hext=None, alphat=None, t=None, m0=None, tol=1e-4, thres=1e-6, t_conv=None, conv_tol=1e-4,
                 demag_fp64=False, K_in=None, return_fields=True, demag_approx=1, dem_thres=0.0,
                 grid_pts=None, grid_abc=None, n_ave=(1, 1, 1), exch=None, grid_nodes=None, grid_elements=None):
        """grid_pts / grid_abc ((n, 3) centres and full edges) select the
        unstructured-prism grid; exch = (rows_start, rows_end, cols, vals) is
        its exchange operator as a 1-based CSR (rows_start / rows_end of
        mkl_sparse_d_create_csr).  grid_nodes ((nnod, 3)) with grid_elements
        ((n, 4), 1-based node indices) select the tetrahedral grid instead of
        grid_abc; grid_pts are then the evaluation points."""
        n = int(np.prod(res))
        self.n = n
        self.res = list(res)
        self.L = list(L)
        self._Ms = _c64(np.broadcast_to(np.asarray(Ms, dtype=np.float64), (n,)))
        self._K0 = _c64(np.broadcast_to(np.asarray(K0, dtype=np.float64), (n,)))
        self._u = _c64(np.zeros((n, 3)) if u_ea is None else np.asarray(u_ea).reshape(n, 3))
        self._hext = _c64(np.zeros((1, 4)) if hext is None else hext)
        self._alphat = _c64(np.zeros((1, 2)) if alphat is None else alphat)
        self._t = _c64(np.linspace(0, 1e-9, 2) if t is None else t)
        self._m0 = _c64(np.zeros(3 * n) if m0 is None else np.asarray(m0).reshape(-1))
        self._tconv = _c64(np.zeros(1) if t_conv is None else t_conv)
        self._K = None if K_in is None else np.ascontiguousarray(K_in, dtype=np.float32)
        q = OrProblem()
        q.nx, q.ny, q.nz = (int(r) for r in res)
        q.Lx, q.Ly, q.Lz = (float(v) for v in L)
        q.u_ea = _d(self._u); q.Ms = _d(self._Ms); q.K0 = _d(self._K0)
        q.A0 = A0; q.gamma = gamma; q.alpha0 = alpha
        q.solver = solver
        q.nt_hext = self._hext.shape[0]; q.hext = _d(self._hext)
        q.nt_alpha = self._alphat.shape[0]; q.alphat = _d(self._alphat)
        q.nt = self._t.shape[0]; q.t = _d(self._t)
        q.m0 = _d(self._m0)
        q.tol = tol; q.thres = thres
        q.nt_conv = self._tconv.shape[0]; q.t_conv = _d(self._tconv); q.conv_tol = conv_tol
        q.demag_fp64 = int(bool(demag_fp64))
        q.K_in = self._K.ctypes.data_as(_fp) if self._K is not None else None
        q.return_fields = int(bool(return_fields))
        q.demag_approx = int(demag_approx)
        q.demag_threshold = float(dem_thres)
        q.grid_type = 1
        if grid_pts is not None:
            self._gp = _c64(np.asarray(grid_pts, dtype=np.float64).reshape(n, 3))
            self._ga = _c64(np.zeros((n, 3)) if grid_abc is None else np.asarray(grid_abc, dtype=np.float64).reshape(n, 3))
            rs, re_, cols, vals = exch
            self._ex = [np.ascontiguousarray(rs, dtype=np.int32), np.ascontiguousarray(re_, dtype=np.int32),
                        np.ascontiguousarray(cols, dtype=np.int32), _c64(vals)]
            q.grid_type = 3
            q.grid_pts = _d(self._gp); q.grid_abc = _d(self._ga)
            q.n_ave[0], q.n_ave[1], q.n_ave[2] = (int(v) for v in n_ave)
            q.exch_nnz = self._ex[2].size
            q.exch_rows_start = self._ex[0].ctypes.data_as(_ip)
            q.exch_rows_end = self._ex[1].ctypes.data_as(_ip)
            q.exch_cols = self._ex[2].ctypes.data_as(_ip)
            q.exch_vals = _d(self._ex[3])
            if grid_nodes is not None:
                self._gn = _c64(np.asarray(grid_nodes, dtype=np.float64).reshape(-1, 3))
                self._ge = np.ascontiguousarray(np.asarray(grid_elements).reshape(n, 4), dtype=np.int32)
                q.grid_type = 2
                q.grid_nnod = self._gn.shape[0]
                q.grid_nodes = _d(self._gn)
                q.grid_elements = self._ge.ctypes.data_as(_ip)
        self.q = q

    def dmdt(self, t, m):
        m = _c64(m).reshape(-1)
        out = np.zeros(3 * self.n)
        H = [np.zeros(3 * self.n) for _ in range(4)]
        rc = lib().or_dmdt(C.byref(self.q), float(t), _d(m), _d(out), *(_d(h) for h in H))
        if rc:
            raise RuntimeError(f"or_dmdt failed: {rc}")
        return out, H  # H = [H_exc, H_ext, H_dem, H_ani] (negated fields, SoA)

    def rk4_adjoint(self, dt, nsteps, hext3, m0, dL_dmT=None):
        """RK4 forward (and discrete adjoint when dL_dmT is given); SoA vectors."""
        hext3 = _c64(hext3)
        m0 = _c64(m0).reshape(-1)
        mT = np.zeros_like(m0)
        out = [None, None, None]
        if dL_dmT is not None:
            g = _c64(dL_dmT).reshape(-1)
            out = [np.zeros_like(m0), np.zeros(3), np.zeros(self.n)]
            rc = lib().or_rk4_adjoint(C.byref(self.q), dt, nsteps, _d(hext3), _d(m0), _d(mT), _d(g),
                                      *(_d(o) for o in out))
        else:
            rc = lib().or_rk4_adjoint(C.byref(self.q), dt, nsteps, _d(hext3), _d(m0), _d(mT), None, None, None, None)
        if rc:
            raise RuntimeError(f"or_rk4_adjoint failed: {rc}")
        return mT, out[0], out[1], out[2]

    def run(self):
        nt = self.q.nt
        nh = self.q.nt_hext if self.q.solver == 1 else 1
        t_out = np.zeros(nh * nt)
        y = np.zeros(nh * nt * 3 * self.n)
        H = [np.zeros(nh * nt * 3 * self.n) for _ in range(4)]
        st = OrStats()
        rc = lib().or_run_micromag(C.byref(self.q), _d(t_out), _d(y), *(_d(h) for h in H), C.byref(st))
        if rc:
            raise RuntimeError(f"or_run_micromag failed: {rc}")
        shape = (nh, nt, 3, self.n)
        return (t_out.reshape(nh, nt), y.reshape(shape), [h.reshape(shape) for h in H],
                dict(f_calls=st.f_calls, steps=st.steps, bad_steps=st.bad_steps, flag3=st.flag3, flag4=st.flag4,
                     h_last=st.h_last))


def rk_integrate(f, t, y0, tol=1e-6, thres=1e-6):
    """RKSUITE 'M' on y' = f(t, y) (numpy callback), outputs at t (t[0] = start)."""
    y0 = _c64(y0).reshape(-1)
    neq = y0.size
    t = _c64(t)

    def cb(tt, yp, dp, _user):
        y = np.ctypeslib.as_array(yp, shape=(neq,))
        d = np.ctypeslib.as_array(dp, shape=(neq,))
        d[:] = f(tt, y.copy())

    cbf = RHS_CB(cb)
    t_out = np.zeros(t.size)
    y_out = np.zeros(t.size * neq)
    st = OrStats()
    rc = lib().or_rk_integrate(neq, cbf, None, t.size, _d(t), _d(y0), tol, thres, _d(t_out), _d(y_out), C.byref(st))
    if rc:
        raise RuntimeError(f"or_rk_integrate failed: {rc}")
    return t_out, y_out.reshape(t.size, neq), dict(f_calls=st.f_calls, steps=st.steps, bad_steps=st.bad_steps,
                                                   flag3=st.flag3, flag4=st.flag4, h_last=st.h_last)


def rk_tables():
    a = np.zeros((11, 11)); b = np.zeros(8); c = np.zeros(11); e = np.zeros(7); r = np.zeros((11, 6))
    lib().or_rk_tables(_d(a), _d(b), _d(c), _d(e), _d(r))
    return a, b, c, e, r


def demag_rows(Krows, ms_rows, m):
    """FP32-semantics demag field for R receiver rows: Krows (6, R, n) float32, m (3n,) -> (3, R)."""
    Krows = np.ascontiguousarray(Krows, dtype=np.float32)
    ms_rows = np.ascontiguousarray(ms_rows, dtype=np.float32)
    m = _c64(m).reshape(-1)
    _, R, n = Krows.shape
    out = np.zeros(3 * R)
    lib().or_demag_rows(n, R, Krows.ctypes.data_as(_fp), ms_rows.ctypes.data_as(_fp), _d(m), _d(out))
    return out.reshape(3, R)


def rotation_matrices(angles):
    """(rotMat, rotMatInv) of DemagFieldGetSolution.f90:925-948, row-major."""
    R = np.zeros((3, 3))
    Ri = np.zeros((3, 3))
    lib().or_rotation_matrices(_d(_c64(np.asarray(angles, dtype=np.float64))), _d(R), _d(Ri))
    return R, Ri


def _tile_args(size, offset, rot, M, tile_type, use_sym, sym_ops):
    nt = np.asarray(size).reshape(-1, 3).shape[0]

    def f3(a, default):
        a = np.full((nt, 3), default, dtype=np.float64) if a is None else np.asarray(a, dtype=np.float64)
        return np.asfortranarray(a.reshape(nt, 3))

    def i1(a, default):
        a = np.full(nt, default, dtype=np.int32) if a is None else np.asarray(a, dtype=np.int32)
        return np.ascontiguousarray(a.reshape(nt))

    return nt, f3(size, 0.0), f3(offset, 0.0), f3(rot, 0.0), f3(M, 0.0), i1(tile_type, 2), i1(use_sym, 0), \
        f3(sym_ops, 1.0)


def _vertices(nt, vertices):
    if vertices is None:
        return None
    return np.asfortranarray(np.asarray(vertices, dtype=np.float64).reshape(nt, 3, 4))


def tiles_tensor(size, offset, pts, rot=None, tile_type=None, use_sym=None, sym_ops=None, vertices=None):
    """getNFromTiles for prism and tetrahedron tiles: N(n_tiles, n_pts, 3, 3),
    Fortran order; vertices (n_tiles, 3, 4) for tile_type 5."""
    nt, sz, off, rt, _, tt, us, so = _tile_args(size, offset, rot, None, tile_type, use_sym, sym_ops)
    vv = _vertices(nt, vertices)
    P = np.asfortranarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    npt = P.shape[0]
    N = np.zeros((nt, npt, 3, 3), order="F")
    rc = lib().or_tiles_tensor_v(nt, _d(sz), _d(off), _d(rt), tt.ctypes.data_as(_ip), us.ctypes.data_as(_ip),
                                 _d(so), None if vv is None else _d(vv), npt, _d(P), _d(N))
    if rc:
        raise ValueError("only prism (2) and tetrahedron (5, with vertices) tiles are restated")
    return N


def tiles_field(size, offset, M, pts, rot=None, tile_type=None, use_sym=None, sym_ops=None, N=None, vertices=None):
    """getHFromTiles for prism and tetrahedron tiles: H(n_pts, 3), Fortran order."""
    nt, sz, off, rt, mm, tt, us, so = _tile_args(size, offset, rot, M, tile_type, use_sym, sym_ops)
    vv = _vertices(nt, vertices)
    P = np.asfortranarray(np.asarray(pts, dtype=np.float64).reshape(-1, 3))
    npt = P.shape[0]
    H = np.zeros((npt, 3), order="F")
    Nin = None if N is None else np.asfortranarray(np.asarray(N, dtype=np.float64))
    rc = lib().or_tiles_field_v(nt, _d(sz), _d(off), _d(rt), _d(mm), tt.ctypes.data_as(_ip), us.ctypes.data_as(_ip),
                                _d(so), None if vv is None else _d(vv), npt, _d(P), None if Nin is None else _d(Nin),
                                _d(H))
    if rc:
        raise ValueError("only prism (2) and tetrahedron (5, with vertices) tiles are restated")
    return H


def threshold_fraction(K, fraction):
    """FindThresholdFraction over the positive entries of an FP32 tensor (any shape)."""
    K = np.ascontiguousarray(K, dtype=np.float32)
    return float(lib().or_threshold_fraction(K.ctypes.data_as(_fp), K.size, float(fraction)))


def interp1(x, y, xv):
    """interp1_MagTense of one value: (value, locate's 1-based bracket)."""
    x = _c64(x); y = _c64(y)
    j = C.c_int()
    v = lib().or_interp1(x.size, _d(x), _d(y), float(xv), C.byref(j))
    return v, j.value
