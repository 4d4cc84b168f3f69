"""Differentiable Landau-Lifshitz integration on the MI355X (torch.autograd).

``ll_rk4(m0, h_ext, ms, ...)`` integrates the MagTense right-hand side
(LandauLifshitzEquationSolver.f90:220-273, same field terms and conventions as
the solver) with ``nsteps`` fixed RK4 steps of ``dt`` and returns m(T).  Its
backward pass is the hand-written discrete adjoint of the integration in
``libmagtense_hip.so`` (``mt_rk4_backward``): checkpoints every
``checkpoint_every`` steps, each segment recomputed once, four
vector-Jacobian products per step.  Gradients flow to m0, to the applied field
(A/m) and to the per-cell saturation magnetization.  There is no reference
counterpart (MagTense has no gradients); this extends the drop-in.

Tensors live on the host (float64); the integration itself runs on the GPU.
``demag_precision='fp32'`` (default) forms the demag field as the reference
does (FP32 tensor and product, FP32 Ms scaling); ``'fp64'`` keeps the tensor,
product and Ms scaling in FP64, which makes the integration a smooth function
of its inputs down to FP64 rounding (what finite-difference checks need).
"""

import ctypes as C

import numpy as np
import torch

from . import _lib


class _Rk4Handle:
    def __init__(self, res, L, ms, h_ext, dt, nsteps, checkpoint_every, A0, alpha, gamma, K0, u_ea, n_states,
                 demag_precision="fp32"):
        L_ = _lib.lib()
        n = int(np.prod(res))
        self.n, self.B = n, n_states
        self.keep = {
            "Ms": np.ascontiguousarray(np.broadcast_to(np.asarray(ms, dtype=np.float64), (n,))),
            "K0": np.ascontiguousarray(np.broadcast_to(np.asarray(K0, dtype=np.float64), (n,))),
            "u": np.ascontiguousarray(np.zeros((n, 3)) if u_ea is None else np.asarray(u_ea, dtype=np.float64)),
        }
        hext = np.zeros((n_states, 1, 4))
        hext[:, 0, 1:] = np.asarray(h_ext, dtype=np.float64).reshape(n_states, 3)
        self.keep["hext"] = hext
        q = _lib.MtProblem()
        q.nx, q.ny, q.nz = (int(v) for v in res)
        q.Lx, q.Ly, q.Lz = (float(v) for v in L)
        q.n_states = n_states
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        q.Ms = dp(self.keep["Ms"])
        q.K0 = dp(self.keep["K0"])
        q.u_ea = dp(self.keep["u"])
        q.A0, q.gamma, q.alpha0 = float(A0), float(gamma), float(alpha)
        q.nt_hext = 1
        q.hext = dp(self.keep["hext"])
        q.hext_per_state = 1
        try:
            q.precision = {"fp32": _lib.PREC_FP32, "fp64": _lib.PREC_FP64}[demag_precision]
        except KeyError:
            raise ValueError("demag_precision must be 'fp32' or 'fp64'") from None
        h = C.c_void_p()
        _lib.check(L_.mt_rk4_create(C.byref(q), float(dt), int(nsteps), int(checkpoint_every), C.byref(h)))
        self.h = h
        self.L = L_

    def forward(self, m0):
        m0 = np.ascontiguousarray(m0, dtype=np.float64)
        mT = np.zeros_like(m0)
        _lib.check(self.L.mt_rk4_forward(self.h, m0.ctypes.data, mT.ctypes.data))
        return mT

    def backward(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        gm0 = np.zeros_like(g)
        gh = np.zeros((self.B, 3))
        gms = np.zeros((self.B, self.n))
        _lib.check(self.L.mt_rk4_backward(self.h, g.ctypes.data, gm0.ctypes.data, gh.ctypes.data, gms.ctypes.data))
        return gm0, gh, gms

    def __del__(self):
        if getattr(self, "h", None):
            self.L.mt_rk4_destroy(self.h)
            self.h = None


def _soa(m):  # (B, N, 3) -> (B, 3N)
    return np.transpose(m, (0, 2, 1)).reshape(m.shape[0], -1)


def _aos(y, n):  # (B, 3N) -> (B, N, 3)
    return np.transpose(y.reshape(y.shape[0], 3, n), (0, 2, 1))


class LLRK4(torch.autograd.Function):
    """m(T) = RK4^nsteps(m0; H_ext, Ms); gradients by the discrete adjoint on the GPU."""

    @staticmethod
    def forward(ctx, m0, h_ext, ms, cfg):
        B = m0.shape[0]
        handle = _Rk4Handle(cfg["res"], cfg["L"], ms.detach().cpu().numpy(), h_ext.detach().cpu().numpy(),
                            cfg["dt"], cfg["nsteps"], cfg["checkpoint_every"], cfg["A0"], cfg["alpha"],
                            cfg["gamma"], cfg["K0"], cfg["u_ea"], B, cfg["demag_precision"])
        mT = handle.forward(_soa(m0.detach().cpu().numpy()))
        ctx.handle = handle
        ctx.ms_shape = ms.shape
        return torch.from_numpy(_aos(mT, handle.n).copy())

    @staticmethod
    def backward(ctx, grad_mT):
        h = ctx.handle
        gm0, gh, gms = h.backward(_soa(grad_mT.detach().cpu().numpy().astype(np.float64)))
        gms_total = gms.sum(axis=0)  # Ms is shared by the states
        if len(ctx.ms_shape) == 0:
            gms_total = gms_total.sum()
        return (torch.from_numpy(_aos(gm0, h.n).copy()), torch.from_numpy(gh), torch.as_tensor(gms_total), None)


def ll_rk4(m0, h_ext, ms, res, L, dt, nsteps, checkpoint_every=100, A0=1.3e-11, alpha=0.02 * 2.21e5,
           gamma=2.21e5, K0=0.0, u_ea=None, demag_precision="fp32"):
    """Differentiable RK4 LL integration.

    m0: (B, N, 3) or (N, 3) float64 tensor; h_ext: (B, 3) or (3,) applied field
    in A/m (constant in time); ms: (N,) or scalar Ms in A/m.  ``alpha`` is the
    API's alpha (alpha_LL * gamma), as in MicromagProblem.  demag_precision:
    'fp32' (the reference's demag arithmetic) or 'fp64'.
    Returns m(T) with the shape of m0.
    """
    single = m0.dim() == 2
    if single:
        m0 = m0.unsqueeze(0)
        h_ext = h_ext.reshape(1, 3)
    cfg = dict(res=list(res), L=list(L), dt=float(dt), nsteps=int(nsteps), checkpoint_every=int(checkpoint_every),
               A0=A0, alpha=alpha, gamma=gamma, K0=K0, u_ea=u_ea, demag_precision=demag_precision)
    out = LLRK4.apply(m0, h_ext.reshape(m0.shape[0], 3), ms, cfg)
    return out[0] if single else out
