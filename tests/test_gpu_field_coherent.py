"""Demag-field error of the GPU products on coherent magnetizations.

Random unit vectors make every receiver's field a sum of terms with random
signs, dominated by the self term.  A uniform or smooth magnetization makes
the far field of a film add up coherently (and for in-plane m on a thin film
the total field is a small difference of large terms), so the field error is
measured here on those states too: uniform +x, +z and (1, 1, 1)/sqrt(3), an
analytic smooth in-plane pattern, a relaxed state (gamma = 0 relaxation of a
random start under a decaying field, as std_problem_4.py:12-21 relaxes its
s-state), and random unit vectors, on the 128 x 32 x 1 film (N = 4096,
3.9 x 3.9 x 3 nm cells).

Reference: the FP64 product of the FP64 tensor (oracle).  Reported: the
norm-wise relative error ||H - H_ref|| / ||H_ref|| over all receivers, the
largest per-receiver error relative to the largest field max_i ||H_i -
H_ref_i|| / max_i ||H_ref_i||, and the largest per-receiver relative error
||H_i - H_ref_i|| / ||H_ref_i|| (ill-conditioned where a receiver's field
nearly cancels: the FP32 product itself reaches 2e-5 there).

The three split products alone leave 1.6e-5 norm-wise on the smooth in-plane
state (the rounding residuals of m and K add up coherently where the film's
in-plane field is a small difference of large near-field terms); with the
near-band correction (mt_demag_near_correction, radius 2: the FP32 product's
difference from the split products over the 5 x 5 neighbourhood) every state
is held to 1e-5 norm-wise and 1e-5 of the largest field per receiver.  Bars:
corrected split-BF16 1e-5 / 1e-5 / 2e-4, FP32 1e-6 / 1e-6 / 1e-4, the
uncorrected split products 1e-5 norm-wise on random states only.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COMP = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
RES, L = [128, 32, 1], [500e-9, 125e-9, 3e-9]
MU0 = 4e-7 * np.pi


def _fp64_product(K, m):
    S, _, n = m.shape
    out = np.zeros((S, 3, n))
    for a in range(3):
        for b in range(3):
            out[:, a] -= m[:, b] @ K[COMP[a][b]].T
    return out


def _states(pts, rng):
    n = pts.shape[0]
    x, y = pts[:, 0] / L[0], pts[:, 1] / L[1]
    out = {}
    for name, v in (("uniform +x", [1, 0, 0]), ("uniform +z", [0, 0, 1]), ("uniform (1,1,1)", [1, 1, 1])):
        u = np.array(v, dtype=np.float64) / np.linalg.norm(v)
        out[name] = np.repeat(u[:, None], n, axis=1)
    phi = 0.6 * np.sin(np.pi * x) * np.cos(np.pi * y) + 0.3 * np.tanh(6 * y)
    sm = np.stack([np.cos(phi), np.sin(phi), 0.05 * np.cos(2 * np.pi * x)])
    out["smooth in-plane"] = sm / np.linalg.norm(sm, axis=0)
    r = rng.normal(size=(3, n))
    out["random"] = r / np.linalg.norm(r, axis=0)
    return out


def _relaxed():
    from magtense.micromag import MicromagProblem

    n = int(np.prod(RES))
    rng = np.random.default_rng(21)
    m0 = rng.normal(size=(n, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    h = np.array([1.0, 1.0, 1.0]) / MU0 * 0.05

    def fh(t):
        return np.expand_dims(np.where(t < 1e-9, 1.0 - t / 1e-9, 0), axis=1) * h

    p = MicromagProblem(RES, L, m0=m0, alpha=4.42e3, gamma=0.0, tol=1e-5)
    _, M, *_ = p.run_simulation(3e-9, 4, fh, 50)
    return np.squeeze(M, axis=2)[-1].T.copy()


def test_field_error_on_coherent_states(oracle):
    import torch
    from magtense_amd import _lib

    Lb = _lib.lib()
    pts, d = oracle.setup_grid(RES, L)
    n = pts.shape[0]
    K = oracle.tensor_uniform(RES, L)
    states = _states(pts, np.random.default_rng(5))
    states["relaxed"] = _relaxed()
    names = list(states)
    m = np.stack([states[k] for k in names])  # (S, 3, n)
    S = m.shape[0]
    m32 = m.astype(np.float32)
    ref = _fp64_product(K, m32.astype(np.float64))

    dp = torch.as_tensor(pts).cuda()
    da = torch.as_tensor(d).cuda()
    K32 = torch.empty(6 * n * n, dtype=torch.float32, device="cuda")
    split = torch.empty(Lb.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda")
    _lib.check(Lb.mt_prism_tensor(dp.data_ptr(), da.data_ptr(), 0, None, dp.data_ptr(), n, n, K32.data_ptr(), None,
                                  split.data_ptr()))
    tab = torch.empty(Lb.mt_toeplitz_tensor_bytes(*RES), dtype=torch.uint8, device="cuda")
    _lib.check(Lb.mt_toeplitz_tensor(*RES, *L, tab.data_ptr()))
    dm = torch.as_tensor(m32).cuda()
    out = {}
    for kind in ("fp32", "bf16x2 dense", "bf16x2 table", "bf16x2 dense+near", "bf16x2 table+near"):
        h = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        if kind == "fp32":
            _lib.check(Lb.mt_demag_apply_f32(K32.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
        elif kind.startswith("bf16x2 dense"):
            _lib.check(Lb.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
        else:
            _lib.check(Lb.mt_demag_apply_bf16x2_toeplitz(tab.data_ptr(), *RES, dm.data_ptr(), h.data_ptr(), S, -1.0))
        if kind.endswith("+near"):
            _lib.check(Lb.mt_demag_near_correction(*RES, *L, 2, dm.data_ptr(), h.data_ptr(), S, -1.0))
        out[kind] = h.cpu().numpy().reshape(S, 3, n).astype(np.float64)

    bars = {"fp32": (1e-6, 1e-6, 1e-4), "bf16x2 dense+near": (1e-5, 1e-5, 2e-4),
            "bf16x2 table+near": (1e-5, 1e-5, 2e-4)}
    failures = []
    for kind, g in out.items():
        for s, name in enumerate(names):
            e = g[s] - ref[s]
            hn = np.linalg.norm(ref[s], axis=0)
            norm_rel = np.linalg.norm(e) / np.linalg.norm(ref[s])
            to_max = (np.linalg.norm(e, axis=0) / hn.max()).max()
            row = np.linalg.norm(e, axis=0) / hn
            print(f"{kind:18s} {name:16s} |H_ref| {hn.mean():.3e}: norm-wise {norm_rel:.2e}, per receiver / max "
                  f"{to_max:.2e}, per receiver max {row.max():.2e} (99.9% {np.quantile(row, 0.999):.2e})")
            bar = bars.get(kind, (1e-5, 1.0, 1.0) if name == "random" else (1.0, 1.0, 1.0))
            if norm_rel >= bar[0] or to_max >= bar[1] or row.max() >= bar[2]:
                failures.append((kind, name, norm_rel, to_max, row.max()))
    assert not failures, failures
