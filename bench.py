"""Benchmark: LL steps/s at N = 32768 cells with a 64-state batch on one MI355X.

Workload (the precision + batching configuration): a 256 x 64 x 2 film
(2 x 2 x 1.5 nm cells, N = 32768), Ms = 8e5 A/m, A = 1.3e-11 J/m, no
anisotropy, alpha = 0.02 (API alpha = 0.02 * gamma), gamma = 2.21e5.  64
independent states, each seeded random unit vectors, at 64 applied fields
evenly spaced from -100 to +100 mT along x (a field sweep).  RKSUITE 'M',
tol 1e-6.  One benchmark step = one accepted RK45 step of all 64 states (7
right-hand sides, each one pass over the demag tensor).

The timed path is the split-BF16 MFMA demag product over the dense 25.8 GB
fragment-tiled tensor (HBM-bound) with its radius-2 near-band FP32
correction.  Beside it, on the same states: the same product over the uniform
grid's translation-invariant table (140 MB, L2/MALL resident), the plain
per-state FP32 mat-vec path (the drop-in), the plain-method configurations
end to end (32 x 8 x 1 and 128 x 32 x 1 films, with the CPU oracle), the
product at one row shard's shape (N/8 receivers x N sources, the per-GPU work
of the row-sharded design), the gradient wall time of the differentiation
configuration (RK4 adjoint, N = 4096, 2000 steps), the FP64 field error on
random and coherent states, the 16384-prism tensor assembly, the widened rows
(threshold-sparse products, magnetostatic field, tetrahedron tiles, tensor
file), and the FP64 CPU oracle timed on a bounded sample on the host cores.
Prints one JSON line.
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MU0 = 4e-7 * np.pi
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA, MI355X
FP32_PEAK_TFLOPS = 157.3  # FP32 matrix, MI355X


def traffic_from_profile(kind):
    """HBM bytes per GEMM launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "gemm_bf16x2_n32768_b64_pmc.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_{kind}:"
    with open(path) as fh:
        for ln in fh:
            if ln.strip().startswith(key):
                return float(ln.split(":", 1)[1])
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(B=64, seed=1234, res=(256, 64, 2), L=(512e-9, 128e-9, 3e-9)):
    n = int(np.prod(res))
    m0 = np.empty((B, 3 * n))
    for b in range(B):
        rng = np.random.default_rng(seed + b)
        m = rng.normal(size=(n, 3))
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        m0[b] = np.concatenate([m[:, 0], m[:, 1], m[:, 2]])
    fields = np.zeros((B, 3))
    fields[:, 0] = (-100.0 + 200.0 * np.arange(B) / max(B - 1, 1)) / 1000.0 / MU0
    return dict(res=res, L=L, n=n, B=B, m0=m0, fields=fields)


def make_solver(W, precision, t_end=1e-9, toeplitz=0):
    from magtense_amd import _lib

    L = _lib.lib()
    n, B = W["n"], W["B"]
    keep = {}
    keep["Ms"] = np.full(n, 8e5)
    keep["t"] = np.array([0.0, t_end])
    hext = np.zeros((B, 2, 4))
    hext[:, 0, 0] = 0.0
    hext[:, 1, 0] = t_end
    hext[:, :, 1:] = W["fields"][:, None, :]
    keep["hext"] = np.ascontiguousarray(hext)
    keep["m0"] = np.ascontiguousarray(W["m0"])
    keep["alphat"] = np.zeros((1, 2))
    keep["tconv"] = np.zeros(1)
    q = _lib.MtProblem()
    q.nx, q.ny, q.nz = W["res"]
    q.Lx, q.Ly, q.Lz = W["L"]
    q.n_states = B
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    q.Ms = dp(keep["Ms"])
    q.A0 = 1.3e-11
    q.gamma = 2.21e5
    q.alpha0 = 0.02 * 2.21e5
    q.nt_alpha = 1
    q.alphat = dp(keep["alphat"])
    q.solver = 2
    q.nt_hext = 2
    q.hext = dp(keep["hext"])
    q.hext_per_state = 1
    q.nt = 2
    q.t = dp(keep["t"])
    q.m0 = dp(keep["m0"])
    q.m0_per_state = 1
    q.tol = 1e-6
    q.thres = 1e-6
    q.nt_conv = 1
    q.t_conv = dp(keep["tconv"])
    q.conv_tol = 1e-4
    q.precision = precision
    q.toeplitz = toeplitz
    q.return_fields = 0
    h = C.c_void_p()
    _lib.check(L.mt_solver_create(C.byref(q), C.byref(h)))
    return L, h, keep


def stats(L, h, B):
    from magtense_amd import _lib

    st = (_lib.MtStats * B)()
    _lib.check(L.mt_solver_results(h, None, None, None, C.cast(st, C.c_void_p)))
    return st


def time_steps(L, h, steps, B):
    """Advance until B*steps accepted state-steps were taken (a full batch the whole time)."""
    from magtense_amd import _lib

    if steps <= 0:
        return 0.0, 0.0
    s0 = L.mt_solver_steps_total(h)
    t0 = time.perf_counter()
    _lib.check(L.mt_solver_advance_total(h, steps * B))
    wall = time.perf_counter() - t0
    return wall, (L.mt_solver_steps_total(h) - s0) / B


def step_breakdown(st, gemm_launches, steps):
    """(c) LL right-hand side and (d) stepper vector operations of a timed run:
    event time and algorithmic HBM bytes per launch and per batch step."""
    def part(ms, launches, nbytes, kernel):
        per = max(launches, 1)
        return {"kernel": kernel, "launches_per_batch_step": launches / steps,
                "ms_per_batch_step": ms / steps, "ms_per_launch": ms / per,
                "alg_bytes_per_launch": nbytes / per,
                "achieved_gbs": nbytes / max(ms, 1e-9) / 1e6,
                "frac": nbytes / max(ms, 1e-9) / 1e6 / HBM_PEAK_GBS}
    if st.llg_launches == 0:  # evaluated by the near-band kernel after each GEMM (its time is in near_band)
        rhs = {"kernel": "in near_corr_kernel<true, 2, true>", "launches_per_batch_step": 0, "ms_per_batch_step": 0.0,
               "note": "the LL right-hand side is evaluated by the near-band correction kernel; see near_band"}
    else:
        rhs = part(st.llg_ms, st.llg_launches, st.llg_bytes, "llg_kernel")
    return {"batch_steps_timed": steps, "gemm_launches_per_batch_step": gemm_launches / steps, "rhs": rhs,
            "stepper": part(st.vec_ms, st.vec_launches, st.vec_bytes, "vec_kernel + finalize_kernel")}


def coherent_states(pts, L):
    """Magnetizations whose far fields add up coherently: uniform +x, +z and
    (1, 1, 1)/sqrt(3), and a smooth in-plane pattern; (S, 3, N)."""
    n = pts.shape[0]
    x, y = pts[:, 0] / L[0], pts[:, 1] / L[1]
    out = []
    for v in ([1, 0, 0], [0, 0, 1], [1, 1, 1]):
        u = np.array(v, dtype=np.float64) / np.linalg.norm(v)
        out.append(np.repeat(u[:, None], n, axis=1))
    phi = 0.6 * np.sin(np.pi * x) * np.cos(np.pi * y) + 0.3 * np.tanh(6 * y)
    sm = np.stack([np.cos(phi), np.sin(phi), 0.05 * np.cos(2 * np.pi * x)])
    out.append(sm / np.linalg.norm(sm, axis=0))
    return ["uniform +x", "uniform +z", "uniform (1,1,1)", "smooth in-plane"], np.stack(out)


def field_error_vs_fp64(W, n_rows=1024, n_states=64, seed=7):
    """Relative demag-field error of the split-BF16 products (dense fragments and
    translation-invariant table) and the FP32 product vs an FP64 product on
    sampled receiver rows (FP64 tensor rows from the GPU kernel), for the
    benchmark's random states and for coherent states: norm-wise over all rows
    and states, and the largest per-receiver relative error |H_i - H_ref_i| /
    |H_ref_i|."""
    import torch
    from magtense_amd import _lib

    L = _lib.lib()
    n = W["n"]
    res, Lg = W["res"], W["L"]
    d = [Lg[k] / res[k] for k in range(3)]
    ax = [np.array([-Lg[k] / 2 + i * d[k] + d[k] / 2 for i in range(res[k])]) for k in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(n, n_rows, replace=False))
    dp = torch.as_tensor(pts).cuda()
    dsub = torch.as_tensor(np.ascontiguousarray(pts[rows])).cuda()
    dabc = torch.as_tensor(np.array(d)).cuda()
    K64 = torch.empty(6 * n_rows * n, dtype=torch.float64, device="cuda")
    _lib.check(L.mt_prism_tensor(dp.data_ptr(), dabc.data_ptr(), 0, None, dsub.data_ptr(), n, n_rows, None,
                                 K64.data_ptr(), None))
    K32 = torch.empty(6 * n * n, dtype=torch.float32, device="cuda")
    split = torch.empty(L.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda")
    _lib.check(L.mt_prism_tensor(dp.data_ptr(), dabc.data_ptr(), 0, None, dp.data_ptr(), n, n, K32.data_ptr(), None,
                                 split.data_ptr()))
    tab = torch.empty(L.mt_toeplitz_tensor_bytes(*res), dtype=torch.uint8, device="cuda")
    _lib.check(L.mt_toeplitz_tensor(*res, *Lg, tab.data_ptr()))
    f32f = torch.empty(L.mt_split_tensor_bytes(n), dtype=torch.uint8, device="cuda")
    _lib.check(L.mt_f32_mfma_tensor(K32.data_ptr(), n, f32f.data_ptr()))
    names, coh = coherent_states(pts, Lg)
    sets = {"random": W["m0"][:n_states].reshape(n_states, 3, n), "coherent": coh}
    Kr = K64.cpu().numpy().reshape(6, n_rows, n)
    comp = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    out = dict(rows=n_rows, random_states=n_states, coherent_states=names)
    for set_name, m in sets.items():
        S = m.shape[0]
        dm = torch.as_tensor(m.astype(np.float32)).cuda()
        ref = np.zeros((S, 3, n_rows))
        for a in range(3):
            for b in range(3):
                ref[:, a] -= m[:, b] @ Kr[comp[a][b]].T
        h = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        for kind in ("fp32", "fp32_mfma", "bf16x2", "bf16x2_table", "bf16x2_near", "bf16x2_table_near"):
            if kind == "fp32":
                _lib.check(L.mt_demag_apply_f32(K32.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
            elif kind == "fp32_mfma":
                _lib.check(L.mt_demag_apply_f32_mfma(f32f.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
            elif kind.startswith("bf16x2_table"):
                _lib.check(L.mt_demag_apply_bf16x2_toeplitz(tab.data_ptr(), *res, dm.data_ptr(), h.data_ptr(), S,
                                                            -1.0))
            else:
                _lib.check(L.mt_demag_apply_bf16x2(split.data_ptr(), n, dm.data_ptr(), h.data_ptr(), S, -1.0))
            if kind.endswith("_near"):  # the solver's default on uniform grids (radius 2)
                _lib.check(L.mt_demag_near_correction(*res, *Lg, 2, dm.data_ptr(), h.data_ptr(), S, -1.0))
            g = h.cpu().numpy().reshape(S, 3, n)[:, :, rows]
            e = g - ref
            norm_rel = [float(np.linalg.norm(e[s]) / np.linalg.norm(ref[s])) for s in range(S)]
            hn = np.linalg.norm(ref, axis=1)
            row_rel = np.linalg.norm(e, axis=1) / hn
            row_to_max = np.linalg.norm(e, axis=1) / hn.max(axis=1, keepdims=True)
            key = f"{kind}_{set_name}"
            out[key + "_rel"] = float(np.linalg.norm(e) / np.linalg.norm(ref))
            out[key + "_row_max_rel"] = float(row_rel.max())
            out[key + "_row_max_rel_to_max_field"] = float(row_to_max.max())
            if set_name == "coherent":
                out[key + "_rel_per_state"] = norm_rel
                out[key + "_row_max_rel_per_state"] = [float(v) for v in row_rel.max(axis=1)]
                out[key + "_row_max_rel_to_max_field_per_state"] = [float(v) for v in row_to_max.max(axis=1)]
    del K32, split, tab, f32f
    # headline figures: the solver's product (dense split-BF16 + near band) over every state, and FP32
    out["bf16x2_rel"] = max(out["bf16x2_near_random_rel"], max(out["bf16x2_near_coherent_rel_per_state"]))
    out["fp32_rel"] = max(out["fp32_random_rel"], max(out["fp32_coherent_rel_per_state"]))
    out["fp32_mfma_rel"] = max(out["fp32_mfma_random_rel"], max(out["fp32_mfma_coherent_rel_per_state"]))
    out["note"] = ("rel: norm-wise over the sampled rows; row_max_rel: largest per-receiver relative error "
                   "(ill-conditioned where a receiver's field nearly cancels); row_max_rel_to_max_field: largest "
                   "per-receiver error over the largest field; _near: with the radius-2 near-band correction the "
                   "solver applies on uniform grids")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(W, rows=2048):
    """FP64 oracle (FP32 demag semantics, 9 gemv-equivalents per RHS) on the host
    cores: the demag products of one full batch step (7 RHS x all B states) for
    `rows` of the N receiver rows, extrapolated to all rows."""
    from oracle import oracle as O

    n = W["n"]
    res, Lg = W["res"], W["L"]
    pts, d = O.setup_grid(list(res), list(Lg))
    rng = np.random.default_rng(3)
    sel = np.sort(rng.choice(n, rows, replace=False))
    abc = np.tile(d, (n, 1))
    t0 = time.perf_counter()
    Krows = O.tensor_points(pts, abc, pts[sel]).astype(np.float32)
    t_tensor = time.perf_counter() - t0
    ms = np.full(rows, 8e5, dtype=np.float32)
    t0 = time.perf_counter()
    for _rhs in range(7):
        for b in range(W["B"]):
            O.demag_rows(Krows, ms, W["m0"][b])
    t_sample = time.perf_counter() - t0
    t_batch_step = t_sample * n / rows
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=1.0 / t_batch_step, unit="LL batch-steps/s (64 states)", cores=cores, kind="port",
                extrapolated=True, cpu_model=cpu_model(), nproc=os.cpu_count(),
                sample=(f"FP64 oracle demag products (FP32 tensor, FP64-exact sums, reference gemv order) of one "
                        f"batch step (7 RHS x {W['B']} states) for {rows} of {n} receiver rows: {t_sample:.2f} s "
                        f"on {cores} threads; extrapolated x{n / rows:g} rows (tensor rows took {t_tensor:.1f} s, "
                        f"not counted)"))


def shard_shape(W, shards=8, shard=3, reps=10):
    """The batched split-BF16 product at one row shard's shape: N/shards receiver
    rows x all N sources x B states, from the shard's own block of fragment rows
    (assembled for its receivers only, as a GPU holding one shard would)."""
    import torch
    from magtense_amd import _lib

    L = _lib.lib()
    n, B = W["n"], W["B"]
    rows = n // shards
    row0 = shard * rows
    res, Lg = W["res"], W["L"]
    d = [Lg[k] / res[k] for k in range(3)]
    ax = [np.array([-Lg[k] / 2 + i * d[k] + d[k] / 2 for i in range(res[k])]) for k in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    np32 = (n + 31) // 32 * 32
    blk = torch.empty((rows + 31) // 32 * 32 * np32 * 24, dtype=torch.uint8, device="cuda")
    dp = torch.as_tensor(pts).cuda()
    dsub = torch.as_tensor(np.ascontiguousarray(pts[row0:row0 + rows])).cuda()
    dabc = torch.as_tensor(np.array(d)).cuda()
    t0 = time.perf_counter()
    _lib.check(L.mt_prism_tensor(dp.data_ptr(), dabc.data_ptr(), 0, None, dsub.data_ptr(), n, rows, None, None,
                                 blk.data_ptr()))
    t_asm = time.perf_counter() - t0
    dm = torch.as_tensor(W["m0"].astype(np.float32)).cuda()
    hd = torch.empty(B * 3 * rows, dtype=torch.float32, device="cuda")
    args = (blk.data_ptr(), n, row0, rows, dm.data_ptr(), hd.data_ptr(), B, -1.0)
    _lib.check(L.mt_demag_apply_bf16x2_rows(*args))
    t0 = time.perf_counter()
    for _ in range(reps):
        _lib.check(L.mt_demag_apply_bf16x2_rows(*args))  # synchronous
    ms = (time.perf_counter() - t0) / reps * 1e3
    alg = 6.0 * rows * n * 4 + 3.0 * n * B * 4 + 3.0 * rows * B * 4
    del blk
    return dict(shape=f"{rows} receiver rows (shard {shard} of {shards}) x {n} sources x {B} states",
                block_bytes=float((rows + 31) // 32 * 32 * np32 * 24), assembly_ms=1e3 * t_asm, call_ms=ms,
                achieved_gbs=alg / (ms * 1e-3) / 1e9, alg_bytes=alg,
                note="wall time per synchronous mt_demag_apply_bf16x2_rows call: state packing, the GEMM, the "
                     "partial-sum combine and the call's workspace allocation; the RCCL all-gather of the states "
                     "is not built")


def adjoint_c5(seed=5, nsteps=2000, dt=1e-13, every=100):
    """The differentiation configuration: 128 x 32 x 1 film (N = 4096), per-cell
    Ms ~ U[7e5, 9e5] A/m, FP32 demag, 2000 fixed RK4 steps of 0.1 ps; loss =
    mean m_z(T) + |m(T) - m_target|; gradients w.r.t. the applied field (3) and
    Ms (4096) by the discrete adjoint with a checkpoint every 100 steps.  Wall
    times of the forward pass and of forward + backward."""
    import torch
    from magtense_amd.adjoint import ll_rk4

    res, Lg = [128, 32, 1], [500e-9, 125e-9, 3e-9]
    n = 4096
    rng = np.random.default_rng(seed)
    unit = lambda a: a / np.linalg.norm(a, axis=1, keepdims=True)  # noqa: E731
    # smooth near-uniform start (from random unit vectors the exchange-driven
    # 2000-step trajectory is chaotic and the gradients meaningless)
    x, y = np.meshgrid(np.arange(res[0]) / res[0], np.arange(res[1]) / res[1])
    m0 = np.stack([np.ones(n), 0.2 * np.sin(2 * np.pi * x).ravel(), 0.1 * np.cos(2 * np.pi * y).ravel()], axis=1)
    m0 = torch.tensor(unit(m0))
    target = torch.tensor(unit(rng.normal(size=(n, 3))))
    ms = torch.tensor(rng.uniform(7e5, 9e5, n), requires_grad=True)
    h = torch.tensor(np.array([-24.6, 4.3, 0.0]) / 1000 / MU0, requires_grad=True)
    t0 = time.perf_counter()
    mT = ll_rk4(m0, h, ms, res, Lg, dt, nsteps, checkpoint_every=every)
    t_fwd = time.perf_counter() - t0
    loss = mT[:, 2].mean() + torch.linalg.norm(mT - target)
    loss.backward()
    t_all = time.perf_counter() - t0
    return dict(workload=f"128x32x1 film, N = {n}, {nsteps} RK4 steps of {dt:g} s, checkpoint every {every}, FP32 demag",
                forward_s=t_fwd, forward_backward_s=t_all, forward_rhs_per_s=4 * nsteps / t_fwd,
                grad_h_norm=float(h.grad.norm()), grad_ms_norm=float(ms.grad.norm()),
                note="errors vs central differences and the FP64 oracle adjoint: tests/test_gpu_adjoint.py")


def next_rows(W, reps=5):
    """The widened rows beside the hot path, timed on the GPU (wall time of the C-ABI
    and Python-API calls, inputs resident where the API takes device arrays):
    * the threshold-sparse demag product (threshold fraction 0.9) on the benchmark film
      (N = 32768), 64 states: build time, stored entries, time per product;
    * the magnetostatic field of the 16384-prism lattice at its 16384 centres
      (`magstatics.get_H_field`, host arrays in and out);
    * the tetrahedron tile tensor of 2048 tetrahedra at 2048 points
      (`magstatics.get_demag_tensor`, tile type 5, host arrays);
    * the demag-tensor file: save and load of the 128 x 32 film's six FP32 N^2 records."""
    import tempfile

    import torch
    from magtense_amd import _lib
    from magtense import magstatics as ms
    from magtense import micromag

    L = _lib.lib()
    out = {}
    n, res, Lg = W["n"], W["res"], W["L"]
    d = [Lg[k] / res[k] for k in range(3)]
    ax = [np.array([-Lg[k] / 2 + i * d[k] + d[k] / 2 for i in range(res[k])]) for k in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    dp = torch.as_tensor(pts).cuda()
    dabc = torch.as_tensor(np.array(d)).cuda()
    K32 = torch.empty(6 * n * n, dtype=torch.float32, device="cuda")
    _lib.check(L.mt_prism_tensor(dp.data_ptr(), dabc.data_ptr(), 0, None, dp.data_ptr(), n, n, K32.data_ptr(),
                                 None, None))
    torch.cuda.synchronize()
    h = C.c_void_p()
    t0 = time.perf_counter()
    _lib.check(L.mt_demag_sparse_create(K32.data_ptr(), n, 4, 0.9, C.byref(h)))
    t_build = time.perf_counter() - t0
    hb = C.c_void_p()
    t0 = time.perf_counter()
    _lib.check(L.mt_demag_sparse_create_bf16x2(K32.data_ptr(), n, 4, 0.9, res[0], C.byref(hb)))
    t_build_b = time.perf_counter() - t0
    hf = C.c_void_p()
    t0 = time.perf_counter()
    _lib.check(L.mt_demag_sparse_create_f32_tiles(K32.data_ptr(), n, 4, 0.9, res[0], C.byref(hf)))
    t_build_f = time.perf_counter() - t0
    del K32
    torch.cuda.empty_cache()
    try:
        nnz = (C.c_longlong * 6)()
        thr = C.c_float()
        _lib.check(L.mt_demag_sparse_info(h, nnz, C.byref(thr)))
        S = W["B"]
        dm = torch.as_tensor(W["m0"][:S].reshape(S, 3 * n).astype(np.float32)).cuda()
        dh = torch.empty(S * 3 * n, dtype=torch.float32, device="cuda")
        _lib.check(L.mt_demag_apply_sparse(h, dm.data_ptr(), dh.data_ptr(), S, -1.0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.mt_demag_apply_sparse(h, dm.data_ptr(), dh.data_ptr(), S, -1.0))
        torch.cuda.synchronize()
        ms_apply = (time.perf_counter() - t0) / reps * 1e3
        # one state (the reference's per-state product)
        _lib.check(L.mt_demag_apply_sparse(h, dm.data_ptr(), dh.data_ptr(), 1, -1.0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.mt_demag_apply_sparse(h, dm.data_ptr(), dh.data_ptr(), 1, -1.0))
        torch.cuda.synchronize()
        ms_apply1 = (time.perf_counter() - t0) / reps * 1e3
        # the same kept entries as split-BF16 tiles: batched MFMA product over the stored tiles
        ent, tiles_ne, nbytes = C.c_longlong(), C.c_longlong(), C.c_longlong()
        ksb = C.c_int()
        _lib.check(L.mt_demag_sparse_blocks(hb, C.byref(ent), C.byref(tiles_ne), C.byref(nbytes), C.byref(ksb)))
        dhb = torch.empty_like(dh)
        _lib.check(L.mt_demag_apply_sparse(hb, dm.data_ptr(), dhb.data_ptr(), S, -1.0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.mt_demag_apply_sparse(hb, dm.data_ptr(), dhb.data_ptr(), S, -1.0))
        torch.cuda.synchronize()
        ms_apply_b = (time.perf_counter() - t0) / reps * 1e3
        rel_b = float(torch.linalg.norm(dhb - dh) / torch.linalg.norm(dh))
        # the same kept entries as FP32 tiles (FP32 MFMAs): the FP32 product for many states
        entf = C.c_longlong()
        ksf = C.c_int()
        _lib.check(L.mt_demag_sparse_blocks(hf, C.byref(entf), None, None, C.byref(ksf)))
        dhf = torch.empty_like(dh)
        _lib.check(L.mt_demag_apply_sparse(hf, dm.data_ptr(), dhf.data_ptr(), S, -1.0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.mt_demag_apply_sparse(hf, dm.data_ptr(), dhf.data_ptr(), S, -1.0))
        torch.cuda.synchronize()
        ms_apply_f = (time.perf_counter() - t0) / reps * 1e3
        rel_f = float(torch.linalg.norm(dhf - dh) / torch.linalg.norm(dh))
    finally:
        L.mt_demag_sparse_destroy(h)
        L.mt_demag_sparse_destroy(hb)
        L.mt_demag_sparse_destroy(hf)
    stored = float(sum(nnz))
    alg = stored * 8.0 + 2.0 * S * 3 * n * 4  # value + column per stored entry, states in, fields out
    out["sparse_demag"] = {"workload": f"{res[0]}x{res[1]}x{res[2]} film, N = {n}, {S} states, threshold fraction 0.9",
                           "build_s": t_build, "stored_entries": stored, "fraction_of_dense": stored / (6.0 * n * n),
                           "threshold": thr.value, "ms_per_product": ms_apply,
                           "alg_bytes_per_product": alg, "achieved_gbs": alg / (ms_apply * 1e-3) / 1e9,
                           "ms_per_product_one_state": ms_apply1,
                           "one_state_gbs": (stored * 8.0 + 2.0 * 3 * n * 4) / (ms_apply1 * 1e-3) / 1e9}
    np32 = (n + 31) // 32 * 32
    tiles_all = (np32 // 32) * (np32 // 16)
    alg_b = float(nbytes.value) + 2.0 * S * 3 * n * 4  # stored fragments + states in, fields out
    out["sparse_demag_bf16x2_tiles"] = {
        "workload": out["sparse_demag"]["workload"] + ", kept entries as split-BF16 32x16 tiles",
        "build_s": t_build_b, "tiles_nonempty": int(tiles_ne.value), "tile_density": tiles_ne.value / tiles_all,
        "list_entries": int(ent.value), "stored_fraction_of_dense": 4.0 * ent.value / tiles_all,
        "fragment_bytes": float(nbytes.value), "ksplit": int(ksb.value), "ms_per_product": ms_apply_b,
        "rel_diff_vs_fp32_csr": rel_b,
        "roofline": {"bound": "hbm", "kernel": "gemm_bf16x2_reg_kernel<4, false, true, false>",
                     "achieved": alg_b / (ms_apply_b * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg_b / (ms_apply_b * 1e-3) / 1e9 / HBM_PEAK_GBS, "alg_bytes_per_launch": alg_b},
        "note": "wall time per synchronous mt_demag_apply_sparse call (state packing, GEMM over the stored tiles, "
                "partial-sum combine, workspace allocation)"}
    # FP32 work over the stored tiles: 9 block products x 2 flop x 32 rows x 16 sources per tile and state
    flops_f = 18.0 * 4 * entf.value * 32 * 16 * S
    tf_f = flops_f / (ms_apply_f * 1e-3) / 1e12
    out["sparse_demag_f32_tiles"] = {
        "workload": out["sparse_demag"]["workload"] + ", kept entries as FP32 32x16 tiles (FP32 MFMAs)",
        "build_s": t_build_f, "list_entries": int(entf.value), "ksplit": int(ksf.value), "ms_per_product": ms_apply_f,
        "speedup_vs_csr": ms_apply / ms_apply_f, "rel_diff_vs_fp32_csr": rel_f,
        "roofline": {"bound": "mfma", "kernel": "gemm_bf16x2_reg_kernel<4, false, true, true> (FP32 fragments)",
                     "achieved": tf_f, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf_f / FP32_PEAK_TFLOPS,
                     "alg_flops_per_launch": flops_f},
        "note": "wall time per synchronous mt_demag_apply_sparse call; FP32 products with FP32 accumulation "
                "(the reference's arithmetic class), one pass over the stored tiles for all states"}
    # magnetostatics: prism lattice field
    c, abc = lattice_prisms()
    nt = c.shape[0]
    rng = np.random.default_rng(11)
    t = ms.Tiles(nt, size=abc.tolist(), offset=c.tolist(), tile_type=2, M_rem=1.2 / MU0,
                 easy_axis=rng.normal(size=(nt, 3)).tolist())
    ms.get_H_field(t, c[:64])
    t0 = time.perf_counter()
    H = ms.get_H_field(t, c)
    t_h = time.perf_counter() - t0
    out["magnetostatic_field"] = {"workload": f"{nt} prism tiles (5 nm lattice, jittered) at their {nt} centres",
                                  "s": t_h, "pairs_per_s": nt * nt / t_h, "finite": bool(np.isfinite(H).all())}
    # tetrahedron tiles: tensor of 2048 tetrahedra at 2048 points
    ntt = 2048
    tt = ms.Tiles(ntt)
    tt.tile_type = [5] * ntt
    base = np.array([[0, 0, 0], [4e-9, 0, 0], [0, 3e-9, 0], [0, 0, 2e-9]], dtype=float)
    grid = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    verts = np.zeros((ntt, 3, 4))
    for i in range(ntt):
        verts[i] = (base + grid[i] * 5e-9 + rng.uniform(-0.3e-9, 0.3e-9, (4, 3))).T
    tt.vertices = verts
    tp = grid * 5e-9 + 1e-9
    ms.get_demag_tensor(tt, tp[:16])
    t0 = time.perf_counter()
    N = ms.get_demag_tensor(tt, tp)
    t_n = time.perf_counter() - t0
    out["tetrahedron_tensor"] = {"workload": f"{ntt} tetrahedron tiles at {tp.shape[0]} points, N(n_tiles, n_pts, 3, 3) FP64",
                                 "s": t_n, "pairs_per_s": ntt * tp.shape[0] / t_n,
                                 "finite": bool(np.isfinite(N).all())}
    # demag-tensor file (six FP32 N^2 records) for the 128 x 32 x 1 film
    nf = 4096
    Kf = np.random.default_rng(3).normal(size=(6, nf, nf)).astype(np.float32)
    with tempfile.TemporaryDirectory() as tmp:
        f = os.path.join(tmp, "K.dat")
        t0 = time.perf_counter()
        micromag.save_demag_tensor(f, Kf)
        t_w = time.perf_counter() - t0
        t0 = time.perf_counter()
        K2 = micromag.load_demag_tensor(f, nf)
        t_r = time.perf_counter() - t0
    out["tensor_file"] = {"workload": f"N = {nf}: six FP32 {nf}^2 records ({Kf.nbytes / 1e6:.1f} MB)",
                          "save_gbs": Kf.nbytes / t_w / 1e9, "load_gbs": Kf.nbytes / t_r / 1e9,
                          "round_trip_exact": bool(np.array_equal(K2, Kf))}
    return out


def lattice_prisms(seed=4, side=128):
    """16384 prisms: side x side lattice of pitch 5 nm, centres jittered by
    U[-0.5, 0.5] nm in x and y, edges U[0.8, 1.2] * (5, 5, 2) nm."""
    rng = np.random.default_rng(seed)
    ix, iy = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    c = np.zeros((side * side, 3))
    c[:, 0] = ix.reshape(-1) * 5e-9 + rng.uniform(-0.5e-9, 0.5e-9, side * side)
    c[:, 1] = iy.reshape(-1) * 5e-9 + rng.uniform(-0.5e-9, 0.5e-9, side * side)
    abc = rng.uniform(0.8, 1.2, (side * side, 3)) * np.array([5e-9, 5e-9, 2e-9])
    return c, abc


def tensor_assembly(reps=3, cpu_rows=64):
    """Prism-tensor assembly on the 16384-prism jittered lattice: all N^2 pairs,
    six FP32 components stored (6.44 GB); pairs/s on the GPU and, for a sample
    of receiver rows, in the FP64 oracle on the host cores."""
    import torch
    from magtense_amd import _lib

    L = _lib.lib()
    c, abc = lattice_prisms()
    n = c.shape[0]
    dc = torch.as_tensor(c).cuda()
    da = torch.as_tensor(abc).cuda()
    K32 = torch.empty(6 * n * n, dtype=torch.float32, device="cuda")
    args = (dc.data_ptr(), da.data_ptr(), 3, None, dc.data_ptr(), n, n, K32.data_ptr(), None, None)
    _lib.check(L.mt_prism_tensor(*args))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.check(L.mt_prism_tensor(*args))  # synchronous
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    del K32
    out = dict(workload="16384 prisms, 128x128 lattice pitch 5 nm, jittered; 6 x N^2 FP32 out",
               pairs=n * n, ms=1e3 * t, pairs_per_s=n * n / t, store_bytes=6.0 * n * n * 4,
               store_floor_ms=6.0 * n * n * 4 / (HBM_PEAK_GBS * 1e9) * 1e3, bound="fp64 valu (transcendentals)")
    if cpu_rows:
        from oracle import oracle as O

        rows = np.linspace(0, n - 1, cpu_rows).astype(int)
        t0 = time.perf_counter()
        O.tensor_points(c, abc, c[rows])
        tc = time.perf_counter() - t0
        # one thread, as the reference assembles (omp_set_num_threads(1),
        # LandauLifshitzEquationSolver.f90:927)
        prev = O.set_threads(1)
        t0 = time.perf_counter()
        O.tensor_points(c, abc, c[rows[:cpu_rows // 4]])
        t1 = time.perf_counter() - t0
        O.set_threads(prev)
        out["cpu_pairs_per_s"] = cpu_rows * n / tc
        out["cpu_pairs_per_s_1thread"] = (cpu_rows // 4) * n / t1
        out["cpu_sample"] = (f"oracle prism tensor, {cpu_rows} rows x {n} sources, {tc:.2f} s on {prev} threads; "
                             f"{cpu_rows // 4} rows on 1 thread, {t1:.2f} s")
    return out


def _unit_rows(rng, n):
    m = rng.normal(size=(n, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def _soa(m):
    return np.concatenate([m[:, 0], m[:, 1], m[:, 2]])


def _gpu_run(res, L, m0, h_fct, t_end, nt, nt_h, timing=False, **kw):
    """One end-to-end run through the drop-in API: (wall s, stats, M)."""
    from magtense.micromag import MicromagProblem

    p = MicromagProblem(list(res), list(L), m0=m0, **kw)
    p._timing = timing
    t0 = time.perf_counter()
    _, M, *_ = p.run_simulation(t_end, nt, h_fct, nt_h)
    return time.perf_counter() - t0, p.last_stats[0], np.squeeze(M, axis=2)


def _oracle_run(res, L, m0, hext, t, **kw):
    from oracle import oracle as O

    q = O.Problem(list(res), list(L), hext=hext, t=t, m0=_soa(m0), return_fields=False, **kw)
    t0 = time.perf_counter()
    _, y, _, st = q.run()
    return time.perf_counter() - t0, st, np.transpose(y[0, -1])


def _line(wall, st, timed_st, n, cells_bytes):
    """steps/s, RHS/s and the per-step time split of one GPU run."""
    steps = max(st["steps"], 1)
    out = dict(wall_s=wall, steps=st["steps"], rhs=st["f_calls"], steps_per_s=st["steps"] / wall,
               rhs_per_s=st["f_calls"] / wall, ms_per_step=1e3 * wall / steps)
    if timed_st is not None:
        ts = max(timed_st["steps"], 1)
        passes = max(timed_st["demag_passes"], 1)
        out["kernel_ms_per_step"] = {"demag": timed_st["demag_ms"] / ts, "llg_kernel": timed_st["llg_ms"] / ts,
                                     "stepper": timed_st["vec_ms"] / ts}
        out["demag_ms_per_launch"] = timed_st["demag_ms"] / passes
        kern = (timed_st["demag_ms"] + timed_st["llg_ms"] + timed_st["vec_ms"]) / ts
        out["gpu_kernel_frac_of_step"] = kern / out["ms_per_step"]
        if cells_bytes:
            gbs = cells_bytes / (out["demag_ms_per_launch"] * 1e-3) / 1e9
            out["roofline"] = {"bound": "hbm", "kernel": "demag_llg_f32_kernel (FP32 demag rows + LL right-hand side)",
                               "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                               "alg_bytes_per_launch": cells_bytes}
    return out


def small_configs(no_cpu=False):
    """The plain-method configurations end to end through MicromagProblem (FP32
    demag product, RKSUITE 'M' tol 1e-6), with the FP64 CPU oracle on the host
    cores:
    * 32 x 8 x 1 film (N = 256): seeded random start relaxed toward +x (gamma =
      0, +x field ramped 100 mT -> 0 over 1 ns, 5 ns), then 1 ns of dynamics at
      (-24.6, 4.3, 0) mT; the oracle runs both stages in full;
    * 128 x 32 x 1 film (N = 4096): seeded random start, seeded in-plane field of
      10-50 mT, 2 ns; the oracle's time is its measured right-hand-side cost on
      this problem times the run's right-hand-side count (extrapolated).
    Each GPU run is repeated with HIP-event kernel timing for the per-step split."""
    out = {}
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # ---- 32 x 8 x 1 -------------------------------------------------------------
    res, L = [32, 8, 1], [500e-9, 125e-9, 3e-9]
    m0 = _unit_rows(np.random.default_rng(1234), 256)
    hx = np.array([100.0, 0.0, 0.0]) / 1000 / MU0
    h = np.array([-24.6, 4.3, 0.0]) / 1000 / MU0

    def h_ini(t):
        return np.expand_dims(np.where(t < 1e-9, 1.0 - t / 1e-9, 0), axis=1) * hx

    def h_dyn(t):
        return np.ones((len(t), 1)) * h

    kw = dict(alpha=4.42e3, tol=1e-6)
    _gpu_run(res, L, m0, h_ini, 1e-10, 5, 20, gamma=0.0, **kw)  # warm-up
    w_rel, st_rel, M1 = _gpu_run(res, L, m0, h_ini, 5e-9, 50, 200, gamma=0.0, **kw)
    m_rel = M1[-1]
    w_dyn, st_dyn, _ = _gpu_run(res, L, m_rel, h_dyn, 1e-9, 200, 2000, gamma=2.21e5, **kw)
    _, st_t, _ = _gpu_run(res, L, m_rel, h_dyn, 1e-9, 200, 2000, timing=True, gamma=2.21e5, **kw)
    c1 = {"workload": "32x8x1 film (15.6x15.6x3 nm), N = 256: relaxation 5 ns (gamma 0) + 1 ns dynamics, FP32 demag",
          "relax": _line(w_rel, st_rel, None, 256, 0), "dynamics": _line(w_dyn, st_dyn, st_t, 256, 0),
          "note": "the 1.6 MB tensor stays in L2: the step is bound by kernel latency and the host's launch "
                  "submissions (one launch per right-hand side: stage combination, FP32 demag rows and LL "
                  "right-hand side in one kernel) and the one host round trip per step of the RKSUITE controller, "
                  "not by HBM; kernel_ms_per_step.demag includes the stage combination and the LL right-hand side"}
    if not no_cpu:
        ht = np.linspace(0, 5e-9, 200)
        hext = np.zeros((200, 4))
        hext[:, 0] = ht
        hext[:, 1:] = h_ini(ht)
        wo_rel, so_rel, mo_rel = _oracle_run(res, L, m0, hext, np.linspace(0, 5e-9, 50), alpha=4.42e3, gamma=0.0,
                                             tol=1e-6)
        ht = np.linspace(0, 1e-9, 2000)
        hext = np.zeros((2000, 4))
        hext[:, 0] = ht
        hext[:, 1:] = h
        wo_dyn, so_dyn, _ = _oracle_run(res, L, mo_rel, hext, np.linspace(0, 1e-9, 200), alpha=4.42e3,
                                        gamma=2.21e5, tol=1e-6)
        c1["cpu_baseline"] = {"kind": "port", "cores": cores, "relax_s": wo_rel, "dynamics_s": wo_dyn,
                              "dynamics_steps_per_s": so_dyn["steps"] / wo_dyn,
                              "dynamics_rhs_per_s": so_dyn["f_calls"] / wo_dyn, "extrapolated": False,
                              "sample": "FP64 oracle, both stages in full"}
        c1["speedup_vs_cpu"] = (wo_rel + wo_dyn) / (w_rel + w_dyn)
    out["32x8x1"] = c1
    # ---- 128 x 32 x 1 -------------------------------------------------------------
    res = [128, 32, 1]
    rng = np.random.default_rng(42)
    m0 = _unit_rows(rng, 4096)
    theta = rng.uniform(0, 2 * np.pi)
    mag = rng.uniform(10, 50)
    h2 = mag * np.array([np.cos(theta), np.sin(theta), 0.0]) / 1000 / MU0

    def h_c2(t):
        return np.ones((len(t), 1)) * h2

    n = 4096
    kw2 = dict(alpha=4.42e3, gamma=2.21e5, tol=1e-6)
    _gpu_run(res, L, m0, h_c2, 1e-12, 2, 2, **kw2)  # warm-up
    w2, st2, _ = _gpu_run(res, L, m0, h_c2, 2e-9, 21, 10, **kw2)
    _, st2t, _ = _gpu_run(res, L, m0, h_c2, 2e-9, 21, 10, timing=True, **kw2)
    c2 = {"workload": f"128x32x1 film (3.9x3.9x3 nm), N = 4096, {mag:.1f} mT in-plane at "
                      f"{np.degrees(theta):.0f} deg, 2 ns from a random start, FP32 demag (402.7 MB tensor)",
          "run": _line(w2, st2, st2t, n, 6.0 * n * n * 4 + 2 * 3 * n * 4)}
    if not no_cpu:
        # two oracle runs of the same problem over 1 ps and 4 ps: both assemble
        # the tensor once, so the difference of their times over the difference
        # of their right-hand-side counts is the integration cost per RHS
        hext2 = np.array([[0.0, *h2], [2e-9, *h2]])
        wa, sa, _ = _oracle_run(res, L, m0, hext2, np.linspace(0, 1e-12, 2), **kw2)
        wb, sb, _ = _oracle_run(res, L, m0, hext2, np.linspace(0, 4e-12, 2), **kw2)
        t_rhs = (wb - wa) / max(sb["f_calls"] - sa["f_calls"], 1)
        t_asm = max(wa - t_rhs * sa["f_calls"], 0.0)  # tensor assembly and set-up, once per run
        wall_cpu = t_asm + t_rhs * st2["f_calls"]
        c2["cpu_baseline"] = {"kind": "port", "cores": cores, "rhs_s": t_rhs, "setup_s": t_asm, "run_s": wall_cpu,
                              "steps_per_s": st2["steps"] / wall_cpu, "rhs_per_s": 1.0 / t_rhs,
                              "extrapolated": True,
                              "sample": f"FP64 oracle (FP32 demag semantics) runs of this problem over 1 ps "
                                        f"({sa['f_calls']} RHS, {wa:.1f} s) and 4 ps ({sb['f_calls']} RHS, "
                                        f"{wb:.1f} s): per-RHS time from their difference, x the GPU run's "
                                        f"{st2['f_calls']} right-hand sides, plus the set-up time"}
        c2["speedup_vs_cpu"] = wall_cpu / w2
    out["128x32x1"] = c2
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--mfma32-steps", type=int, default=3,
                    help="batch steps of the FP32 MFMA product (precision fp32_mfma) timed beside the others (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-field-error", action="store_true")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--no-shard", action="store_true",
                    help="skip the product at one row shard's shape (N/8 rows x N sources)")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the widened rows (sparse demag, magnetostatic field, tetrahedron tiles, tensor file)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiation configuration (RK4 adjoint gradient wall time)")
    ap.add_argument("--table-steps", type=int, default=30,
                    help="steps of the translation-invariant table path timed beside the dense tensor (0: skip)")
    ap.add_argument("--no-main", action="store_true",
                    help="skip the headline leg (dense split-BF16 solver): profiling runs of the other legs")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 32x8x1 and 128x32x1 plain-method configurations (end to end, with the CPU oracle)")
    ap.add_argument("--table", action="store_true",
                    help="time the translation-invariant table as the main path (default: the dense 25.8 GB tensor)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X (asked for --gpus {args.gpus}): the per-GPU product of the "
              "row-sharded tensor is mt_demag_apply_bf16x2_rows; runs over several GPUs are not launched.",
              file=sys.stderr)
        sys.exit(2)
    # torch ships its own HIP runtime; initialise it before libmagtense_hip's
    # (both then share the device; the reverse order fails to initialise torch)
    try:
        import torch

        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    except ImportError:
        pass
    from magtense_amd import _lib

    W = workload(B=args.batch)
    n, B = W["n"], W["B"]
    # ---- split-BF16 MFMA path (the measured value) ---------------------------------
    L = _lib.lib()
    steps_per_s = ms_step = wall = batch_steps = gemm_ms = t_setup = None
    passes = rejected = 0
    breakdown = None
    if not args.no_main:
        t0 = time.perf_counter()
        L, h, keep = make_solver(W, _lib.PREC_BF16X2, toeplitz=0 if args.table else -1)
        t_setup = time.perf_counter() - t0
        time_steps(L, h, args.warmup, B)
        st0 = stats(L, h, B)
        wall, batch_steps = time_steps(L, h, args.steps, B)
        st1 = stats(L, h, B)
        passes = st1[0].demag_passes - st0[0].demag_passes
        rejected = sum(st1[b].bad_steps - st0[b].bad_steps for b in range(B))
        # kernel duration with HIP events on the solver's stream (separate short run)
        L.mt_solver_set_timing(h, 1)
        sta = stats(L, h, B)
        _, timed_steps = time_steps(L, h, 2, B)
        stb = stats(L, h, B)
        L.mt_solver_set_timing(h, 0)
        gemm_launches = stb[0].demag_passes - sta[0].demag_passes
        near_ms = stb[0].near_ms / max(gemm_launches, 1)  # near-band correction after each GEMM
        gemm_ms = stb[0].demag_ms / max(gemm_launches, 1) - near_ms
        breakdown = step_breakdown(stb[0], gemm_launches, timed_steps)
        breakdown["near_band"] = {"kernel": ("near_corr_kernel<true, 2, true> (radius 2, with the LL right-hand side)"
                                             if stb[0].llg_launches == sta[0].llg_launches
                                             else "near_corr_kernel<true, 2, false> (radius 2)"),
                                  "ms_per_launch": near_ms,
                                  "ms_per_batch_step": near_ms * gemm_launches / timed_steps}
        L.mt_solver_destroy(h)
        del keep
        steps_per_s = batch_steps / wall
        ms_step = 1e3 * wall / batch_steps
        breakdown["other_ms_per_batch_step"] = (ms_step - passes / batch_steps * (gemm_ms + near_ms)
                                                - breakdown["rhs"]["ms_per_batch_step"]
                                                - breakdown["stepper"]["ms_per_batch_step"])
        breakdown["note"] = ("rhs = (c) LL kernel, stepper = (d) RKSUITE vector operations with their descriptor uploads, "
                             "timed with HIP events over 2 batch steps; other = the default run's step time less the "
                             "GEMM passes and those two: host scheduling and launch gaps")
        log(f"step breakdown: {json.dumps(breakdown)}")
        log(f"bf16x2: setup {t_setup:.2f} s, {args.steps} steps in {wall:.3f} s = {steps_per_s:.2f} steps/s; "
            f"demag passes {passes}, rejected {rejected}, gemm {gemm_ms:.3f} ms/launch")
    alg_bytes = 6.0 * n * n * 4 + 2 * 3 * n * B * 4
    alg_flops = 18.0 * n * n * B  # 9 block products x 2 flop per entry and state (split products not counted)
    # ---- the translation-invariant table (140 MB instead of 25.8 GB) beside it -----------
    table = None
    if args.table_steps > 0 and not args.table:
        Lt, ht, keept = make_solver(W, _lib.PREC_BF16X2, toeplitz=1)
        time_steps(Lt, ht, 2, B)
        wallt, bst = time_steps(Lt, ht, args.table_steps, B)
        Lt.mt_solver_set_timing(ht, 1)
        sa = stats(Lt, ht, B)
        time_steps(Lt, ht, 2, B)
        sb = stats(Lt, ht, B)
        Lt.mt_solver_set_timing(ht, 0)
        Lt.mt_solver_destroy(ht)
        del keept
        tpass = max(sb[0].demag_passes - sa[0].demag_passes, 1)
        tms = (sb[0].demag_ms - sb[0].near_ms) / tpass
        tf = alg_flops / (tms * 1e-3) / 1e12
        table = dict(steps_per_s=bst / wallt, steps=args.table_steps, gemm_ms=tms, kernel="gemm_bf16x2_tw_kernel",
                     roofline={"bound": "mfma", "achieved": tf, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": tf / BF16_PEAK_TFLOPS, "mfma_issued_frac": 3.0 * tf / BF16_PEAK_TFLOPS,
                               "traffic": traffic_from_profile("table")},
                     table_bytes=float(L.mt_toeplitz_tensor_bytes(*W["res"])),
                     note="same split-BF16 values; the uniform grid's tensor stored once per cell offset")
        log(f"table path: {bst / wallt:.2f} steps/s, gemm {tms:.3f} ms = {tf:.0f} TFLOP/s algorithmic")
    # ---- plain FP32 per-state path ------------------------------------------------------
    plain = None
    plain_info = None
    if not args.no_plain:
        L2, h2, keep2 = make_solver(W, _lib.PREC_FP32)
        time_steps(L2, h2, 1, B)
        sp0 = stats(L2, h2, B)
        wall2, bs2 = time_steps(L2, h2, args.plain_steps, B)
        sp1 = stats(L2, h2, B)
        L2.mt_solver_destroy(h2)
        del keep2
        plain = bs2 / wall2
        plain_passes = sp1[0].demag_passes - sp0[0].demag_passes
        plain_info = dict(steps_per_s=plain, batch_steps=bs2, wall_s=wall2, demag_passes=int(plain_passes),
                          passes_per_batch_step=plain_passes / max(bs2, 1e-9),
                          ms_per_pass=1e3 * wall2 / max(plain_passes, 1),
                          note="one pass = the 64 per-state FP32 mat-vecs of one right-hand side "
                               "(demag_llg_f32_kernel: the FP32 rows, 6 N^2 4 B per state, fused with the LL "
                               "right-hand side)")
        log(f"fp32 plain: {args.plain_steps} steps in {wall2:.3f} s = {plain:.4f} steps/s, {plain_passes} passes")
    # ---- batched FP32 MFMA (precision fp32_mfma): the reference's arithmetic class, one tensor pass per RHS
    mfma32 = None
    if args.mfma32_steps > 0:
        L3, h3, keep3 = make_solver(W, _lib.PREC_FP32_MFMA)
        time_steps(L3, h3, 1, B)
        s30 = stats(L3, h3, B)
        wall3, bs3 = time_steps(L3, h3, args.mfma32_steps, B)
        s31 = stats(L3, h3, B)
        L3.mt_solver_set_timing(h3, 1)
        s3a = stats(L3, h3, B)
        time_steps(L3, h3, 1, B)
        s3b = stats(L3, h3, B)
        L3.mt_solver_set_timing(h3, 0)
        L3.mt_solver_destroy(h3)
        del keep3
        p3 = max(s3b[0].demag_passes - s3a[0].demag_passes, 1)
        g3 = s3b[0].demag_ms / p3
        tf3 = alg_flops / (g3 * 1e-3) / 1e12
        mfma32 = dict(steps_per_s=bs3 / wall3, batch_steps=bs3, wall_s=wall3,
                      passes_per_batch_step=(s31[0].demag_passes - s30[0].demag_passes) / max(bs3, 1e-9),
                      gemm_ms=g3, speedup_vs_plain_fp32=(bs3 / wall3 / plain) if plain else None,
                      roofline={"bound": "mfma", "kernel": "gemm_bf16x2_reg_kernel<4, false, false, true> (FP32 fragments)",
                                "achieved": tf3, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": tf3 / FP32_PEAK_TFLOPS, "alg_flops_per_launch": alg_flops},
                      note="precision fp32_mfma: the FP32 tensor as MFMA fragments, FP32 products and sums, one "
                           "pass over the 25.8 GB tensor for all 64 states")
        log(f"fp32 mfma: {bs3 / wall3:.3f} steps/s, gemm {g3:.3f} ms = {tf3:.1f} TFLOP/s")
    cfgs = None if args.no_configs else small_configs(no_cpu=args.no_cpu)
    log(f"plain-method configurations: {json.dumps(cfgs)}")
    shard = None if args.no_shard else shard_shape(W)
    log(f"row shard shape: {shard}")
    adj = None if args.no_adjoint else adjoint_c5()
    log(f"adjoint (differentiation configuration): {adj}")
    nxt = None if args.no_next else next_rows(W)
    log(f"widened rows: {nxt}")
    rd, cp = C.c_double(0), C.c_double(0)
    _lib.check(L.mt_measure_bandwidth(8 << 30, 10, C.byref(rd), C.byref(cp)))
    rdr = C.c_double(0)
    _lib.check(L.mt_measure_bandwidth_pattern(8 << 30, 10, 1, C.byref(rdr), None))
    log(f"measured streaming bandwidth: read {rd.value:.0f} GB/s (constant bytes), {rdr.value:.0f} GB/s "
        f"(random bits), copy {cp.value:.0f} GB/s")
    ferr = None if args.no_field_error else field_error_vs_fp64(W)
    log(f"field error vs FP64: {ferr}")
    asm = None if args.no_assembly else tensor_assembly(cpu_rows=0 if args.no_cpu else 64)
    log(f"tensor assembly: {asm}")
    cpu = None if args.no_cpu else cpu_baseline(W)
    log(f"cpu baseline: {cpu}")
    if args.no_main:
        roofline = None
    elif not args.table:
        # the dense fragment-tiled tensor streams 25.8 GB per pass: HBM-bound
        achieved = alg_bytes / (gemm_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile("dense"),
                    "kernel": "gemm_bf16x2_reg_kernel<4, false, false, false>", "kernel_ms": gemm_ms,
                    "alg_bytes_per_launch": alg_bytes, "measured_read_peak": rd.value,
                    "frac_of_measured_read_peak": achieved / rd.value,
                    "measured_read_peak_random_bits": rdr.value,
                    "frac_of_measured_read_peak_random_bits": achieved / rdr.value}
    else:
        # the translation-invariant table (140 MB) stays in L2/MALL: the matrix pipe bounds it
        achieved = alg_flops / (gemm_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / BF16_PEAK_TFLOPS, "traffic": traffic_from_profile("table"),
                    "kernel": "gemm_bf16x2_tw_kernel", "kernel_ms": gemm_ms,
                    "alg_flops_per_launch": alg_flops,
                    "mfma_issued_frac": 3.0 * achieved / BF16_PEAK_TFLOPS,
                    "note": "split-BF16 issues 3 MFMA products per algorithmic flop (hi*hi + hi*lo + lo*hi)"}
    line = {
        "metric": "LL steps/s at N=32768, 64-state batch: split-BF16 MFMA vs FP32 plain, GPU",
        "value": steps_per_s,
        "unit": "LL batch-steps/s (one accepted RK45 step of all 64 states)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x2 demag product (FP32 accumulate), FP64 state/RHS/stepper",
        "data": "synthetic: seeded random unit magnetization, 64-field sweep -100..+100 mT",
        "config": {"workload": "256x64x2 film (2x2x1.5 nm), N=32768, 64-state field sweep, RKSUITE 'M' tol 1e-6",
                   "n_cells": n, "batch": B, "precision": "bf16x2", "tol": 1e-6,
                   "tensor": "translation-invariant table (140 MB)" if args.table
                   else "dense split-BF16 hi+lo fragments (25.8 GB)"},
        "state_steps_per_s": steps_per_s * B if steps_per_s else None,
        "rhs_per_s": passes / wall if wall else None,
        "rejected_attempts": int(rejected),
        "plain_fp32_steps_per_s": plain,
        "plain_fp32": plain_info,
        "fp32_mfma": mfma32,
        "speedup_vs_plain_fp32": (steps_per_s / plain) if (plain and steps_per_s) else None,
        "setup_s": t_setup,
        "field_error_vs_fp64": ferr,
        "tensor_assembly": asm,
        "table_path": table,
        "step_breakdown": breakdown,
        "plain_configs": cfgs,
        "row_shard": shard,
        "adjoint": adj,
        "next_rows": nxt,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
