/*
 * magtense_hip.h - C ABI of the MI355X (gfx950) micromagnetic hot path.
 *
 * Plain C types only (no HIP or torch types in the signatures).  Every
 * function returns 0 on success or a negative status; mt_last_error() returns
 * the message of the last failure on the calling thread.  "device pointer"
 * arguments are HIP device addresses (for example torch.Tensor.data_ptr() of a
 * cuda tensor); "host pointer" arguments are ordinary host memory.
 *
 * Reference interfaces replaced (paths relative to the MagTense tree):
 *   - f2py entry RunMicroMagSimulation
 *       python/src/magtense/lib/FortranToPythonIO.f90:429-489
 *     -> mt_run_micromag / mt_solver_* (batched: n_states independent runs)
 *   - CUDA plug-in C ABI
 *       source/MagTenseFortranCuda/cuda/MagTenseCudaBlasICLWrapper.cxx:13-59
 *       (bound from Fortran in source/MagTenseMicroMag/FortranCuda.f90:15-60)
 *     -> icl_initDemagMatrices, icl_MVMult_GetH, icl_destroy (same signatures)
 *   - prism tensor getNFromTiles / ComputeDemagfieldTensor
 *       python/src/magtense/lib/FortranToPythonIO.f90:25-93,
 *       source/MagTenseMicroMag/LandauLifshitzEquationSolver.f90:883-1164
 *     -> mt_prism_tensor (device arrays), mt_tiles_demag_tensor (host arrays)
 *   - getHFromTiles  FortranToPythonIO.f90:106-223
 *     -> mt_tiles_h_field
 *   - demag product updateDemagfield / cu_MVMult_GetH
 *       LandauLifshitzEquationSolver.f90:558-779, MagTenseCudaBlas.cu:367-407
 *     -> mt_demag_apply_f32, mt_demag_apply_bf16x2
 *   - right-hand side dmdt_fct  LandauLifshitzEquationSolver.f90:220-273
 *     -> mt_ll_rhs
 */
#ifndef MAGTENSE_HIP_H
#define MAGTENSE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MT_PREC_FP32 0   /* plain per-state FP32 mat-vec (reference semantics) */
#define MT_PREC_BF16X2 1 /* split BF16 hi+lo tensor, batched MFMA GEMM, FP32 accumulation */
#define MT_PREC_FP64 2   /* FP64 tensor, product and Ms scaling (mt_rk4 only: gradient checks) */
#define MT_PREC_FP32_MFMA 3 /* FP32 tensor as MFMA fragments, batched FP32 MFMA GEMM (v_mfma_f32_32x32x2f32),
                               FP32 accumulation: the reference's arithmetic class, one tensor pass per 64 states */

/* status codes (every int-returning entry point; MT_ERR_UNSUPPORTED below) */
#define MT_OK 0
#define MT_ERR_ARG (-1)    /* invalid argument, rejected before any launch */
#define MT_ERR_HIP (-2)    /* HIP runtime failure */
#define MT_ERR_SOLVER (-3) /* RKSUITE fatal condition */
#define MT_ERR_ALLOC (-4)  /* device allocation failure */
#define MT_ERR_STATE (-5)  /* entry points called in the wrong order */

/* demag approximations (dem_appr of MicromagProblem, micromag.py:61-68) */
#define MT_DEMAG_DENSE 1
#define MT_DEMAG_THRESHOLD 2
#define MT_DEMAG_THRESHOLD_FRACTION 4

/* ---- library --------------------------------------------------------------- */
const char* mt_last_error(void);
int mt_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int mt_device_count(void);

/* Measures the device's streaming bandwidth on a `bytes` buffer (grid-stride
 * 16-byte loads; copy counts bytes read + written), averaged over `iters`
 * launches.  No reference counterpart: it gives the measured ceiling beside the
 * published 8 TB/s and the known-size launches that calibrate HBM counters.
 * Either output may be null. */
int mt_measure_bandwidth(long long bytes, int iters, double* read_gbps, double* copy_gbps);
/* Same probe over a buffer of constant bytes (pattern 0, what
 * mt_measure_bandwidth reads) or of pseudo-random bits (pattern 1: HBM data
 * lines toggle as with live data, which costs power and hence clock). */
int mt_measure_bandwidth_pattern(long long bytes, int iters, int pattern, double* read_gbps, double* copy_gbps);

/* ---- (a) prism tensor ----------------------------------------------------------
 * K = -N_standard (MagTense sign) of prism j (centre src[3j..], full edges
 * abc[abc_stride*j ..]) at point i (pts[3i..]), evaluated at rot_j (p_i - c_j)
 * when rot (9 doubles per prism, row-major rotMat of getRotationMatrices,
 * DemagFieldGetSolution.f90:925-948) is given, else at p_i - c_j.  All arrays
 * are device pointers, row-major.  Outputs (any may be null):
 *   out32[c*n_pts*n_src + i*n_src + j]  FP32, c = xx,xy,xz,yy,yz,zz
 *   out64                               same layout, FP64
 *   split                               fragment-tiled BF16 hi/lo,
 *                                       round_up(n_pts,32) * round_up(n_src,32) * 24
 *                                       bytes (mt_split_tensor_bytes(n) when square;
 *                                       with pts = receivers row0.. of the sources,
 *                                       the row block of mt_demag_apply_bf16x2_rows)
 * Replaces getN_prism_3D over all pairs (TileNComponents.f90:629-724 and
 * LandauLifshitzEquationSolver.f90:935-984 / 1044-1133). */
int mt_prism_tensor(const double* src, const double* abc, int abc_stride, const double* rot, const double* pts,
                    int n_src, int n_pts, float* out32, double* out64, void* split);

/* ---- magnetostatic prism tiles (host arrays) --------------------------------
 * Mirrors the f2py arguments of getNFromTiles / getHFromTiles
 * (FortranToPythonIO.f90:25-93,106-223) that matter for rectangular prisms.
 * Every (n,3) array is Fortran-ordered as f2py passes it: element (i,k) at
 * i + n*k.  Optional members may be null (rot: no rotation; tile_type: all
 * prisms; use_sym: 0; sym_ops: 1; M: only needed for the field).  A tile
 * whose tile_type is not 2 (prism) fails with MT_ERR_UNSUPPORTED. */
typedef struct {
  int n_tiles;
  const double* size;    /* tile_size (a,b,c), full edges */
  const double* offset;  /* tile centre */
  const double* rot;     /* rotAngles, getRotationMatrices (DemagFieldGetSolution.f90:925-948) */
  const double* M;       /* Mag */
  const int* tile_type;  /* tileType */
  const int* use_sym;    /* exploitSymmetry (1 = eight mirrored evaluations, :744-822) */
  const double* sym_ops; /* symmetryOps */
  const double* vertices; /* (n_tiles,3,4) Fortran order, the vertices of tetrahedron tiles (tile_type 5;
                             getN_tensor_tetrahedron, TileNComponents.f90:944-961); may be null without them */
} mt_tiles;

#define MT_ERR_UNSUPPORTED (-6)

/* N(n_tiles, n_pts, 3, 3), Fortran order ((i,p,a,b) at i + nt*(p + np*(a + 3b))),
 * of every tile at every point (pts: (n_pts,3) Fortran order), in each tile's
 * rotated frame, MagTense sign (getFieldFromTile, DemagFieldGetSolution.f90:685-730). */
int mt_tiles_demag_tensor(const mt_tiles* tiles, int n_pts, const double* pts, double* N);

/* H(n_pts,3), Fortran order: sum over tiles of Rinv (N (R M)) (getHFromTiles).
 * When N is non-null it is the stored tensor of mt_tiles_demag_tensor
 * (useStoredN); the sum over tiles is a fixed-order tree, not the reference's
 * sequential loop (differences at FP64 rounding level). */
int mt_tiles_h_field(const mt_tiles* tiles, int n_pts, const double* pts, const double* N, double* H);
long long mt_split_tensor_bytes(int n);
long long mt_packed_m_bytes(int n);
/* FP32 receiver-major tensor (device) -> split BF16 fragment layout (device). */
int mt_split_tensor(const float* K32, int n, void* split);
/* FP32 receiver-major tensor (device) -> the same fragment order with FP32
 * values for v_mfma_f32_32x32x2f32 (mt_split_tensor_bytes(n) bytes, device).
 * No reference counterpart (an additive batched form of the FP32 product). */
int mt_f32_mfma_tensor(const float* K32, int n, void* frags);

/* Translation-invariant split-BF16 table of a uniform nx x ny x nz grid (nx % 32
 * == 0; cell sizes and centres as setupGrid, LandauLifshitzEquationSolver.f90:
 * 815-872): every tensor entry depends only on the cell offset, so the fragments
 * of all (row tile, k-chunk) pairs with the same offset are stored once
 * (mt_toeplitz_tensor_bytes bytes, device) instead of the dense N x N tensor.
 * No reference counterpart: a storage form of the dense tensor of
 * ComputeDemagfieldTensor (:883-986) for the batched product. */
long long mt_toeplitz_tensor_bytes(int nx, int ny, int nz);
int mt_toeplitz_tensor(int nx, int ny, int nz, double Lx, double Ly, double Lz, void* table);

/* ---- (b) demag product -----------------------------------------------------------
 * h[s*3n + c*n + i] = pref * sum_j K_c(i,j) m[s*3n + c'*n + j] for each of
 * nstates states (SoA per state, device pointers).  pref = -1 reproduces
 * the reference's alpha = -1 (MagTenseCudaBlas.cu:367-407). */
int mt_demag_apply_f32(const float* K32, int n, const float* m, float* h, int nstates, float pref);
/* Same product from the split tensor: packs m (FP32, device) into BF16 hi/lo
 * fragments internally; nstates <= 64 per call (larger batches are chunked). */
int mt_demag_apply_bf16x2(const void* split, int n, const float* m, float* h, int nstates, float pref);
/* Same product from the FP32 fragments of mt_f32_mfma_tensor on FP32 MFMAs
 * (FP32 products and sums: the arithmetic class of the reference's sgemv,
 * LandauLifshitzEquationSolver.f90:706-754, one tensor pass per 64 states). */
int mt_demag_apply_f32_mfma(const void* frags, int n, const float* m, float* h, int nstates, float pref);
/* Row block of the same product, the per-GPU product of a tensor sharded by
 * receiver rows: h[s*3*nrows + c*nrows + r] = pref * sum_j K_c(row0 + r, j) m[...j]
 * for r < nrows, all n sources.  split_rows points at the fragment rows of
 * receivers row0 ... row0 + nrows - 1 (row0 % 32 == 0, row0 + nrows <= n), which
 * are contiguous in the split tensor from byte mt_split_tensor_row_offset(n, row0)
 * on (-1 for an invalid row0) and span round_up(nrows, 32) * round_up(n, 32) * 24
 * bytes; a GPU holding only its shard passes its own copy of those bytes.
 * No reference counterpart: the row split of updateDemagfield's products
 * (LandauLifshitzEquationSolver.f90:706-754). */
long long mt_split_tensor_row_offset(int n, int row0);
int mt_demag_apply_bf16x2_rows(const void* split_rows, int n, int row0, int nrows, const float* m, float* h,
                               int nstates, float pref);
/* Near-band correction of a split-BF16 product on a uniform nx x ny x nz grid
 * of an Lx x Ly x Lz box: h += pref * (the FP32 product minus the three split
 * products) over the sources within `radius` cells (1..4) in x and y and
 * min(radius, nz - 1) layers in z of every receiver, m and h device [s][3n]
 * FP32.  Applied after mt_demag_apply_bf16x2 / _toeplitz it brings the field
 * of coherent magnetizations (uniform, smooth) to the FP32 product's level in
 * the near field, where the split's rounding residuals are largest; the
 * solver applies it with radius 2 on uniform grids (MT_NEAR_RADIUS overrides,
 * 0 switches it off).  No reference counterpart (the reference's product is
 * FP32, LandauLifshitzEquationSolver.f90:706-754). */
int mt_demag_near_correction(int nx, int ny, int nz, double Lx, double Ly, double Lz, int radius, const float* m,
                             float* h, int nstates, float pref);
/* Allocation-free form of mt_demag_apply_bf16x2 (same kernels, same result).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace (packed states and k-split partial sums,
 * mt_demag_apply_bf16x2_noalloc_bytes(n, nstates) bytes) lives in a 64 MiB
 * file-scope device array that every device holds for the life of the
 * process (mt_noalloc_scratch_bytes()); larger problems (about n > 43000 cells)
 * are refused with MT_ERR_ARG before any launch.  Calls that overlap on one
 * device share the array and race. */
long long mt_noalloc_scratch_bytes(void);
long long mt_demag_apply_bf16x2_noalloc_bytes(int n, int nstates);
int mt_demag_apply_bf16x2_noalloc(const void* split, int n, const float* m, float* h, int nstates, float pref);
/* Same product from the translation-invariant table of mt_toeplitz_tensor. */
int mt_demag_apply_bf16x2_toeplitz(const void* table, int nx, int ny, int nz, const float* m, float* h, int nstates,
                                   float pref);

/* Threshold-sparse tensor (ApplyThresholdDense, DemagAuxFunctions.f90:20-122;
 * ConvertDenseToSparse_s, util_micromag.f90:81-138): six CSR matrices built on
 * the device from the FP32 receiver-major tensor K32 (device), keeping entries
 * with |K| > threshold.  approx MT_DEMAG_THRESHOLD uses `threshold` directly;
 * MT_DEMAG_THRESHOLD_FRACTION treats it as the fraction of positive entries to
 * drop (FindThresholdFraction, DemagAuxFunctions.f90:353-428).  K32 may be
 * freed afterwards. */
int mt_demag_sparse_create(const float* K32, int n, int approx, double threshold, void** out);
/* nnz per component (xx xy xz yy yz zz) and the FP32 threshold used. */
int mt_demag_sparse_info(void* sp, long long* nnz6, float* threshold);
/* Same kept entries, held as a block-sparse split-BF16 tensor: the 32-row x
 * 16-source tiles holding a kept entry, stored in the MFMA fragment order of
 * the dense split tensor (entries at or below the threshold zero), with row
 * tiles grouped four to a workgroup over a shared list of source chunks
 * (grid_nx > 0, a multiple of 32: the cells form rows of grid_nx, and tiles of
 * consecutive rows may share a group).  The product is the batched split-BF16
 * MFMA GEMM over the stored tiles.  No reference counterpart (an additive
 * precision mode of the sparse branch, LandauLifshitzEquationSolver.f90:581-624). */
int mt_demag_sparse_create_bf16x2(const float* K32, int n, int approx, double threshold, int grid_nx, void** out);
/* The same block-sparse tiles with FP32 values (the reference's FP32
 * arithmetic): the product runs v_mfma_f32_32x32x2f32 over the stored tiles
 * with FP32 accumulation, for many states at once (the solver takes this form
 * for the threshold modes with precision fp32 and five or more states).  No
 * reference counterpart (the reference multiplies one state at a time by the
 * CSR matrix, MagTenseCudaBlas.cu:106-297, LandauLifshitzEquationSolver.f90:581-624). */
int mt_demag_sparse_create_f32_tiles(const float* K32, int n, int approx, double threshold, int grid_nx, void** out);
/* Block statistics of a tensor from mt_demag_sparse_create_bf16x2 or
 * mt_demag_sparse_create_f32_tiles: list
 * entries (each 4 x 12 KiB of fragments), non-empty tiles, fragment bytes, and
 * the partial-sum slots of its balanced workgroup schedule (ksplit). */
int mt_demag_sparse_blocks(void* sp, long long* entries, long long* tiles_nonempty, long long* bytes, int* ksplit);
/* h = pref * K_sparse m for nstates states, [s][3n] device layout (the
 * product of updateDemagfield's sparse branch, LandauLifshitzEquationSolver.f90:581-624;
 * FP32 CSR, or the tiles of mt_demag_sparse_create_bf16x2 /
 * mt_demag_sparse_create_f32_tiles, whose workspace the handle allocates on
 * first use and keeps: calls on one handle must not overlap). */
int mt_demag_apply_sparse(void* sp, const float* m, float* h, int nstates, float pref);
void mt_demag_sparse_destroy(void* sp);

/* ---- (c) right-hand side ------------------------------------------------------------
 * One dm/dt evaluation per state on a uniform grid (device pointers):
 *   m, dmdt: nstates x 3n doubles (SoA per state); ms, k0: n doubles (host);
 *   u_ea: n*3 doubles (host, may be null); hext: nstates*3 doubles (host, A/m);
 *   K32: FP32 tensor (device).  fields (device, may be null): nstates x 12n
 *   doubles = H_exc, H_ext, H_dem, H_ani (negated, SoA). */
int mt_ll_rhs(int nx, int ny, int nz, double Lx, double Ly, double Lz, const double* ms, const double* k0,
              const double* u_ea, double A0, double gamma, double alpha, const float* K32, const double* m,
              const double* hext, int nstates, double* dmdt, double* fields);

/* ---- (d) batched solver ------------------------------------------------------------
 * Mirrors RunMicroMagSimulation for a uniform grid, for n_states independent
 * magnetization states sharing the grid, materials and tensor.  All pointers
 * are host pointers. */
typedef struct {
  int nx, ny, nz;
  double Lx, Ly, Lz;
  int n_states;
  const double* Ms;      /* n cells */
  const double* K0;      /* n cells, may be null (0) */
  const double* u_ea;    /* n*3 row-major, may be null */
  double A0, gamma, alpha0;
  int nt_alpha;
  const double* alphat;  /* nt_alpha*2 row-major [t, alpha] (used when alpha0 == 0) */
  int solver;            /* 1 explicit (field rows in sequence), 2 dynamic */
  int nt_hext;
  const double* hext;    /* nt_hext*4 row-major [t, hx, hy, hz] (A/m); per state if hext_per_state */
  int hext_per_state;
  int nt;
  const double* t;       /* output times, t[0] = start */
  const double* m0;      /* 3n SoA (per state if m0_per_state) */
  int m0_per_state;
  double tol, thres;
  int nt_conv;
  const double* t_conv;
  double conv_tol;
  int precision;         /* MT_PREC_FP32, MT_PREC_BF16X2 or MT_PREC_FP32_MFMA */
  int return_fields;     /* compute H_exc/H_ext/H_dem/H_ani at every output time */
  long long max_steps;   /* > 0: pause every state after this many accepted steps */
  const float* K_in;     /* optional host FP32 tensor, 6*n*n receiver-major (skips assembly) */
  float* K_out;          /* optional host buffer receiving the FP32 tensor */
  int keep_fp32_tensor;  /* keep an FP32 copy beside the split tensor */
  int ksplit;            /* 0: automatic k-split of the MFMA GEMM */
  int demag_approx;      /* MT_DEMAG_DENSE (0 or 1), MT_DEMAG_THRESHOLD (2), MT_DEMAG_THRESHOLD_FRACTION (4);
                            the sparse modes keep an FP32 CSR (MT_PREC_FP32) or split-BF16 tiles (MT_PREC_BF16X2) */
  double demag_threshold; /* dem_thres: |K| threshold (mode 2) or fraction of the positive entries (mode 4) */
  int toeplitz;          /* MT_PREC_BF16X2 tensor storage: 0 automatic (the translation-invariant fragment
                            table when nx % 32 == 0 and no K_in), 1 require it, -1 dense fragments */
  /* grid of unstructured prisms (grid_type 3, gridTypeUnstructuredPrisms; loadMicroMagProblem,
     MagTenseMicroMagPyIO.f90:76-85,163-180): n = nx*ny*nz cells at grid_pts with full edges
     grid_abc; the tensor is averaged over n_ave sub-points of each receiver cell
     (LandauLifshitzEquationSolver.f90:1044-1133) and the exchange operator is the loaded CSR
     (ConvertExchangeTerm3D_NonUniform, :1530-1544).  grid_type 0 or 1: the uniform grid. */
  int grid_type;
  const double* grid_pts;      /* n*3 row-major */
  const double* grid_abc;      /* n*3 row-major */
  int n_ave[3];
  int exch_nnz;
  const int* exch_rows_start;  /* n, 1-based (mkl_sparse_d_create_csr rows_start) */
  const int* exch_rows_end;    /* n, 1-based */
  const int* exch_cols;        /* exch_nnz, 1-based */
  const double* exch_vals;     /* exch_nnz (double; the reference's Python path passes int32, Q7) */
  /* tetrahedral grid (grid_type 2, gridTypeTetrahedron; LandauLifshitzEquationSolver.f90:988-1041):
     element i (evaluation point grid_pts[i]) has the 1-based nodes grid_elements[4i .. 4i+3];
     n_ave is not used; the exchange operator is the loaded CSR as for grid_type 3 */
  int grid_nnod;
  const double* grid_nodes;    /* grid_nnod*3 row-major */
  const int* grid_elements;    /* 4n, 1-based */
} mt_micromag_problem;

typedef struct {
  long long f_calls, steps, bad_steps, flag3, flag4;
  double h_last, t_reached;
  long long evals, ticks, demag_passes;
  /* mt_solver_set_timing on: event-timed milliseconds since it was switched on
   * of the demag products, of the stepper's vector operations (descriptor
   * uploads included) and of the LL right-hand-side kernel, their launch counts
   * and the algorithmic HBM bytes of the last two (every vector a launch reads
   * or writes counted once) */
  double demag_ms;
  double vec_ms, llg_ms, vec_bytes, llg_bytes;
  long long vec_launches, llg_launches;
  /* of those demag milliseconds, the near-band correction's (uniform grids, bf16x2),
   * which includes the LL right-hand side that kernel evaluates (llg_ms then 0) */
  double near_ms;
} mt_stats;

typedef struct mt_solver mt_solver;

int mt_solver_create(const mt_micromag_problem* problem, mt_solver** out);
/* Advance every state by `steps` accepted steps (0: run to the end). */
int mt_solver_advance(mt_solver* s, long long steps);
int mt_solver_run(mt_solver* s);
/* Advance until the states' accepted steps, summed over the batch, grew by
 * total_steps (states keep independent step sizes; the batch stays full). */
int mt_solver_advance_total(mt_solver* s, long long total_steps);
long long mt_solver_steps_total(mt_solver* s);
/* t_out: n_states*n_h*nt; y_out: n_states*n_h*nt*3n; fields: 4x that (H_exc,
 * H_ext, H_dem, H_ani interleaved per output as 12n); stats: n_states entries.
 * n_h = nt_hext for the explicit solver, 1 otherwise.  Any pointer may be null. */
int mt_solver_results(mt_solver* s, double* t_out, double* y_out, double* fields, mt_stats* stats);
int mt_solver_state(mt_solver* s, double* y);
int mt_solver_set_timing(mt_solver* s, int on);
/* Sparse demag of a solver created with demag_approx 2 or 4: nnz per component
 * and the FP32 threshold used (an argument error when the solver is dense). */
int mt_solver_demag_info(mt_solver* s, long long* nnz6, float* threshold);
/* The form the solver holds its demag tensor in: 0 dense (FP32 or split-BF16,
 * or the translation-invariant table), 1 threshold CSR, 2 threshold tiles in
 * split-BF16, 3 threshold tiles in FP32 (precision fp32 with five or more
 * states).  No reference counterpart (the reference's sparse path is CSR). */
int mt_solver_demag_form(mt_solver* s, int* form);
void mt_solver_destroy(mt_solver* s);
int mt_run_micromag(const mt_micromag_problem* problem, double* t_out, double* y_out, double* fields,
                    mt_stats* stats);

/* ---- differentiable integration (no reference counterpart) -------------------------
 * Fixed-step RK4 of the same LL right-hand side for n_states states (constant
 * applied field per state: the first row of hext; alpha0, gamma, Ms, K0, A0,
 * u_ea and the tensor as in mt_micromag_problem; demag product in FP32 as the
 * reference (precision MT_PREC_FP32) or fully in FP64 (MT_PREC_FP64: FP64 tensor,
 * product and Ms scaling, for gradient checks against finite differences).
 * mt_rk4_forward stores checkpoints every checkpoint_every steps;
 * mt_rk4_backward runs the hand-written discrete adjoint (recomputing each
 * checkpoint segment) for the cotangent dL/dm(T) and returns dL/dm0
 * (n_states*3n), dL/dH_ext (n_states*3, A/m) and dL/dMs (n_states*n).  Host pointers. */
typedef struct mt_rk4 mt_rk4;
int mt_rk4_create(const mt_micromag_problem* problem, double dt, long long nsteps, int checkpoint_every,
                  mt_rk4** out);
int mt_rk4_forward(mt_rk4* h, const double* m0, double* mT);
int mt_rk4_backward(mt_rk4* h, const double* dL_dmT, double* grad_m0, double* grad_hext, double* grad_ms);
void mt_rk4_destroy(mt_rk4* h);

/* ---- reference plug-in ABI (MagTenseCudaBlasICLWrapper.cxx:13-59) -------------------
 * Host arrays, column-major N x N FP32 blocks, scalars by reference, one
 * global instance - exactly what FortranCuda.f90 binds. */
void icl_initDemagMatrices(const float* Kxx, const float* Kxy, const float* Kxz, const float* Kyy, const float* Kyz,
                           const float* Kzz, int* n);
void icl_MVMult_GetH(const float* Mx, const float* My, const float* Mz, float* Hx, float* Hy, float* Hz, int* n,
                     float* pref);
void icl_destroy(void);
/* Sparse plug-in entry points (MagTenseCudaBlasICLWrapper.cxx:13-37,
 * MagTenseCudaBlas.cu:106-297): component mat_no (1..6 = xx xy xz yy yz zz)
 * as a CSR of n rows with nnz entries, 1-based columns and row starts
 * (rowInds has n+1 entries, the last equal to nnz + rowInds[0]); then
 * H = pref * K M over all six components, host arrays. */
void icl_initDemagMatrices_sparse(int* n, int* nnz, int* mat_no, float* values, int* colInds, int* rowInds);
void icl_MVMult_GetH_sparse(const float* Mx, const float* My, const float* Mz, float* Hx, float* Hy, float* Hz,
                            const int* n, const float* pref);

#ifdef __cplusplus
}
#endif

#endif /* MAGTENSE_HIP_H */
