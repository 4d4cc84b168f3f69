// Kernel descriptors and launchers shared by the host runtime.
#pragma once

#include <vector>

#include "mt_common.h"

namespace mt {

// ---- (a) tensor assembly ------------------------------------------------------
int launch_assemble(const double* src, const double* abc, int abc_stride, const double* rot, const double* pts,
                    int n_src, int n_pts, float* out32, double* out64, uint8_t* split, hipStream_t stream);
// Magnetostatic tiles: records of kTileRec doubles (abc, offset, R, Rinv, R*M,
// symmetry ops, symmetry flag, tile type 2 / 5, tetrahedron vertices); points
// Fortran-ordered (npt,3); record size kTileRec (mt_common.h).
int launch_tiles_tensor(const double* tiles, int nt, const double* pts, int npt, double* out, hipStream_t stream);
int launch_tiles_field(const double* tiles, int nt, const double* pts, int npt, const double* Nin, double* H,
                       hipStream_t stream);
int launch_split(const float* K, int n, uint8_t* split, hipStream_t stream);
// Translation-invariant split-BF16 fragment table of a uniform grid (nx % 32 == 0),
// toeplitz_tensor_bytes(nx, ny, nz) bytes.
// Unstructured prisms: pts / abc (n x 3) per cell, receiver averaging over
// nave[0] x nave[1] x nave[2] sub-points; FP32 and/or FP64 receiver-major out.
int launch_assemble_avg(const double* pts, const double* abc, int n, const int nave[3], float* out32, double* out64,
                        hipStream_t stream);
// Tetrahedral grid: evaluation points pts (n x 3), element j's nodes
// elem[4j..4j+3] (0-based) into nodes (x,y,z rows); receiver-major out.
int launch_assemble_tetra(const double* pts, const double* nodes, const int* elem, int n, float* out32,
                          double* out64, hipStream_t stream);
int launch_toeplitz_assemble(int nx, int ny, int nz, double Lx, double Ly, double Lz, uint8_t* out,
                             hipStream_t stream);

// ---- (b) demag product ------------------------------------------------------------
// out[s*out_stride + c*n + i] = sum_j K_c.(i,j) m_s,.(j)   (positive K m)
int launch_matvec_f32(const float* K, int n, const float* m, int64_t m_stride, float* out, int64_t out_stride,
                      int nstates, hipStream_t stream);
// FP64 form of the same product (FP64 tensor planes and states)
int launch_matvec_f64(const double* K, int n, const double* m, int64_t m_stride, double* out, int64_t out_stride,
                      int nstates, hipStream_t stream);
int gemm_ksplit_for(int n, int tiles_per_group);  // k-split of the batched GEMM (two-tile groups)
int gemm_ksplit_rows(int n, int nrows, int tiles_per_group);  // same for nrows receiver rows of n sources
int gemm_ksplit_table(int nx, int ny, int nz);  // k-split of the kernel launch_gemm_bf16x2_toeplitz picks
int gemm_ksplit_dense(int n, int nrows);  // k-split of the dense-tensor kernel launch_gemm_bf16x2{,_rows} picks
// out[ks*split_stride + col*col_stride + c*n + i] = k-split partial of K m
int launch_gemm_bf16x2(const uint8_t* Ksplit, const uint8_t* Mpack, int n, float* out, int64_t col_stride,
                       int64_t split_stride, int ncols, int ksplit, hipStream_t stream);
// Row block: Krows holds the fragment rows of nrows receivers (row tiles
// row0/32 ... of an n-source split tensor, row0 % 32 == 0); output rows 0..nrows-1
// (out[ks*split_stride + col*col_stride + c*nrows + r]).
int launch_gemm_bf16x2_rows(const uint8_t* Krows, const uint8_t* Mpack, int n, int nrows, float* out,
                            int64_t col_stride, int64_t split_stride, int ncols, int ksplit, hipStream_t stream);
// Same product with the tensor given as the translation-invariant fragment table of
// an nx x ny x nz uniform grid (n = nx ny nz, nx % 32 == 0).
int launch_gemm_bf16x2_toeplitz(const uint8_t* Ktab, int nx, int ny, int nz, const uint8_t* Mpack, float* out,
                                int64_t col_stride, int64_t split_stride, int ncols, int ksplit, hipStream_t stream);

// Threshold-sparse demag tensor: six CSR matrices (xx xy xz yy yz zz), rows =
// receivers, FP32 values, entries with |K| > threshold.
struct SparseDemag {
  int n = 0;
  float threshold = 0.0f;
  int64_t nnz[6] = {0, 0, 0, 0, 0, 0};
  int64_t* rowptr = nullptr;  // 6 x (n+1), per component
  int64_t* coff = nullptr;    // 6 component offsets into cols/vals (device)
  int* cols = nullptr;
  float* vals = nullptr;
  float* mT = nullptr;        // transposed staging for the many-state product
  float* hT = nullptr;
  int tcols = 0;
  std::vector<std::vector<int>> host_cols;    // components loaded from host CSR
  std::vector<std::vector<float>> host_vals;
  int build(const float* K32, int n, float thr, hipStream_t stream);
  int load_component(int c, int n, int nnz, const float* values, const int* colInds, const int* rowInds,
                     hipStream_t stream);
  // h[s*ld + a*n + i] = pref * (K m_s)_a(i), s < S <= 64, m and h device [S][3n] with row stride ld
  int apply(const float* m, int64_t ld, int S, float pref, float* h, hipStream_t stream);
  void release();
  ~SparseDemag() { release(); }
};
// FindThresholdFraction over the positive entries of the FP32 receiver-major tensor.
int threshold_for_fraction(const float* K32, int n, double fraction, float* thr, hipStream_t stream);

// ---- (c) LL right-hand side ------------------------------------------------------
struct LlgGrid {
  int n, nx, ny, nz;
  double idx2, idy2, idz2;   // 1/dx^2 etc.
  // loaded exchange operator (unstructured grids, 0-based CSR: row i holds
  // entries xs[i] .. xe[i]-1); null: the uniform-grid stencil
  const int* xs;
  const int* xe;
  const int* xc;
  const double* xv;
  const double* jfact;       // A0 / (mu0 Ms_i)
  const double* kfact;       // K0_i / (mu0 Ms_i)
  const float* ms32;         // Ms_i in FP32 (demag scaling)
  const double* ms64;        // Ms_i in FP64 (FP64 demag mode; may be null)
  const double* aniso;       // 6 per cell (u u^T), or null when K0 == 0 everywhere
};

struct LlgDesc {
  const double* m;           // argument, 3n SoA
  double* out;               // dm/dt, 3n SoA (may be null)
  const float* hpart;        // demag partial sums for this state's column
  int64_t split_stride;      // distance between k-split partials
  int nsplit;
  const double* hpart64;     // FP64 demag sums (3n, replaces hpart when set): H_m = -S * Ms (FP64)
  double hext[3];            // applied field (A/m) at the stage time
  double alpha, gamma;
  double* fields;            // optional 12n: H_exc, H_ext, H_dem, H_ani (negated, SoA)
};

int launch_llg(const LlgGrid& g, const LlgDesc* d_descs, int nstates, hipStream_t stream);

// ---- (d) vector operations -------------------------------------------------------
constexpr int kMaxTerms = 12;

enum RedKind {
  kRedNone = 0,
  kRedStepB = 1,
  kRedFinal = 2,
  kRedStepA = 3,
  kRedYpNorm = 4,
  kRedWeightsAvg = 5,
  kRedWdot = 6,
  kRedMaxDiff = 7,
};

struct VecDesc {
  // combination: dst = post_scale * (bconst + bcoef*base + sum_k coef[k]*src[k])   (skipped when dst is null)
  double* dst;
  const double* base;
  double bcoef;
  double bconst;
  double post_scale;
  int nterm;
  int pack_col;              // >= 0: also pack dst into the demag input column
  const double* src[kMaxTerms];
  double coef[kMaxTerms];
  // reduction
  int rkind;
  int rn;
  const double* ra;
  const double* rb;
  const double* rc;
  double* rd;                // kRedFinal: error-estimate output; kRedStepA: yp (read)
  double* rw;                // weights (written, or read by kRedWdot)
  const double* rsrc[8];
  double rcoef[8];
  double rscale;
  double thresh;
};

enum PackMode { kPackNone = 0, kPackF32 = 1, kPackBf16 = 2 };

struct PackSpec {
  int mode;
  float* f32;                // [col][3n]
  int64_t col_stride;
  uint8_t* bf16;             // fragment layout
};

int vec_blocks_per_state(int n3);
int launch_vec(const VecDesc* d_descs, int nstates, int n, double* partials, double* results, const PackSpec& pack,
               bool reduce, hipStream_t stream);

}  // namespace mt
