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
// FP32 planes -> the dense tensor as FP32 MFMA fragments (split_tensor_bytes(n) bytes)
int launch_f32_fragments(const float* K, int n, uint8_t* frags, hipStream_t stream);
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
// Block-sparse split-BF16 tensor (threshold modes with precision bf16x2): only
// the 32-row x 16-source tiles holding a kept entry are stored, in the dense
// fragment order.  Row tiles are grouped four to a workgroup; a group's tiles
// share one sorted list of k-chunks (the union of their non-empty chunks);
// wave w's fragments of list entry l are fragment set 4 goff[g] + w cnt_g + l
// (12 fragments of 1 KiB), cnt_g = goff[g+1] - goff[g].
struct BlockSparseView {
  int ngroups = 0;
  const int* tiles = nullptr;     // 4 per group: row tile index, -1 none (the first is always a tile)
  const int64_t* goff = nullptr;  // ngroups + 1 list offsets
  const int* chunks = nullptr;    // k-chunk of each list entry
  // balanced schedule: workgroup w runs segments segs[wseg[w] .. wseg[w+1]),
  // each (group, partial-sum slot, first list entry, entries) of one group's list
  int nwg = 0, seg_max = 0;  // seg_max: the longest segment (its chunk indices are staged in LDS)
  const int* wseg = nullptr;
  const int4* segs = nullptr;
};
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
// Same product over a block-sparse tensor (fragment sets `frags`, structure
// `bs`): split-BF16 fragments and states, or with f32 FP32 fragments and
// states (the FP32 layouts of pack_f32_frag) on FP32 MFMAs.
int launch_gemm_bsr(const uint8_t* frags, const BlockSparseView& bs, bool f32, const uint8_t* Mpack, int n, float* out,
                    int64_t col_stride, int64_t split_stride, int ncols, int ksplit, hipStream_t stream);
// The dense product over FP32 fragments (launch_f32_fragments) with FP32
// state fragments, on FP32 MFMAs: partial sums as launch_gemm_bf16x2.
int launch_gemm_f32(const uint8_t* Kf, const uint8_t* Mpack, int n, float* out, int64_t col_stride,
                    int64_t split_stride, int ncols, int ksplit, hipStream_t stream);

struct BlockSparse {
  int n = 0, ngroups = 0;
  int ksplit = 1;                 // partial-sum slots the balanced schedule writes
  int nwg = 0;                    // workgroups of the schedule
  int seg_max = 0;                // longest segment (list entries)
  int max_list = 0;               // longest group list (entries)
  float threshold = 0.0f;
  bool f32 = false;               // FP32 fragments (FP32 MFMAs) instead of split-BF16 pairs
  int64_t entries = 0;            // list entries over all groups (each 4 x 12 KiB of fragments)
  int64_t tiles_nonempty = 0;     // (row tile, chunk) tiles holding a kept entry
  int64_t nnz[6] = {0, 0, 0, 0, 0, 0};  // kept entries per component
  int* tiles = nullptr;
  int64_t* goff = nullptr;
  int* chunks = nullptr;
  int* egroup = nullptr;          // group of each list entry (build only)
  int* wseg = nullptr;
  int4* segs = nullptr;
  uint8_t* frags = nullptr;
  // grid_nx > 0 with grid_nx % 32 == 0: uniform-grid rows of grid_nx cells, so
  // row tiles of consecutive grid rows can share a group
  int build(const float* K32, int n, float thr, int grid_nx, hipStream_t stream, bool fp32 = false);
  int schedule(const std::vector<int64_t>& goff_host, int cus);
  BlockSparseView view() const { return BlockSparseView{ngroups, tiles, goff, chunks, nwg, seg_max, wseg, segs}; }
  int64_t bytes() const { return entries * 4 * 12 * 1024; }
  void release();
  ~BlockSparse() { release(); }
};

// Threshold-sparse demag tensor: six CSR matrices (xx xy xz yy yz zz), rows =
// receivers, FP32 values, entries with |K| > threshold.
// Workspace of one batched split-BF16 product (api.hip): the packed states
// (zeroed before every 64-state group: padded sources must read as zero) and
// the k-split partial sums.
struct Bf16x2Work {
  int ks = 1, cols = 1;
  size_t packed_bytes = 0, part_floats = 0;
  size_t bytes() const { return (size_t)round_up((int64_t)packed_bytes, 256) + part_floats * sizeof(float); }
};
// ks > 0 overrides the k-split of the dense / table kernels (block-sparse tensors carry their own)
Bf16x2Work bf16x2_work(int n, int nrows, int nx, int ny, int nz, int nstates, int ks = 0);
// The batched product h = pref * (K m) for nstates states in groups of 64 on
// the null stream, with the caller's workspace `ws` of w.bytes(); synchronous.
// nx > 0: K is the translation-invariant table of an nx x ny x nz grid; nrows <
// n: K is a row block of the dense tensor; bsv: K holds the fragments of a
// block-sparse tensor (f32: its FP32 form, states packed as FP32 fragments).
int apply_bf16x2_in(const void* K, int n, int nrows, int nx, int ny, int nz, const float* m, float* h, int nstates,
                    float pref, const Bf16x2Work& w, uint8_t* ws, const BlockSparseView* bsv = nullptr,
                    bool f32 = false);
int device_ok();

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
  BlockSparse* blocks = nullptr;  // split-BF16 block-sparse form (precision bf16x2) instead of the CSR
  uint8_t* ws = nullptr;          // the block-sparse product's workspace, kept between calls
  size_t ws_bytes = 0;
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

// Small batches (up to kInlineStates states) pass their descriptors by value
// in the launch arguments: no host -> device copy, and so no copy kernel, in
// front of the launch (a tick of a small problem is a few kernels of a few
// microseconds each).
constexpr int kInlineStates = 4;
struct LlgDescsInline {
  LlgDesc d[kInlineStates];
};
int launch_llg_inline(const LlgGrid& g, const LlgDesc* h_descs, int nstates, hipStream_t stream);
// Plain FP32 demag rows and the LL right-hand side in one launch (the FP32
// dense path): the rows of matvec_f32 for each state, then the LL kernel's
// arithmetic for the block's cells with those sums (bit-identical to
// launch_matvec_f32 + launch_llg with one partial-sum slot).  Descriptors from
// d_descs, or by value from h_descs when d_descs is null (nstates <=
// kInlineStates); m32 + s * m_stride is state s's FP32 demag input.
int launch_demag_llg_f32(const float* K, const LlgGrid& g, const LlgDesc* d_descs, const LlgDesc* h_descs,
                         const float* m32, int64_t m_stride, int nstates, hipStream_t stream);

// Near-band correction of the split-BF16 product on a uniform nx x ny x nz
// grid: the FP32 product's difference from the three split products, summed
// over the sources within R cells in x and y and min(R, nz - 1) layers in z.
// With K32 = khi + klo + dK and m32 = mhi + mlo + dm (BF16 hi and lo, FP32
// rest) that difference is khi dm + klo (mlo + dm) + dK m32.  Every term is
// about 2^-17 of K32 m32, so BF16 operands carry it to about 2^-26 of the
// product: the kernel forms khi bf(dm) + klo mlo + bf(dK) mhi with BF16 pair
// dot products (FP32 sums).
//
// Table: per offset kNearWords words of BF16 pairs (first element in the low
// half), for output component a = x, y, z at 5a: (khi, klo) of K_a,x, K_a,y,
// K_a,z, then (dK_a,x, dK_a,y), (dK_a,z, 0); word 15 is zero.  Offsets o =
// (ox + R) + (oy + R) w + (oz + Rz) w^2, w = 2R + 1, o = receiver - source.
constexpr int kNearWords = 16;
int near_band_rz(int R, int nz);
int64_t near_table_words(int R, int nz);

// round to the nearest BF16, ties to even (finite inputs)
__host__ __device__ inline uint32_t bf16_rne_bits(float v) {
  uint32_t u;
  __builtin_memcpy(&u, &v, 4);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__host__ __device__ inline float bf16_bits_value(uint32_t b) {
  const uint32_t u = b << 16;
  float v;
  __builtin_memcpy(&v, &u, 4);
  return v;
}
// One offset's table words from its six FP32 tensor entries (xx xy xz yy yz zz).
__host__ __device__ inline void near_pack_offset(const float k32[6], uint32_t* w) {
  constexpr int cab[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  uint32_t hl[6], dk[6];
  for (int c = 0; c < 6; ++c) {
    const uint32_t hi = bf16_rne_bits(k32[c]);
    const float khi = bf16_bits_value(hi);
    const uint32_t lo = bf16_rne_bits(k32[c] - khi);
    const float klo = bf16_bits_value(lo);
    hl[c] = hi | lo << 16;
    dk[c] = bf16_rne_bits((k32[c] - khi) - klo);
  }
  for (int a = 0; a < 3; ++a) {
    for (int b = 0; b < 3; ++b) w[5 * a + b] = hl[cab[a][b]];
    w[5 * a + 3] = dk[cab[a][0]] | dk[cab[a][1]] << 16;
    w[5 * a + 4] = dk[cab[a][2]];
  }
  w[15] = 0;
}
// thr: entries with |K32| <= thr are zero (the threshold-sparse modes); -1 keeps all
int launch_near_table(int nx, int ny, int nz, double Lx, double Ly, double Lz, int R, int Rz, float thr,
                      uint32_t* table, hipStream_t stream);
// out[s*col_stride + a*n + i] (+)= pref * correction of state s; the state is
// read from descs[s].m (FP64) when descs is set, else from m32 + s*m_stride.
// llg (with descs, pref 1, no accumulation): the correction is not stored;
// each cell's LL right-hand side is evaluated with it instead (the LL kernel's
// arithmetic, the GEMM slots of descs[s] and then the correction, in the order
// the LL kernel sums the slots), replacing launch_llg after the product.
int launch_near_correction(const uint32_t* table, int R, int nx, int ny, int nz, const LlgDesc* descs,
                           const float* m32, int64_t m_stride, int nstates, float* out, int64_t col_stride, float pref,
                           bool accumulate, hipStream_t stream, const LlgGrid* llg = nullptr);

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
  int rslot;                 // the reduction's four results go to results[4*rslot .. 4*rslot+3]
};

enum PackMode { kPackNone = 0, kPackF32 = 1, kPackBf16 = 2, kPackF32Frag = 3 };

// Byte offset of state column col's component c at source j in the FP32
// B-fragment layout of the block-sparse FP32 product (same 12 KiB per 16-source
// chunk as the split-BF16 states): fragment ((J*3 + c)*2 + col/32)*2 + s/4,
// lane (col % 32) + 32 (k % 2), float s % 4, with k = j % 16, s = k / 2.
__host__ __device__ inline int64_t f32_frag_offset(int col, int c, int j) {
  const int J = j / kTileK, k = j % kTileK, st = k / 2;
  const int lane = (col & 31) + 32 * (k & 1);
  const int64_t frag = (((int64_t)J * 3 + c) * 2 + col / 32) * 2 + st / 4;
  return frag * kFragBytes + (int64_t)lane * 16 + (st % 4) * 4;
}

struct PackSpec {
  int mode;
  float* f32;                // [col][3n]
  int64_t col_stride;
  uint8_t* bf16;             // fragment layout (split-BF16, or FP32 with kPackF32Frag)
};

int vec_blocks_per_state(int n3);
// With one block per state (3n <= 2048) the vector kernel writes the
// reductions' results itself (no combine launch).
int launch_vec(const VecDesc* d_descs, int nstates, int n, double* partials, double* results, const PackSpec& pack,
               bool reduce, hipStream_t stream);

struct VecDescsInline {
  VecDesc d[kInlineStates];
};
// launch_vec with the descriptors by value (nstates <= kInlineStates)
int launch_vec_inline(const VecDesc* h_descs, int nstates, int n, double* partials, double* results,
                      const PackSpec& pack, bool reduce, hipStream_t stream);

// Small problems on the FP32 dense path: the stage argument formed in the
// demag + LL launch.  Every workgroup evaluates the whole combination of
// h_pre[s] (the vector kernel's arithmetic, in its order) into LDS, writes its
// own cells of it to h_pre[s].dst, and runs the FP32 demag rows and the LL
// right-hand side from LDS: one launch per right-hand side instead of two
// (bit-identical to launch_vec_inline with an FP32 pack followed by
// launch_demag_llg_f32).  Up to kInlineStates states, n <= kStageFusedMaxCells,
// no reduction in h_pre.
constexpr int kStageFusedMaxCells = 1024;
int launch_stage_demag_llg_f32(const float* K, const LlgGrid& g, const VecDesc* h_pre, const LlgDesc* h_descs,
                               int nstates, hipStream_t stream);

}  // namespace mt
