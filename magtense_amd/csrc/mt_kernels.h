// Kernel descriptors and launchers shared by the host runtime.
#pragma once

#include "mt_common.h"

namespace mt {

// ---- (a) tensor assembly ------------------------------------------------------
int launch_assemble(const double* src, const double* abc, int abc_stride, const double* rot, const double* pts,
                    int n_src, int n_pts, float* out32, double* out64, uint8_t* split, hipStream_t stream);
// Prism tiles: tile records of 32 doubles (abc, offset, R, Rinv, R*M, symmetry
// ops, symmetry flag, pad); points Fortran-ordered (npt,3).
int launch_tiles_tensor(const double* tiles, int nt, const double* pts, int npt, double* out, hipStream_t stream);
int launch_tiles_field(const double* tiles, int nt, const double* pts, int npt, const double* Nin, double* H,
                       hipStream_t stream);
int launch_split(const float* K, int n, uint8_t* split, hipStream_t stream);

// ---- (b) demag product ------------------------------------------------------------
// out[s*out_stride + c*n + i] = sum_j K_c.(i,j) m_s,.(j)   (positive K m)
int launch_matvec_f32(const float* K, int n, const float* m, int64_t m_stride, float* out, int64_t out_stride,
                      int nstates, hipStream_t stream);
int gemm_ksplit_for(int n, int waves);
// out[ks*split_stride + col*col_stride + c*n + i] = k-split partial of K m
int launch_gemm_bf16x2(const uint8_t* Ksplit, const uint8_t* Mpack, int n, float* out, int64_t col_stride,
                       int64_t split_stride, int ncols, int ksplit, hipStream_t stream);

// ---- (c) LL right-hand side ------------------------------------------------------
struct LlgGrid {
  int n, nx, ny, nz;
  double idx2, idy2, idz2;   // 1/dx^2 etc.
  const double* jfact;       // A0 / (mu0 Ms_i)
  const double* kfact;       // K0_i / (mu0 Ms_i)
  const float* ms32;         // Ms_i in FP32 (demag scaling)
  const double* aniso;       // 6 per cell (u u^T), or null when K0 == 0 everywhere
};

struct LlgDesc {
  const double* m;           // argument, 3n SoA
  double* out;               // dm/dt, 3n SoA (may be null)
  const float* hpart;        // demag partial sums for this state's column
  int64_t split_stride;      // distance between k-split partials
  int nsplit;
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
