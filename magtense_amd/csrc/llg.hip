// (c) Fused Landau-Lifshitz right-hand side and (d) the batched vector
// operations of the RKSUITE 'M' stepper.
//
// LL right-hand side, per cell and state, in FP64 (demag term FP32):
//   reference dmdt_fct and helpers, LandauLifshitzEquationSolver.f90:220-273,
//   exchange :440-474 (Laplacian with the Neumann rows of :1194-1522),
//   external field :485-514, anisotropy :523-548, demag scaling :749-754.
// All effective fields are kept negated as in the reference; the API's alpha
// is alpha_LL * gamma.
//
// Vector operations: one generic kernel forms linear combinations of state
// vectors (stage arguments, interpolants, copies) and, in the same pass,
// the weighted max-norms / inner products the step control needs
// (RKSuitef90.f90 step :1790-1874, stepa :1876-1955, stepb :1957-1976,
// step_integrate :1262-1270, stiff :2047-2357).  Reductions are two-pass
// (per-block partials, then a fixed-order combine), so sums are deterministic.

#include <hip/hip_bf16.h>

#include "llg_cell.h"
#include "matvec_rows.h"
#include "mt_common.h"
#include "mt_kernels.h"

namespace mt {

namespace {

__device__ inline uint16_t bf16_bits(float v) {
  __hip_bfloat16 h = __float2bfloat16(v);
  return *reinterpret_cast<uint16_t*>(&h);
}
__device__ inline float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// One thread per (cell, state).  The grid is 1-D with a multiple of 8 blocks;
// block b runs on XCD b % 8, and XCD x takes the contiguous x-th eighth of the
// (state, cell block) range, so the exchange stencil's neighbour rows and
// planes are read through the L2 that already holds them.  INL: descriptors
// by value (inl) instead of from descs.
template <bool INL>
__global__ void __launch_bounds__(256) llg_kernel(const LlgGrid g, const LlgDesc* __restrict__ descs,
                                                  LlgDescsInline inl, int cell_blocks, int nstates) {
  const int total = cell_blocks * nstates;
  const int per_xcd = (int)(gridDim.x / 8);
  const int w = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
  if (w >= total) return;
  const LlgDesc& D = INL ? inl.d[w / cell_blocks] : descs[w / cell_blocks];
  const int i = (w % cell_blocks) * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  llg_cell(g, D, i, [&](int c) {
    float s = 0.0f;
    for (int k = 0; k < D.nsplit; ++k) s += D.hpart[(int64_t)k * D.split_stride + (int64_t)c * g.n + i];
    return s;
  });
}

// The FP32 dense path's demag rows and LL right-hand side in one launch:
// block = kFusedRows receiver rows of state blockIdx.y (the sums of
// matvec_rows_f32), then one thread per row evaluates llg_cell with them.
template <bool INL, bool VEC4, int RB>
__global__ void __launch_bounds__(256) demag_llg_f32_kernel(const float* __restrict__ K, const LlgGrid g,
                                                            const LlgDesc* __restrict__ descs, LlgDescsInline inl,
                                                            const float* __restrict__ m32, int64_t m_stride) {
  const int s = blockIdx.y;
  const int i0 = blockIdx.x * RB;
  __shared__ float red[4][RB * 3];
  __shared__ float sums[RB * 3];
  matvec_rows_f32<RB, VEC4>(K, g.n, m32 + s * m_stride, i0, red, sums);
  const int r = threadIdx.x;
  if (r >= RB || i0 + r >= g.n) return;
  const LlgDesc& D = INL ? inl.d[s] : descs[s];
  // one partial-sum slot: S = 0 + row sum, as the LL kernel forms it
  llg_cell(g, D, i0 + r, [&](int c) {
    float v = 0.0f;
    v += sums[r * 3 + c];
    return v;
  });
}

// Stage combination of a small problem, carried by value into the fused launch.
struct StageDesc {
  double* dst;
  const double* base;
  double bcoef, bconst, post_scale;
  int nterm;
  const double* src[kMaxTerms];
  double coef[kMaxTerms];
};
struct StageDescsInline {
  StageDesc d[kInlineStates];
};

// Stage argument + FP32 demag rows + LL right-hand side (launch_stage_demag_llg_f32).
// Block = RB receiver rows of state blockIdx.y.  The whole stage vector goes to
// LDS (3n FP64, then its FP32 rounding); each element sums bconst, bcoef*base
// and the terms in the vector kernel's order, all of its operands loaded
// before the sum (an index past nterm reloads operand 0, which is not used).
template <bool VEC4, int RB>
__global__ void __launch_bounds__(256) stage_demag_llg_f32_kernel(const float* __restrict__ K, const LlgGrid g,
                                                                  StageDescsInline sd, LlgDescsInline inl) {
  extern __shared__ double stage_lds[];
  const int n = g.n, n3 = 3 * n;
  const int s = blockIdx.y;
  const int i0 = blockIdx.x * RB;
  double* m64 = stage_lds;
  float* m32 = reinterpret_cast<float*>(stage_lds + n3);
  const StageDesc& S = sd.d[s];
  const int nt = S.nterm;
  for (int e = threadIdx.x; e < n3; e += blockDim.x) {
    double x[kMaxTerms];
    if (nt > 0) {
#pragma unroll
      for (int k = 0; k < kMaxTerms; ++k) x[k] = S.src[k < nt ? k : 0][e];
    }
    const double b = S.base ? S.base[e] : 0.0;
    double v = S.bconst;
    if (S.base) v += S.bcoef * b;
#pragma unroll
    for (int k = 0; k < kMaxTerms; ++k)
      if (k < nt) v += S.coef[k] * x[k];
    v = S.post_scale * v;
    m64[e] = v;
    m32[e] = (float)v;
  }
  __syncthreads();
  if (threadIdx.x < 3 * RB) {  // this block's cells of the stage argument
    const int c = threadIdx.x / RB, i = i0 + threadIdx.x % RB;
    if (i < n) S.dst[(int64_t)c * n + i] = m64[c * n + i];
  }
  __shared__ float red[4][RB * 3];
  __shared__ float sums[RB * 3];
  matvec_rows_f32<RB, VEC4>(K, n, m32, i0, red, sums);
  const int r = threadIdx.x;
  if (r >= RB || i0 + r >= n) return;
  LlgDesc D = inl.d[s];
  D.m = m64;  // the stage argument from LDS (other blocks write its global copy)
  llg_cell(g, D, i0 + r, [&](int c) {
    float v = 0.0f;
    v += sums[r * 3 + c];
    return v;
  });
}

__device__ inline double nan_to_inf(double v) { return isnan(v) ? INFINITY : v; }

// Generic vector op.  A block covers 2048 consecutive elements of one state;
// thread t handles elements t + 256q (q < 8), so every load and store of the
// combination and the reductions is coalesced.  Each operand vector is loaded
// for all eight elements before any of them is used: a wave keeps eight
// 512-byte loads in flight per operand instead of waiting on one at a time.
// Per element the arithmetic (and its order) is that of the scalar form
//   v = post_scale * (bconst + bcoef*base + sum_k coef[k]*src[k]).
// The BF16 pack needs eight consecutive elements per 16-byte store and goes
// through LDS.
constexpr int kVecE = 8;

__device__ inline void load_e(const double* p, int e0, int n3, double (&x)[kVecE]) {
#pragma unroll
  for (int q = 0; q < kVecE; ++q) {
    const int e = e0 + 256 * q;
    x[q] = e < n3 ? p[e] : 0.0;
  }
}

// acc[q] += coef(k) * ptr(k)[e_q] for k = 0 .. cnt-1, in k order.  Operand
// k+1's loads are issued before operand k is accumulated, so two operands'
// loads are in flight; the last iteration reloads its own operand (cached) to
// keep the loop body free of branches, so the wait before each accumulation
// leaves exactly the next operand's eight loads outstanding.
template <typename PF, typename CF>
__device__ inline void accumulate(double (&acc)[kVecE], int cnt, PF ptr, CF coef, int e0, int n3) {
  if (cnt <= 0) return;
  double x[kVecE], y[kVecE];
  load_e(ptr(0), e0, n3, x);
  for (int k = 0; k < cnt; ++k) {
    load_e(ptr(k + 1 < cnt ? k + 1 : k), e0, n3, y);
    const double c = coef(k);
#pragma unroll
    for (int q = 0; q < kVecE; ++q) acc[q] += c * x[q];
#pragma unroll
    for (int q = 0; q < kVecE; ++q) x[q] = y[q];
  }
}

// INL: descriptors by value (inl) instead of from descs.  With one block
// per state the block's reduction is the result (written to results).
template <bool INL>
__global__ void __launch_bounds__(256) vec_kernel(const VecDesc* __restrict__ descs, VecDescsInline inl, int n,
                                                  int n3, int blocks_per_state, double* __restrict__ partials,
                                                  double* __restrict__ results, PackSpec pack) {
  const VecDesc& D = INL ? inl.d[blockIdx.y] : descs[blockIdx.y];
  const int t0 = blockIdx.x * 2048;
  const int e0 = t0 + threadIdx.x;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
  const int kind = D.rkind;
  const bool do_pack = D.pack_col >= 0 && pack.mode != kPackNone && D.dst;
  __shared__ float tile[2048];
  bool ok[kVecE];
#pragma unroll
  for (int q = 0; q < kVecE; ++q) ok[q] = e0 + 256 * q < n3;
  double v[kVecE], x[kVecE];
#pragma unroll
  for (int q = 0; q < kVecE; ++q) v[q] = 0.0;
  if (D.dst) {
#pragma unroll
    for (int q = 0; q < kVecE; ++q) v[q] = D.bconst;
    // base first, then the terms in order
    const int nb = D.base ? 1 : 0;
    accumulate(
        v, nb + D.nterm, [&](int k) { return k < nb ? D.base : D.src[k - nb]; },
        [&](int k) { return k < nb ? D.bcoef : D.coef[k - nb]; }, e0, n3);
    const double ps = D.post_scale;
#pragma unroll
    for (int q = 0; q < kVecE; ++q) {
      v[q] = ps * v[q];
      if (ok[q]) D.dst[e0 + 256 * q] = v[q];
    }
  }
  if (do_pack) {
    if (pack.mode == kPackF32) {
#pragma unroll
      for (int q = 0; q < kVecE; ++q)
        if (ok[q]) pack.f32[(int64_t)D.pack_col * pack.col_stride + e0 + 256 * q] = (float)v[q];
    } else if (pack.mode == kPackF32Frag) {
#pragma unroll
      for (int q = 0; q < kVecE; ++q)
        if (ok[q]) {
          const int e = e0 + 256 * q, c = e / n, j = e - c * n;
          *reinterpret_cast<float*>(pack.bf16 + f32_frag_offset(D.pack_col, c, j)) = (float)v[q];
        }
    } else {
#pragma unroll
      for (int q = 0; q < kVecE; ++q) tile[threadIdx.x + 256 * q] = (float)v[q];
    }
  }
  switch (kind) {
    case kRedNone:
      break;
    case kRedStepB:    // max |sum c_k s_k| / max(0.5(|y|+|ynew|), thresh)
    case kRedFinal: {  // errest = scale * sum c_k s_k; w = max(0.5(|y|+|ynew|), thresh)
      double s[kVecE], a[kVecE];
#pragma unroll
      for (int q = 0; q < kVecE; ++q) s[q] = 0.0;
      accumulate(
          s, D.rn, [&](int k) { return D.rsrc[k]; }, [&](int k) { return D.rcoef[k]; }, e0, n3);
      load_e(D.ra, e0, n3, a);
      load_e(D.rb, e0, n3, x);
      const double th = D.thresh, sc = D.rscale;
#pragma unroll
      for (int q = 0; q < kVecE; ++q) {
        if (!ok[q]) continue;
        const double w = fmax(0.5 * (fabs(a[q]) + fabs(x[q])), th);
        if (kind == kRedStepB) {
          r0 = fmax(r0, nan_to_inf(fabs(s[q]) / w));
        } else {
          const double sv = sc * s[q];
          const int e = e0 + 256 * q;
          if (D.rd) D.rd[e] = sv;
          if (D.rw) D.rw[e] = w;
          r0 = fmax(r0, nan_to_inf(fabs(sv / w)));
        }
      }
      break;
    }
    case kRedStepA: {  // w = max(w, |ynew|); max |y|/w, |ynew|/w, |ynew-y|/w, |f-yp|/w
      double ya[kVecE], yn[kVecE], wo[kVecE], fc[kVecE];
      load_e(D.ra, e0, n3, ya);
      load_e(D.rb, e0, n3, yn);
      load_e(D.rw, e0, n3, wo);
      load_e(D.rc, e0, n3, fc);
      load_e(D.rd, e0, n3, x);
#pragma unroll
      for (int q = 0; q < kVecE; ++q) {
        if (!ok[q]) continue;
        const double w = fmax(wo[q], fabs(yn[q]));
        D.rw[e0 + 256 * q] = w;
        r0 = fmax(r0, nan_to_inf(fabs(ya[q]) / w));
        r1 = fmax(r1, nan_to_inf(fabs(yn[q]) / w));
        r2 = fmax(r2, nan_to_inf(fabs(yn[q] - ya[q]) / w));
        r3 = fmax(r3, nan_to_inf(fabs(fc[q] - x[q]) / w));
      }
      break;
    }
    case kRedYpNorm: {  // w = max(|y|, thresh); max over y != 0 of |yp|/w
      double ya[kVecE];
      load_e(D.ra, e0, n3, ya);
      load_e(D.rb, e0, n3, x);
#pragma unroll
      for (int q = 0; q < kVecE; ++q) {
        if (!ok[q]) continue;
        const double w = fmax(fabs(ya[q]), D.thresh);
        D.rw[e0 + 256 * q] = w;
        if (ya[q] != 0.0) r0 = fmax(r0, nan_to_inf(fabs(x[q]) / w));
      }
      break;
    }
    case kRedWeightsAvg: {  // w = max(0.5(|a|+|b|), thresh)
      double a[kVecE];
      load_e(D.ra, e0, n3, a);
      load_e(D.rb, e0, n3, x);
#pragma unroll
      for (int q = 0; q < kVecE; ++q)
        if (ok[q]) D.rw[e0 + 256 * q] = fmax(0.5 * (fabs(a[q]) + fabs(x[q])), D.thresh);
      break;
    }
    case kRedWdot: {  // sums of (p/w)(q/w) for up to three pairs (rsrc[2k], rsrc[2k+1])
      double w[kVecE], p[kVecE];
      load_e(D.rw, e0, n3, w);
      for (int k = 0; k < D.rn && k < 3; ++k) {
        load_e(D.rsrc[2 * k], e0, n3, p);
        load_e(D.rsrc[2 * k + 1], e0, n3, x);
        double r = 0.0;
#pragma unroll
        for (int q = 0; q < kVecE; ++q)
          if (ok[q]) r += (p[q] / w[q]) * (x[q] / w[q]);
        if (k == 0) r0 = r;
        else if (k == 1) r1 = r;
        else r2 = r;
      }
      break;
    }
    case kRedMaxDiff: {
      double a[kVecE];
      load_e(D.ra, e0, n3, a);
      load_e(D.rb, e0, n3, x);
#pragma unroll
      for (int q = 0; q < kVecE; ++q)
        if (ok[q]) r0 = fmax(r0, nan_to_inf(fabs(a[q] - x[q])));
      break;
    }
    default:
      break;
  }
  if (do_pack && pack.mode == kPackBf16) {
    __syncthreads();
    const int l0 = threadIdx.x * 8;  // eight consecutive elements of this block's tile
    const int p0 = t0 + l0;
    if (p0 < n3) {
      if ((n & 7) == 0 && p0 + 8 <= n3) {
        const int c = p0 / n, j0 = p0 - c * n;
        uint32_t hw[4], lw[4];
        for (int q = 0; q < 4; ++q) {
          float f0 = tile[l0 + 2 * q], f1 = tile[l0 + 2 * q + 1];
          uint16_t h0 = bf16_bits(f0), h1 = bf16_bits(f1);
          uint16_t g0 = bf16_bits(f0 - bf16_value(h0)), g1 = bf16_bits(f1 - bf16_value(h1));
          hw[q] = h0 | ((uint32_t)h1 << 16);
          lw[q] = g0 | ((uint32_t)g1 << 16);
        }
        int J = j0 / kTileK, half = (j0 % kTileK) / 8;
        int lane = (D.pack_col & 31) + 32 * half;
        int sb = D.pack_col / 32;
        int64_t frag = (((int64_t)J * 3 + c) * 2 + sb) * 2;
        uint4* p = reinterpret_cast<uint4*>(pack.bf16 + frag * kFragBytes + lane * 16);
        p[0] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        p[kFragBytes / 16] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
      } else {
        for (int q = 0; q < 8 && p0 + q < n3; ++q) {
          int e = p0 + q;
          int c = e / n, j = e - c * n;
          float f = tile[l0 + q];
          uint16_t hi = bf16_bits(f);
          uint16_t lo = bf16_bits(f - bf16_value(hi));
          int J = j / kTileK, half = (j % kTileK) / 8;
          int lane = (D.pack_col & 31) + 32 * half;
          int sb = D.pack_col / 32;
          int64_t frag = (((int64_t)J * 3 + c) * 2 + sb) * 2;
          uint16_t* p = reinterpret_cast<uint16_t*>(pack.bf16 + frag * kFragBytes + lane * 16) + (j % 8);
          p[0] = hi;
          p[kFragBytes / 2] = lo;
        }
      }
    }
  }
  if (kind == kRedNone || kind == kRedWeightsAvg) return;
  const bool is_sum = kind == kRedWdot;
  // block reduction of four values
  __shared__ double red[4][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double rv[4] = {r0, r1, r2, r3};
  for (int k = 0; k < 4; ++k) {
    double x = rv[k];
    for (int off = 32; off > 0; off >>= 1) {
      double o = __shfl_xor(x, off, 64);
      x = is_sum ? x + o : fmax(x, o);
    }
    if (lane == 0) red[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    int k = threadIdx.x;
    double x = red[0][k];
    for (int w = 1; w < 4; ++w) x = is_sum ? x + red[w][k] : fmax(x, red[w][k]);
    if (blocks_per_state == 1 && results) results[D.rslot * 4 + k] = x;
    else partials[((int64_t)blockIdx.y * blocks_per_state + blockIdx.x) * 4 + k] = x;
  }
}

// One thread per (state, value): fixed-order combine of the block partials.
template <bool INL>
__global__ void finalize_kernel(const VecDesc* __restrict__ descs, VecDescsInline inl, int nstates,
                                int blocks_per_state, const double* __restrict__ partials,
                                double* __restrict__ results) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nstates * 4) return;
  int s = t / 4, k = t % 4;
  const VecDesc& D = INL ? inl.d[s] : descs[s];
  int kind = D.rkind;
  // a descriptor without a reduction leaves its slot alone: a result written
  // by an earlier launch may still be waiting to be read
  if (kind == kRedNone || kind == kRedWeightsAvg) return;
  bool is_sum = kind == kRedWdot;
  double x = partials[((int64_t)s * blocks_per_state) * 4 + k];
  for (int b = 1; b < blocks_per_state; ++b) {
    double o = partials[((int64_t)s * blocks_per_state + b) * 4 + k];
    x = is_sum ? x + o : fmax(x, o);
  }
  results[D.rslot * 4 + k] = x;
}

}  // namespace

// kernel arguments are limited to 4 KiB
static_assert(sizeof(VecDescsInline) + 256 <= 4096, "vector kernel arguments exceed 4 KiB");
static_assert(sizeof(LlgDescsInline) + sizeof(LlgGrid) + 256 <= 4096, "LL kernel arguments exceed 4 KiB");

namespace {
template <bool INL>
int launch_llg_t(const LlgGrid& g, const LlgDesc* descs, const LlgDescsInline& inl, int nstates, hipStream_t stream) {
  const int cell_blocks = (g.n + 255) / 256;
  const int64_t total = (int64_t)cell_blocks * nstates;
  const int64_t blocks = (total + 7) / 8 * 8;
  if (blocks > INT32_MAX) {
    set_error("LL kernel: too many cells x states for one launch");
    return kErrArg;
  }
  hipLaunchKernelGGL(llg_kernel<INL>, dim3((unsigned)blocks), dim3(256), 0, stream, g, descs, inl, cell_blocks,
                     nstates);
  MT_HIP(hipGetLastError());
  return kOk;
}

template <bool INL>
int launch_vec_t(const VecDesc* descs, const VecDescsInline& inl, int nstates, int n, double* partials,
                 double* results, const PackSpec& pack, bool reduce, hipStream_t stream) {
  int n3 = 3 * n;
  int bps = vec_blocks_per_state(n3);
  dim3 grid((unsigned)bps, (unsigned)nstates);
  hipLaunchKernelGGL(vec_kernel<INL>, grid, dim3(256), 0, stream, descs, inl, n, n3, bps, partials, results, pack);
  MT_HIP(hipGetLastError());
  if (reduce && bps > 1) {
    int t = nstates * 4;
    hipLaunchKernelGGL(finalize_kernel<INL>, dim3((unsigned)((t + 127) / 128)), dim3(128), 0, stream, descs, inl,
                       nstates, bps, partials, results);
    MT_HIP(hipGetLastError());
  }
  return kOk;
}
}  // namespace

int launch_llg(const LlgGrid& g, const LlgDesc* d_descs, int nstates, hipStream_t stream) {
  return launch_llg_t<false>(g, d_descs, LlgDescsInline{}, nstates, stream);
}

int launch_llg_inline(const LlgGrid& g, const LlgDesc* h_descs, int nstates, hipStream_t stream) {
  if (nstates < 1 || nstates > kInlineStates) {
    set_error("LL kernel: descriptors by value hold at most kInlineStates states");
    return kErrArg;
  }
  LlgDescsInline inl{};
  for (int s = 0; s < nstates; ++s) inl.d[s] = h_descs[s];
  return launch_llg_t<true>(g, nullptr, inl, nstates, stream);
}

namespace {
template <int RB>
int launch_demag_llg_rb(const float* K, const LlgGrid& g, const LlgDesc* d_descs, const LlgDescsInline& inl,
                        const float* m32, int64_t m_stride, int nstates, hipStream_t stream) {
  dim3 grid((unsigned)((g.n + RB - 1) / RB), (unsigned)nstates);
  const bool vec4 = (g.n % 4) == 0 && (m_stride % 4) == 0;
  if (d_descs) {
    if (vec4)
      hipLaunchKernelGGL((demag_llg_f32_kernel<false, true, RB>), grid, dim3(256), 0, stream, K, g, d_descs, inl, m32,
                         m_stride);
    else
      hipLaunchKernelGGL((demag_llg_f32_kernel<false, false, RB>), grid, dim3(256), 0, stream, K, g, d_descs, inl,
                         m32, m_stride);
  } else {
    if (vec4)
      hipLaunchKernelGGL((demag_llg_f32_kernel<true, true, RB>), grid, dim3(256), 0, stream, K, g, nullptr, inl, m32,
                         m_stride);
    else
      hipLaunchKernelGGL((demag_llg_f32_kernel<true, false, RB>), grid, dim3(256), 0, stream, K, g, nullptr, inl, m32,
                         m_stride);
  }
  MT_HIP(hipGetLastError());
  return kOk;
}
}  // namespace

int launch_demag_llg_f32(const float* K, const LlgGrid& g, const LlgDesc* d_descs, const LlgDesc* h_descs,
                         const float* m32, int64_t m_stride, int nstates, hipStream_t stream) {
  LlgDescsInline inl{};
  if (!d_descs) {
    if (nstates < 1 || nstates > kInlineStates) {
      set_error("demag + LL kernel: descriptors by value hold at most kInlineStates states");
      return kErrArg;
    }
    for (int s = 0; s < nstates; ++s) inl.d[s] = h_descs[s];
  }
  // four receiver rows per block (the 128 x 32 film's 402.7 MB tensor: 2 / 4 / 8
  // rows -> 72.2 / 66.8 / 114.4 us per launch)
  return launch_demag_llg_rb<4>(K, g, d_descs, inl, m32, m_stride, nstates, stream);
}

static_assert(sizeof(StageDescsInline) + sizeof(LlgDescsInline) + sizeof(LlgGrid) + 256 <= 4096,
              "stage + demag + LL kernel arguments exceed 4 KiB");

int launch_stage_demag_llg_f32(const float* K, const LlgGrid& g, const VecDesc* h_pre, const LlgDesc* h_descs,
                               int nstates, hipStream_t stream) {
  if (nstates < 1 || nstates > kInlineStates || g.n < 1 || g.n > kStageFusedMaxCells) {
    set_error("stage + demag + LL kernel: 1..kInlineStates states, 1..kStageFusedMaxCells cells");
    return kErrArg;
  }
  StageDescsInline sd{};
  LlgDescsInline inl{};
  for (int s = 0; s < nstates; ++s) {
    const VecDesc& v = h_pre[s];
    if (!v.dst || v.rkind != kRedNone || v.nterm < 0 || v.nterm > kMaxTerms) {
      set_error("stage + demag + LL kernel: a stage combination with a destination and no reduction");
      return kErrArg;
    }
    StageDesc& d = sd.d[s];
    d.dst = v.dst;
    d.base = v.base;
    d.bcoef = v.bcoef;
    d.bconst = v.bconst;
    d.post_scale = v.post_scale;
    d.nterm = v.nterm;
    for (int k = 0; k < v.nterm; ++k) {
      d.src[k] = v.src[k];
      d.coef[k] = v.coef[k];
    }
    inl.d[s] = h_descs[s];
  }
  constexpr int RB = 4;
  const dim3 grid((unsigned)((g.n + RB - 1) / RB), (unsigned)nstates);
  const size_t lds = (size_t)3 * g.n * (sizeof(double) + sizeof(float));
  if (g.n % 4 == 0)
    hipLaunchKernelGGL((stage_demag_llg_f32_kernel<true, RB>), grid, dim3(256), lds, stream, K, g, sd, inl);
  else
    hipLaunchKernelGGL((stage_demag_llg_f32_kernel<false, RB>), grid, dim3(256), lds, stream, K, g, sd, inl);
  MT_HIP(hipGetLastError());
  return kOk;
}

int vec_blocks_per_state(int n3) { return (n3 + 256 * 8 - 1) / (256 * 8); }

int launch_vec(const VecDesc* d_descs, int nstates, int n, double* partials, double* results, const PackSpec& pack,
               bool reduce, hipStream_t stream) {
  return launch_vec_t<false>(d_descs, VecDescsInline{}, nstates, n, partials, results, pack, reduce, stream);
}

int launch_vec_inline(const VecDesc* h_descs, int nstates, int n, double* partials, double* results,
                      const PackSpec& pack, bool reduce, hipStream_t stream) {
  if (nstates < 1 || nstates > kInlineStates) {
    set_error("vector kernel: descriptors by value hold at most kInlineStates states");
    return kErrArg;
  }
  VecDescsInline inl{};
  for (int s = 0; s < nstates; ++s) inl.d[s] = h_descs[s];
  return launch_vec_t<true>(nullptr, inl, nstates, n, partials, results, pack, reduce, stream);
}

}  // namespace mt
