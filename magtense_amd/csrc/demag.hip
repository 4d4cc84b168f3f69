// (b) Dense demag-field product H = -K m over the six unique tensor blocks.
//
// Reference semantics: LandauLifshitzEquationSolver.f90:558-579,706-755 (MKL
// sgemv x9 per right-hand side) and the CUDA plug-in cu_MVMult_GetH
// (source/MagTenseFortranCuda/cuda/MagTenseCudaBlas.cu:367-407): m is cast to
// FP32, each field component is -(K_a1 m_1 + K_a2 m_2 + K_a3 m_3) in FP32, and
// the off-diagonal blocks serve both of their symmetric positions.
//
// Device paths:
//   * matvec_f32  - the plain per-state FP32 drop-in: one pass over the FP32
//                   receiver-major blocks per state, six row streams per cell.
//   * gemm_bf16x2 - the batched path: the tensor stored as BF16 hi+lo MFMA
//                   A-fragments, up to 64 states as the N dimension, three
//                   split products (hi*hi + hi*lo + lo*hi) per block on
//                   v_mfma_f32_32x32x16_bf16 with FP32 accumulation; the
//                   fragments come either from the dense fragment-tiled tensor
//                   (streamed once per pass) or, on a uniform grid, from the
//                   translation-invariant table (looked up by cell offset).
//   * matvec_f64  - the FP64 product of the adjoint's gradient-check mode.
// All write partial sums S = K m (positive); the LL kernel applies the sign,
// sums k-split partials and scales by the receiver cell's Ms.

#include <hip/hip_bf16.h>

#include <cstdlib>
#include <type_traits>

#include "llg_cell.h"
#include "matvec_rows.h"
#include "mt_common.h"
#include "mt_kernels.h"

namespace mt {

namespace {

// ---------------------------------------------------------------------------
// Plain FP32 per-state mat-vec: block = RB receiver rows of one state (the
// row sums of matvec_rows.h).
// ---------------------------------------------------------------------------
template <int RB, bool VEC4>
__global__ void __launch_bounds__(256) matvec_f32_kernel(const float* __restrict__ K, int n,
                                                         const float* __restrict__ m, int64_t m_stride,
                                                         float* __restrict__ out, int64_t out_stride) {
  const int s = blockIdx.y;
  const int i0 = blockIdx.x * RB;
  __shared__ float red[4][RB * 3];
  __shared__ float sums[RB * 3];
  matvec_rows_f32<RB, VEC4>(K, n, m + s * m_stride, i0, red, sums);
  if (threadIdx.x < RB * 3) {
    int r = threadIdx.x / 3, c = threadIdx.x % 3;
    int i = i0 + r;
    if (i < n) out[s * out_stride + (int64_t)c * n + i] = sums[threadIdx.x];
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 batched MFMA GEMM.
//   A (tensor): fragment ((I*nJ + J)*6 + comp)*2 + part, 1 KiB each, lane l
//     holds row 32I + l%32, k = 16J + 8(l/32) + [0,8).
//   B (states): fragment ((J*3 + beta)*2 + sb)*2 + part, lane l holds column
//     32sb + l%32, same k mapping.
//   C: 32x32 FP32 per (alpha, sb); lane l, register r -> column l%32,
//     row 8(r/4) + 4(l/32) + r%4.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ inline bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ---------------------------------------------------------------------------
// Split-BF16 GEMM, two-tile LDS-ring version.  A workgroup of four waves owns two
// 32-row tiles; waves (tile, sb) split the 64 state columns in two halves.
// Every k-chunk (16 columns of the tensor) is streamed HBM -> LDS with
// global_load_lds_dwordx4 (no VGPRs held by loads in flight) through a ring of
// STAGES buffers: 24 KiB of tensor fragments (two tiles x 6 components x
// hi/lo) and 12 KiB of state fragments per stage, each fragment loaded once
// for the workgroup.  One barrier per chunk: after it, the chunk is complete
// for every wave and the slot consumed one iteration earlier is free for the
// chunk STAGES-1 ahead.
// ---------------------------------------------------------------------------
// Asynchronous 16-byte-per-lane copy global -> LDS (global_load_lds_dwordx4):
// lane l's 16 bytes land at LDS byte address lds + 16*l.  Issued as inline
// assembly so that the compiler does not serialise it with a blanket vmcnt(0)
// before every LDS read; completion is tracked explicitly with s_waitcnt
// vmcnt and s_barrier by the caller.
// M0 (the LDS base of the copy, sampled at issue) is saved and restored around
// the instruction, with the one wait state an M0 write needs before an LDS DMA.
__device__ inline void lds_dma16(const void* src, uint32_t lds) {
  uint32_t saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %2\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %1, off\n\t"
      "s_mov_b32 m0, %0\n\t"
      "s_nop 0"
      : "=&s"(saved)
      : "v"(src), "s"(__builtin_amdgcn_readfirstlane(lds))
      : "memory");
}
// Same transfer with the non-temporal hint: for data read once per pass (the
// dense tensor), so that it does not displace reused lines in L2/MALL.
__device__ inline void lds_dma16_nt(const void* src, uint32_t lds) {
  uint32_t saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\t"
      "s_mov_b32 m0, %2\n\t"
      "s_nop 0\n\t"
      "global_load_lds_dwordx4 %1, off nt\n\t"
      "s_mov_b32 m0, %0\n\t"
      "s_nop 0"
      : "=&s"(saved)
      : "v"(src), "s"(__builtin_amdgcn_readfirstlane(lds))
      : "memory");
}

__device__ inline uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// s_waitcnt immediate (gfx9 encoding) that waits for vmcnt <= v only:
// vmcnt[3:0] bits 3:0, vmcnt[5:4] bits 15:14, expcnt 7 and lgkmcnt 15 = no wait.
constexpr int vmcnt_imm(int v) { return 0x0F70 | (v & 15) | ((v >> 4) << 14); }

// TOEP: Kt is the translation-invariant fragment table of an nx x ny x nz
// uniform grid (nx % 32 == 0) instead of the dense fragment-tiled tensor; the
// fragment set of (row tile I, k-chunk J) is looked up from the cell offset.
template <int STAGES, bool TOEP>
__global__ void __launch_bounds__(256) gemm_bf16x2_lds_kernel(const uint4* __restrict__ Kt,
                                                              const uint4* __restrict__ Mp, int64_t np, int ksplit,
                                                              float* __restrict__ out, int n, int64_t col_stride,
                                                              int64_t split_stride, int ncols, int gnx, int gny,
                                                              int gnz) {
  constexpr int FR = 36;             // fragments per stage: 2 tiles x 12 + 12 state fragments
  constexpr int LOADS = 9;           // per wave per stage: 6 tensor + 3 state fragments
  __shared__ uint4 smem[STAGES][FR][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tl = wave >> 1, sb = wave & 1;
  const int nJ = (int)(np / kTileK);                      // np: padded k extent (source cells)
  const int n_row_tiles = (int)((n + kTileM - 1) / kTileM);  // n: receiver rows (np for a square tensor)
  int ks = blockIdx.x % ksplit;
  int group = blockIdx.x / ksplit;
  if (TOEP) {
    // XCD-aware order (block b runs on XCD b % 8): with a k-split dividing 8,
    // XCD x takes k-split x % ksplit for every row-tile pair, so its resident
    // workgroups walk the same cell-offset sequence (below) and share table
    // slabs in its L2; otherwise each XCD takes a contiguous range of work.
    const int groups2 = (n_row_tiles + 1) / 2;
    if (ksplit <= 8 && 8 % ksplit == 0) {
      const int x = (int)(blockIdx.x % 8);
      ks = x % ksplit;
      group = (int)(blockIdx.x / 8) * (8 / ksplit) + x / ksplit;
      if (group >= groups2) return;  // whole workgroup, before any barrier
    } else {
      const int total = groups2 * ksplit;
      const int per_xcd = (total + 7) / 8;
      const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
      if (lin >= total) return;
      ks = lin % ksplit;
      group = lin / ksplit;
    }
  }
  const int jb = (int)((int64_t)nJ * ks / ksplit);
  const int je = (int)((int64_t)nJ * (ks + 1) / ksplit);
  const int64_t I = (int64_t)group * 2 + tl;
  const bool tile_ok = I < n_row_tiles;
  const bool cols_ok = 32 * sb < ncols;

  // this wave's share of the loads of one stage
  int64_t ksrc_tile[6];
  int kdst[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int f = wave * 6 + q;            // 0..23
    int t = f / 12, kf = f % 12;
    int64_t It = (int64_t)group * 2 + t;
    if (It >= n_row_tiles) It = 0;   // keep the load count uniform; never consumed
    ksrc_tile[q] = (It * nJ) * 12 + kf;
    kdst[q] = t * 12 + kf;
  }
  // translation-invariant lookup: this wave's row tile (x0 tile, iy, iz)
  const int cpr = gnx / 16, tpr = gnx / 32, U = gnx / 8 - 2;
  int tx0 = 0, tiy = 0, tiz = 0, wy = 0, wz = 0;
  if (TOEP) {
    int64_t It = (int64_t)group * 2 + (wave >> 1);
    if (It >= n_row_tiles) It = 0;
    const int irow = (int)(It / tpr);
    tx0 = (int)(It - (int64_t)irow * tpr);
    tiy = irow % gny;
    tiz = irow / gny;
    const int wrow = (int)((int64_t)group * 2 / tpr);  // the pair's first tile sets the shared chunk order
    wy = wrow % gny;
    wz = wrow / gny;
  }
  // Chunk v of the k range -> k-chunk J.  With the table the k-chunks are
  // visited by cell offset: v = ((s ny + r) cpr + x1c) takes source row
  // jy = (wy - r) mod ny in plane jz = (wz - s) mod nz, so every workgroup with
  // the same k-split reads the same (dy, dz) slab at the same step.  For each
  // tile this is a permutation of the k-chunks.
  auto issue = [&](int v, int slot) {
    int J = v;
    if (TOEP) {
      const int x1c = v % cpr;
      const int rs = v / cpr;
      const int r = rs % gny, sp = rs / gny;
      J = (((wz - sp + gnz) % gnz) * gny + (wy - r + gny) % gny) * cpr + x1c;
    }
    int64_t tset = 0;
    if (TOEP) {
      const int jrow = J / cpr;
      const int x1c = J - jrow * cpr;
      const int jy = jrow % gny, jz = jrow / gny;
      tset = ((int64_t)(tiz - jz + gnz - 1) * (2 * gny - 1) + (tiy - jy + gny - 1)) * U + (2 * tx0 - x1c + cpr - 1);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      if (TOEP)
        lds_dma16(Kt + (tset * 12 + (wave * 6 + q) % 12) * 64 + lane, lds_addr(&smem[slot][kdst[q]][0]));
      else  // the dense tensor is read once per pass
        lds_dma16_nt(Kt + (ksrc_tile[q] + (int64_t)J * 12) * 64 + lane, lds_addr(&smem[slot][kdst[q]][0]));
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      int f = wave * 3 + q;          // 0..11
      const uint4* src = Mp + ((int64_t)J * 12 + f) * 64 + lane;
      lds_dma16(src, lds_addr(&smem[slot][24 + f][0]));
    }
  };

  floatx16 acc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[a][r] = 0.0f;

  const int nchunks = je - jb;
  if (nchunks > 0) {
#pragma unroll
    for (int s = 0; s < STAGES - 1; ++s) issue(jb + (s < nchunks ? s : nchunks - 1), s);
    for (int j = 0; j < nchunks; ++j) {
      // my loads of chunk j are done once at most STAGES-2 later stages are pending
      __builtin_amdgcn_s_waitcnt(vmcnt_imm((STAGES - 2) * LOADS));
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
      {
        int jn = j + STAGES - 1;
        issue(jb + (jn < nchunks ? jn : nchunks - 1), jn % STAGES);
      }
      const int slot = j % STAGES;
      if (tile_ok && cols_ok) {
        bf16x8 kf[12], bh[3], bl[3];
#pragma unroll
        for (int f = 0; f < 12; ++f) kf[f] = as_bf16x8(smem[slot][tl * 12 + f][lane]);
#pragma unroll
        for (int beta = 0; beta < 3; ++beta) {
          bh[beta] = as_bf16x8(smem[slot][24 + (beta * 2 + sb) * 2 + 0][lane]);
          bl[beta] = as_bf16x8(smem[slot][24 + (beta * 2 + sb) * 2 + 1][lane]);
        }
        const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int beta = 0; beta < 3; ++beta) {
            const int c = comp[a][beta];
            acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kf[2 * c], bh[beta], acc[a], 0, 0, 0);
            acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kf[2 * c], bl[beta], acc[a], 0, 0, 0);
            acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kf[2 * c + 1], bh[beta], acc[a], 0, 0, 0);
          }
      }
    }
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // no LDS write may outlive the workgroup
  }
  if (!tile_ok || !cols_ok) return;
  float* o = out + (int64_t)ks * split_stride;
  const int col = 32 * sb + (lane & 31);
  if (col >= ncols) return;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      int64_t row = I * kTileM + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
      if (row < n) o[col * col_stride + (int64_t)a * n + row] = acc[a][r];
    }
}

constexpr int kFlushChunks = 128;  // chunks summed in the MFMA accumulators before they are flushed
// FP32 fragments: 24 roundings per chunk and element (eight FP32 MFMAs per
// block, three blocks per output component) instead of nine; flushing every 8
// chunks keeps the inner chain at 192 (the 256x64x2 film's smooth in-plane
// state: 1.05e-5 of the FP64 field with flushes every 128 chunks)
constexpr int kFlushChunksF32 = 8;
constexpr int kTableChainChunks = 256;  // longest partial sum of the entry-reuse table kernel

// Register-ring kernel for the dense tensor.  Each wave owns one 32-row tile
// and all 64 states (six accumulators); its A fragments (12 per k-chunk,
// 12 KiB) stream from HBM straight into a D-deep register ring with
// non-temporal loads, so D-1 chunks per wave, four waves per CU, are in
// flight without the LDS capacity bounding the lookahead.  The state
// fragments of a chunk (12 KiB, L2/MALL resident) ride in the same ring
// (3 x 16 B per thread) and are shared through a two-slot LDS buffer.  All
// loads are ordinary vector loads issued in chunk order, so the compiler's
// in-order vmcnt wait for chunk j lets chunks j+1 .. j+D-1 stay in flight.
// Row tiles come from n (receiver rows), k-chunks from np (padded sources).
// TOEP: Kt is the translation-invariant fragment table of a gnx x gny x gnz
// uniform grid; each wave looks up its tile's fragment set from the cell
// offset and the k-chunks are visited in offset order (as in the LDS kernel),
// with ordinary loads so the table stays in L2.
// BSR: Kt holds only the (row tile, k-chunk) tiles of a block-sparse tensor
// (BlockSparseView bsv): group g owns the row tiles bsv.tiles[4g .. 4g+3]
// (-1: none) and the shared chunk list bsv.chunks[bsv.goff[g] ..
// bsv.goff[g+1]); wave w's fragments of list entry l sit at fragment set
// 4 goff[g] + w cnt_g + l, so each wave still streams one contiguous region.
// Workgroups follow the balanced schedule: workgroup b runs the segments
// bsv.segs[bsv.wseg[b] .. bsv.wseg[b+1]), each a run of one group's list
// whose sums go to that segment's partial-sum slot.
// F32: the fragments hold FP32 values instead of BF16 hi/lo pairs, in the
// same 12 KiB sets: component c's two 1 KiB fragments give lane l the eight
// k-steps s = 0..7 of v_mfma_f32_32x32x2f32 (s = 4p + q: fragment 2c + p,
// float q), A = K_c(row l % 32, k = 2s + l / 32) and, for the states, B =
// m_beta(k = 2s + l / 32, column l % 32 of the half) — eight FP32 MFMAs per
// block instead of three split products.
template <int D, bool TOEP, bool BSR = false, bool F32 = false>
__global__ void __launch_bounds__(256, 1) gemm_bf16x2_reg_kernel(const uint4* __restrict__ Kt,
                                                                 const uint4* __restrict__ Mp, int64_t np,
                                                                 int ksplit, float* __restrict__ out, int n,
                                                                 int64_t col_stride, int64_t split_stride,
                                                                 int ncols, int gnx, int gny, int gnz,
                                                                 BlockSparseView bsv,
                                                                 const int* __restrict__ bchunks,
                                                                 const int4* __restrict__ bsegs,
                                                                 const int64_t* __restrict__ bgoff,
                                                                 const int* __restrict__ btiles) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  __shared__ u32x4 bs[2][12][64];
  extern __shared__ int sidx[];  // BSR: the current segment's source chunks (bsv.seg_max)
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nJ = (int)(np / kTileK);
  const int n_row_tiles = (int)((n + kTileM - 1) / kTileM);
  const int groups = (n_row_tiles + 3) / 4;
  // BSR: workgroup w runs the segments of its balanced range, one after another
  // (the BSR arrays come in as separate read-only arguments as well, so that
  // the compiler reads them with scalar loads: a vector load of the chunk index
  // would make every chunk wait for the whole ring of fragment loads)
  // XCD x (blocks b with b % 8 == x) runs the x-th eighth of the ranges
  // (gridDim.x % 8 == 0), consecutive groups of neighbouring rows (measured
  // within 1 % of the plain order: the workgroups of an XCD still sit at
  // different sources of their lists at any one time)
  const int wgi = BSR ? (int)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3)) : 0;
  const int seg0 = BSR ? bsv.wseg[wgi] : 0, seg1 = BSR ? bsv.wseg[wgi + 1] : 1;
  for (int sgi = seg0; sgi < seg1; ++sgi) {
  // XCD-aware order (block b runs on XCD b % 8): the workgroups resident on
  // one XCD share k-splits, hence the state slab they read through its L2.
  int ks, group, jb, nchunks, cnt;
  int64_t lbase = 0;
  if (BSR) {
    const int4 sg = bsegs[sgi];  // (group, slot, first entry, entries)
    group = sg.x;
    ks = sg.y;
    lbase = bgoff[group];
    cnt = (int)(bgoff[group + 1] - lbase);
    jb = sg.z;
    nchunks = sg.w;
    // the segment's source chunks go to LDS: an index read from memory inside
    // the ring (a vector load, or a scalar load sharing its counter with the
    // LDS traffic) would make chunks wait behind the fragment loads in flight
    __syncthreads();  // the previous segment's reads of the LDS buffers are done
    for (int i = threadIdx.x; i < nchunks; i += blockDim.x) sidx[i] = bchunks[lbase + jb + i];
    __syncthreads();
  } else {
    const int x = (int)(blockIdx.x % 8), q = (int)(blockIdx.x / 8);
    if (ksplit % 8 == 0) {
      const int per = ksplit / 8;
      ks = x + 8 * (q % per);
      group = q / per;
    } else if (8 % ksplit == 0) {
      ks = x % ksplit;
      group = q * (8 / ksplit) + x / ksplit;
    } else {
      ks = (int)(blockIdx.x % ksplit);
      group = (int)(blockIdx.x / ksplit);
    }
    if (group >= groups) return;  // whole workgroup, before any barrier
    cnt = nJ;
    jb = (int)((int64_t)cnt * ks / ksplit);
    nchunks = (int)((int64_t)cnt * (ks + 1) / ksplit) - jb;
  }
  const int64_t I = BSR ? (int64_t)btiles[4 * group + wave] : (int64_t)group * 4 + wave;
  const bool tile_ok = BSR ? I >= 0 : I < n_row_tiles;
  // a wave past the last tile loads the group's first tile (kept uniform, never consumed)
  const int64_t Il = tile_ok ? I : (BSR ? (int64_t)btiles[4 * group] : (int64_t)group * 4);
  const int64_t aoff = BSR ? (4 * lbase + (int64_t)(tile_ok ? wave : 0) * cnt) * 12 * 64
                           : (TOEP ? 0 : Il * nJ * 12 * 64);
  const u32x4* A = reinterpret_cast<const u32x4*>(Kt) + aoff + lane;
  const u32x4* B = reinterpret_cast<const u32x4*>(Mp) + threadIdx.x;
  const bool two_halves = ncols > 32;
  // translation-invariant lookup: this wave's tile (x0 tile, iy, iz) and the
  // group's first tile row (wy, wz), which sets the shared chunk order
  const int cpr = gnx / 16, tpr = gnx / 32, U = gnx / 8 - 2;
  int tx0 = 0, tiy = 0, tiz = 0, wy = 0, wz = 0;
  if (TOEP) {
    const int irow = (int)(Il / tpr);
    tx0 = (int)(Il - (int64_t)irow * tpr);
    tiy = irow % gny;
    tiz = irow / gny;
    const int wrow = (int)((int64_t)group * 4 / tpr);
    wy = wrow % gny;
    wz = wrow / gny;
  }

  // native vector element types keep both rings in registers
  u32x4 ar[D][12];
  u32x4 br[D][3];
  // acc sums at most kFlushChunks chunks before it is added into sum and
  // restarted: two-level FP32 summation, so a long row's partial sum does not
  // swallow the rounding of thousands of additions
  floatx16 acc[2][3], sum[2][3];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[h][a][r] = sum[h][a][r] = 0.0f;

#define MT_REG_ISSUE(v, slot)                                                                            \
  do {                                                                                                    \
    int J_ = jb + ((v) < nchunks ? (v) : nchunks - 1);                                                    \
    const u32x4* a_;                                                                                      \
    if (TOEP) {                                                                                           \
      const int x1c_ = J_ % cpr, rs_ = J_ / cpr;                                                          \
      const int jy_ = (wy - rs_ % gny + gny) % gny, jz_ = (wz - rs_ / gny + gnz) % gnz;                   \
      J_ = (jz_ * gny + jy_) * cpr + x1c_;                                                                \
      const int64_t tset_ = ((int64_t)(tiz - jz_ + gnz - 1) * (2 * gny - 1) + (tiy - jy_ + gny - 1)) * U +  \
                            (2 * tx0 - x1c_ + cpr - 1);                                                   \
      a_ = A + tset_ * 12 * 64;                                                                           \
      _Pragma("unroll") for (int f = 0; f < 12; ++f) ar[slot][f] = a_[f * 64];                            \
    } else if (BSR) {                                                                                     \
      a_ = A + (int64_t)J_ * 12 * 64;                                                                     \
      _Pragma("unroll") for (int f = 0; f < 12; ++f) ar[slot][f] = __builtin_nontemporal_load(a_ + f * 64); \
      J_ = sidx[J_ - jb];                                                                                 \
    } else {                                                                                              \
      a_ = A + (int64_t)J_ * 12 * 64;                                                                     \
      _Pragma("unroll") for (int f = 0; f < 12; ++f) ar[slot][f] = __builtin_nontemporal_load(a_ + f * 64); \
    }                                                                                                     \
    const u32x4* b_ = B + (int64_t)J_ * 12 * 64;                                                          \
    _Pragma("unroll") for (int q = 0; q < 3; ++q) br[slot][q] = b_[q * 256];                              \
  } while (0)

  // chunk j -> LDS slot j & 1, barrier, MFMAs from register slot `slot`
#define MT_REG_CONSUME(j, slot)                                                                         \
  do {                                                                                                   \
    u32x4* dst_ = &bs[(j) & 1][0][0] + threadIdx.x;                                                      \
    _Pragma("unroll") for (int q = 0; q < 3; ++q) dst_[q * 256] = br[slot][q];                          \
    __syncthreads();                                                                                     \
    if (tile_ok) {                                                                                       \
      if (F32) {                                                                                         \
        f32x4 b0[3][2], b1[3][2];                                                                        \
        _Pragma("unroll") for (int beta = 0; beta < 3; ++beta)                                           \
        _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                                  \
          b0[beta][h] = __builtin_bit_cast(f32x4, bs[(j) & 1][(beta * 2 + h) * 2 + 0][lane]);            \
          b1[beta][h] = __builtin_bit_cast(f32x4, bs[(j) & 1][(beta * 2 + h) * 2 + 1][lane]);            \
        }                                                                                                \
        _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                                  \
          if (h == 1 && !two_halves) break;                                                              \
          _Pragma("unroll") for (int a = 0; a < 3; ++a)                                                  \
          _Pragma("unroll") for (int beta = 0; beta < 3; ++beta) {                                       \
            const int c = (a == 0 ? beta : (a == 1 ? (beta == 0 ? 1 : 2 + beta) : (beta == 0 ? 2 : 3 + beta))); \
            const f32x4 k0 = __builtin_bit_cast(f32x4, ar[slot][2 * c]);                                 \
            const f32x4 k1 = __builtin_bit_cast(f32x4, ar[slot][2 * c + 1]);                             \
            _Pragma("unroll") for (int q = 0; q < 4; ++q)                                                \
              acc[h][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(k0[q], b0[beta][h][q], acc[h][a], 0, 0, 0); \
            _Pragma("unroll") for (int q = 0; q < 4; ++q)                                                \
              acc[h][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(k1[q], b1[beta][h][q], acc[h][a], 0, 0, 0); \
          }                                                                                              \
        }                                                                                                \
      } else {                                                                                           \
        bf16x8 bh[3][2], bl[3][2];                                                                       \
        _Pragma("unroll") for (int beta = 0; beta < 3; ++beta)                                           \
        _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                                  \
          bh[beta][h] = __builtin_bit_cast(bf16x8, bs[(j) & 1][(beta * 2 + h) * 2 + 0][lane]);           \
          bl[beta][h] = __builtin_bit_cast(bf16x8, bs[(j) & 1][(beta * 2 + h) * 2 + 1][lane]);           \
        }                                                                                                \
        _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                                  \
          if (h == 1 && !two_halves) break;                                                              \
          _Pragma("unroll") for (int a = 0; a < 3; ++a)                                                  \
          _Pragma("unroll") for (int beta = 0; beta < 3; ++beta) {                                       \
            const int c = (a == 0 ? beta : (a == 1 ? (beta == 0 ? 1 : 2 + beta) : (beta == 0 ? 2 : 3 + beta))); \
            const bf16x8 khi = __builtin_bit_cast(bf16x8, ar[slot][2 * c]);                              \
            const bf16x8 klo = __builtin_bit_cast(bf16x8, ar[slot][2 * c + 1]);                          \
            acc[h][a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(khi, bh[beta][h], acc[h][a], 0, 0, 0);   \
            acc[h][a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(khi, bl[beta][h], acc[h][a], 0, 0, 0);   \
            acc[h][a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(klo, bh[beta][h], acc[h][a], 0, 0, 0);   \
          }                                                                                              \
        }                                                                                                \
      }                                                                                                  \
    }                                                                                                    \
  } while (0)

  if (nchunks > 0) {
    // chunk j0 + s always sits in register slot s: no branch inside the main
    // loop, so every slot keeps one register assignment around the loop
#pragma unroll
    for (int s = 0; s < D - 1; ++s) MT_REG_ISSUE(s, s);
    int j0 = 0;
    for (; j0 + D <= nchunks; j0 += D) {
#pragma unroll
      for (int s = 0; s < D; ++s) {
        MT_REG_ISSUE(j0 + s + D - 1, (s + D - 1) % D);
        MT_REG_CONSUME(j0 + s, s);
      }
      if ((j0 + D) % (F32 ? kFlushChunksF32 : kFlushChunks) == 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            sum[h][a] += acc[h][a];
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[h][a][r] = 0.0f;
          }
      }
    }
    // fewer than D chunks left, all already issued into slots 0 .. rem - 1
    const int rem = nchunks - j0;
#pragma unroll
    for (int s = 0; s < D - 1; ++s)
      if (s < rem) MT_REG_CONSUME(j0 + s, s);
  }
#undef MT_REG_CONSUME
#undef MT_REG_ISSUE
  if (tile_ok) {
    float* o = out + (int64_t)ks * split_stride;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = 32 * h + (lane & 31);
      if (col >= ncols) break;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          int64_t row = I * kTileM + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
          if (row < n) o[col * col_stride + (int64_t)a * n + row] = sum[h][a][r] + acc[h][a][r];
        }
    }
  }
  }  // segments
}

template <int D, bool TOEP = false, bool BSR = false, bool F32 = false>
int launch_reg(const uint8_t* K, const uint8_t* Mpack, int n, int nrows, float* out, int64_t col_stride,
               int64_t split_stride, int ncols, int ksplit, hipStream_t stream, int gnx = 0, int gny = 1,
               int gnz = 1, BlockSparseView bs = BlockSparseView{}) {
  const int64_t np = round_up(n, kTileM);
  const int groups = BSR ? bs.ngroups : (int)((round_up(nrows, kTileM) / kTileM + 3) / 4);
  int64_t blocks = (int64_t)groups * ksplit;
  if (ksplit % 8 != 0 && 8 % ksplit == 0) blocks = round_up(blocks, 8);
  if (BSR) blocks = bs.nwg;  // the balanced schedule's ranges
  if (blocks == 0) return kOk;
  const size_t lds = BSR ? (size_t)bs.seg_max * sizeof(int) : 0;
  hipLaunchKernelGGL((gemm_bf16x2_reg_kernel<D, TOEP, BSR, F32>), dim3((unsigned)blocks), dim3(256), lds, stream,
                     reinterpret_cast<const uint4*>(K), reinterpret_cast<const uint4*>(Mpack), np, ksplit, out, nrows,
                     col_stride, split_stride, ncols, gnx, gny, gnz, bs, bs.chunks, bs.segs, bs.goff, bs.tiles);
  MT_HIP(hipGetLastError());
  return kOk;
}

}  // namespace

// FP64 product for the gradient-check mode of the RK4 adjoint: one wavefront
// per (receiver row, state), lanes stride over the sources, the six planes'
// rows are read once for the three output components.
__global__ void __launch_bounds__(256) matvec_f64_kernel(const double* __restrict__ K, int n,
                                                         const double* __restrict__ m, int64_t m_stride,
                                                         double* __restrict__ out, int64_t out_stride) {
  const int s = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const size_t nn = (size_t)n * n;
  const double* row = K + (size_t)i * n;
  const double* mx = m + (int64_t)s * m_stride;
  const double* my = mx + n;
  const double* mz = my + n;
  double hx = 0.0, hy = 0.0, hz = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double xx = row[j], xy = row[nn + j], xz = row[2 * nn + j];
    const double yy = row[3 * nn + j], yz = row[4 * nn + j], zz = row[5 * nn + j];
    const double a = mx[j], b = my[j], c = mz[j];
    hx += xx * a + xy * b + xz * c;
    hy += xy * a + yy * b + yz * c;
    hz += xz * a + yz * b + zz * c;
  }
  for (int o = 32; o > 0; o >>= 1) {
    hx += __shfl_xor(hx, o, 64);
    hy += __shfl_xor(hy, o, 64);
    hz += __shfl_xor(hz, o, 64);
  }
  if (lane == 0) {
    double* o = out + (int64_t)s * out_stride;
    o[i] = hx;
    o[n + i] = hy;
    o[2 * n + i] = hz;
  }
}

int launch_matvec_f64(const double* K, int n, const double* m, int64_t m_stride, double* out, int64_t out_stride,
                      int nstates, hipStream_t stream) {
  dim3 grid((unsigned)((n + 3) / 4), (unsigned)nstates);
  hipLaunchKernelGGL(matvec_f64_kernel, grid, dim3(256), 0, stream, K, n, m, m_stride, out, out_stride);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_matvec_f32(const float* K, int n, const float* m, int64_t m_stride, float* out, int64_t out_stride,
                      int nstates, hipStream_t stream) {
  constexpr int RB = 4;
  dim3 grid((unsigned)((n + RB - 1) / RB), (unsigned)nstates);
  bool vec4 = (n % 4) == 0 && (m_stride % 4) == 0;
  if (vec4)
    hipLaunchKernelGGL((matvec_f32_kernel<RB, true>), grid, dim3(256), 0, stream, K, n, m, m_stride, out, out_stride);
  else
    hipLaunchKernelGGL((matvec_f32_kernel<RB, false>), grid, dim3(256), 0, stream, K, n, m, m_stride, out,
                       out_stride);
  MT_HIP(hipGetLastError());
  return kOk;
}

// Kernel choice for measurements (MT_GEMM_VARIANT): 0 (default) picks the
// fastest measured kernel for each storage.  Dense tensor: 0 = register-ring
// kernel with 4 chunks in flight, 6 = two-tile kernel with a four-stage LDS
// DMA ring.  Table: 0 = entry-reuse kernel where nx % 128 == 0, else the
// four-tile single-stage LDS kernel, 3 = that four-tile kernel everywhere,
// 2 = two-tile four-stage kernel, 4 = register-ring kernel.
// Read at every launch (a getenv per GEMM), so that one process can compare
// the kernels.
int gemm_variant() {
  const char* e = getenv("MT_GEMM_VARIANT");
  return e ? atoi(e) : 0;
}

namespace {
int forced_ksplit() {
  static const int forced = [] {
    const char* e = getenv("MT_GEMM_KSPLIT");  // measurement override
    return e ? atoi(e) : 0;
  }();
  return forced;
}
int clamp_ksplit(int ks, int nJ) {
  if (forced_ksplit() > 0) ks = forced_ksplit();
  if (ks > nJ) ks = nJ > 0 ? nJ : 1;
  return ks;
}
}  // namespace

int gemm_ksplit_rows(int n, int nrows, int tiles_per_group);

int gemm_ksplit_for(int n, int tiles_per_group) { return gemm_ksplit_rows(n, n, tiles_per_group); }

int gemm_ksplit_rows(int n, int nrows, int tiles_per_group) {
  int64_t np = round_up(n, kTileM);
  int nJ = (int)(np / kTileK);
  int groups = (int)((round_up(nrows, kTileM) / kTileM + tiles_per_group - 1) / tiles_per_group);
  // enough workgroups to fill 256 CUs several times (measured on N = 32768 with
  // two-tile groups, 2048 workgroups at k-split 4: k-split 4 streams 6.0 TB/s,
  // 8 5.9 TB/s, 16 5.5 TB/s); fewer k-splits also mean fewer partial sums
  int ks = 4;
  while (groups * ks < 2048 && ks * 2 <= nJ) ks *= 2;
  return clamp_ksplit(ks, nJ);
}

int gemm_ksplit_dense(int n, int nrows) {
  if (gemm_variant() == 6) return gemm_ksplit_rows(n, nrows, 2);
  // register-ring kernel: one resident workgroup (four row tiles) per CU; the
  // smallest k-split that gives every CU a workgroup (measured at N = 32768:
  // k-split 1 / 2 / 4 -> 4.21 / 4.23 / 4.27 ms in one run, 4 / 8 / 16 -> 4.19 /
  // 4.33 / 4.52 ms in another, and fewer partial sums for the LL kernel to read;
  // a 4096-row block at 8 -> 0.69 ms per call, 4 (half the CUs idle) -> 0.99 ms)
  const int nJ = (int)(round_up(n, kTileM) / kTileK);
  const int groups = (int)((round_up(nrows, kTileM) / kTileM + 3) / 4);
  int ks = 1;
  while (groups * ks < 256 && ks * 2 <= nJ) ks *= 2;
  return clamp_ksplit(ks, nJ);
}

int launch_gemm_bf16x2_rows(const uint8_t* Krows, const uint8_t* Mpack, int n, int nrows, float* out,
                            int64_t col_stride, int64_t split_stride, int ncols, int ksplit, hipStream_t stream) {
  if (gemm_variant() != 6)
    return launch_reg<4>(Krows, Mpack, n, nrows, out, col_stride, split_stride, ncols, ksplit, stream);
  // variant 6: the two-tile kernel with the four-stage LDS DMA ring
  const int64_t np = round_up(n, kTileM);
  const int groups2 = (int)((round_up(nrows, kTileM) / kTileM + 1) / 2);
  hipLaunchKernelGGL((gemm_bf16x2_lds_kernel<4, false>), dim3((unsigned)(groups2 * ksplit)), dim3(256), 0, stream,
                     reinterpret_cast<const uint4*>(Krows), reinterpret_cast<const uint4*>(Mpack), np, ksplit, out,
                     nrows, col_stride, split_stride, ncols, 0, 1, 1);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_gemm_bf16x2(const uint8_t* Ksplit, const uint8_t* Mpack, int n, float* out, int64_t col_stride,
                       int64_t split_stride, int ncols, int ksplit, hipStream_t stream) {
  return launch_gemm_bf16x2_rows(Ksplit, Mpack, n, n, out, col_stride, split_stride, ncols, ksplit, stream);
}

int launch_gemm_f32(const uint8_t* Kf, const uint8_t* Mpack, int n, float* out, int64_t col_stride,
                    int64_t split_stride, int ncols, int ksplit, hipStream_t stream) {
  return launch_reg<4, false, false, true>(Kf, Mpack, n, n, out, col_stride, split_stride, ncols, ksplit, stream);
}

// Four-tile kernel for the translation-invariant table: a workgroup of four
// waves owns four 32-row tiles, each wave one tile and all 64 states (six
// accumulators).  Tensor and state fragments of a k-chunk (48 + 12 KiB) are
// DMA'd into a single LDS slot; each wave copies its 24 fragments of the chunk
// into registers, the workgroup releases the slot and DMAs the next chunk, and
// the chunk's 54 MFMAs run while that copy is in flight.  Two workgroups per CU
// interleave, so one's MFMAs cover the other's LDS reads.  (The table is
// L2-resident, so the kernel is bound by the MFMA and LDS pipes and the
// L2 -> LDS transfer rather than by HBM.)
__global__ void __launch_bounds__(256, 2) gemm_bf16x2_t4s_kernel(const uint4* __restrict__ Kt,
                                                                const uint4* __restrict__ Mp, int64_t np,
                                                                int ksplit, float* __restrict__ out, int n,
                                                                int64_t col_stride, int64_t split_stride, int ncols,
                                                                int gnx, int gny, int gnz) {
  constexpr int FR = 60;     // fragments per slot: 4 tiles x 12 + 12 state fragments
  __shared__ uint4 smem[1][FR][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nJ = (int)(np / kTileK);
  const int n_row_tiles = (int)(np / kTileM);
  const int groups = (n_row_tiles + 3) / 4;
  // XCD-aware order (block b runs on XCD b % 8).  With a k-split dividing 8,
  // XCD x takes k-split x % ksplit for every group: its resident workgroups
  // then walk the same offset sequence (below) and share the table slabs in
  // its L2.  Otherwise each XCD takes a contiguous range of work.
  int ks, group;
  const bool by_ks = ksplit <= 8 && 8 % ksplit == 0;
  if (by_ks) {
    const int xpk = 8 / ksplit;  // XCDs per k-split
    const int x = (int)(blockIdx.x % 8);
    ks = x % ksplit;
    group = (int)(blockIdx.x / 8) * xpk + x / ksplit;
    if (group >= groups) return;  // whole workgroup, before any barrier
  } else {
    const int total = groups * ksplit;
    const int per_xcd = (total + 7) / 8;
    const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (lin >= total) return;
    ks = lin % ksplit;
    group = lin / ksplit;
  }
  const int jb = (int)((int64_t)nJ * ks / ksplit);
  const int je = (int)((int64_t)nJ * (ks + 1) / ksplit);
  const int64_t I = (int64_t)group * 4 + wave;
  const bool tile_ok = I < n_row_tiles;
  const int64_t Il = tile_ok ? I : 0;  // loads stay uniform; never consumed
  const int cpr = gnx / 16, tpr = gnx / 32, U = gnx / 8 - 2;
  int tx0 = 0, tiy = 0, tiz = 0, wy = 0, wz = 0;
  {
    const int irow = (int)(Il / tpr);
    tx0 = (int)(Il - (int64_t)irow * tpr);
    tiy = irow % gny;
    tiz = irow / gny;
    // the workgroup's first tile sets the chunk order shared by its four waves
    const int64_t I0 = (int64_t)group * 4 < n_row_tiles ? (int64_t)group * 4 : 0;
    const int wrow = (int)(I0 / tpr);
    wy = wrow % gny;
    wz = wrow / gny;
  }
  // Chunk v of the k range -> k-chunk J.  The k-chunks are visited by cell
  // offset: v = ((s ny + r) cpr + x1c) takes source row jy = (wy - r) mod ny
  // in plane jz = (wz - s) mod nz, so every workgroup with the same k-split
  // reads the same (dy, dz) slab at the same step.  For each tile this is a
  // permutation of the k-chunks.
  auto chunk_of = [&](int v) {
    const int x1c = v % cpr;
    const int rs = v / cpr;
    const int r = rs % gny, sp = rs / gny;
    const int jy = (wy - r + gny) % gny, jz = (wz - sp + gnz) % gnz;
    return (jz * gny + jy) * cpr + x1c;
  };
  auto issue = [&](int v, int slot) {
    const int J = chunk_of(v);
    const int jrow = J / cpr;
    const int x1c = J - jrow * cpr;
    const int jy = jrow % gny, jz = jrow / gny;
    const int64_t base =
        (((int64_t)(tiz - jz + gnz - 1) * (2 * gny - 1) + (tiy - jy + gny - 1)) * U + (2 * tx0 - x1c + cpr - 1)) * 12;
#pragma unroll
    for (int q = 0; q < 12; ++q) lds_dma16(Kt + (base + q) * 64 + lane, lds_addr(&smem[slot][wave * 12 + q][0]));
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int f = wave * 3 + q;  // 0..11
      lds_dma16(Mp + ((int64_t)J * 12 + f) * 64 + lane, lds_addr(&smem[slot][48 + f][0]));
    }
  };
  floatx16 acc[2][3];
#pragma unroll
  for (int sb = 0; sb < 2; ++sb)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[sb][a][r] = 0.0f;

  uint4 ra[12], rb[12];
  auto mfmas = [&]() {
    const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int beta = 0; beta < 3; ++beta)
#pragma unroll
      for (int t = 0; t < 3; ++t)  // hi*hi, hi*lo, lo*hi
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int sb = 0; sb < 2; ++sb) {
            const int c = comp[a][beta];
            const bf16x8 kf = as_bf16x8(ra[2 * c + (t == 2 ? 1 : 0)]);
            const bf16x8 mf = as_bf16x8(rb[(beta * 2 + sb) * 2 + (t == 1 ? 1 : 0)]);
            acc[sb][a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kf, mf, acc[sb][a], 0, 0, 0);
          }
  };

  const int nchunks = je - jb;
  if (nchunks > 0) {
    issue(jb, 0);
    for (int j = 0; j < nchunks; ++j) {
      __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // my loads of chunk j
      __builtin_amdgcn_s_barrier();              // everyone's
      asm volatile("" ::: "memory");
#pragma unroll
      for (int q = 0; q < 12; ++q) ra[q] = smem[0][wave * 12 + q][lane];
#pragma unroll
      for (int q = 0; q < 12; ++q) rb[q] = smem[0][48 + q][lane];
      __builtin_amdgcn_s_waitcnt(0xC07F);        // lgkmcnt(0): my reads done
      __builtin_amdgcn_s_barrier();              // everyone's: the slot is free
      asm volatile("" ::: "memory");
      issue(jb + (j + 1 < nchunks ? j + 1 : j), 0);
      if (tile_ok) mfmas();
    }
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // no LDS write may outlive the workgroup
  }
  if (!tile_ok) return;
  float* o = out + (int64_t)ks * split_stride;
#pragma unroll
  for (int sb = 0; sb < 2; ++sb) {
    const int col = 32 * sb + (lane & 31);
    if (col >= ncols) continue;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        int64_t row = I * kTileM + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
        if (row < n) o[col * col_stride + (int64_t)a * n + row] = acc[sb][a][r];
      }
  }
}

// Entry-reuse kernel for the translation-invariant table (nx % 128 == 0):
// a workgroup of four waves owns four x-consecutive 32-row tiles of one grid
// row (x tiles tx0 .. tx0+3) and walks the 16-column chunks of each source row
// in x order.  Wave w at chunk x1c needs the table entry (12 fragments) of x
// offset e = 2w - x1c + cpr - 1 (relative to 2 tx0), so over a source row the
// four waves need cpr + 6 distinct entries where the four-tile kernel loads
// 4 cpr.  Entries live in an eight-slot LDS ring (slot e mod 8): a step brings
// one new entry (four at the first two chunks of a row) and one state chunk,
// DMA'd one step ahead so that the transfer runs under the current step's
// MFMAs; one barrier per step.  Slot safety (cpr % 8 == 0): the entries of
// steps g and g+1 are eight consecutive values of e at most, so the DMA for
// step g+1 (issued after the barrier that opens step g) writes only slots that
// step g does not read, and the slots it overwrites were last read at step g-1,
// whose reads every wave finished before that barrier.
__global__ void __launch_bounds__(256, 1) gemm_bf16x2_tw_kernel(const uint4* __restrict__ Kt,
                                                                const uint4* __restrict__ Mp, int ksplit,
                                                                float* __restrict__ out, int n, int64_t col_stride,
                                                                int64_t split_stride, int ncols, int gnx, int gny,
                                                                int gnz) {
  __shared__ uint4 ent[8][12][64];  // 96 KiB: table entries, slot e mod 8
  __shared__ uint4 bst[2][12][64];  // 24 KiB: state chunks, slot g mod 2
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cpr = gnx / 16, tpr = gnx / 32, U = gnx / 8 - 2;
  const int rows = gny * gnz;
  const int groups = n / kTileM / 4;
  // XCD-aware order as in the four-tile kernel (block b runs on XCD b % 8)
  int ks, group;
  if (ksplit <= 8 && 8 % ksplit == 0) {
    const int x = (int)(blockIdx.x % 8);
    ks = x % ksplit;
    group = (int)(blockIdx.x / 8) * (8 / ksplit) + x / ksplit;
  } else {
    const int per_xcd = (groups * ksplit + 7) / 8;
    const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    ks = lin % ksplit;
    group = lin < groups * ksplit ? lin / ksplit : groups;
  }
  if (group >= groups) return;  // whole workgroup, before any barrier
  const int sb0 = rows * ks / ksplit, sb1 = rows * (ks + 1) / ksplit;  // source rows of this split
  const int I0 = group * 4;
  const int irow = I0 / tpr;
  const int tx0 = I0 - irow * tpr;
  const int tiy = irow % gny, tiz = irow / gny;
  const int G = (sb1 - sb0) * cpr;
  // source row s (visited by cell offset, as in the four-tile kernel) ->
  // source cell row (jy, jz)
  auto src_row = [&](int s, int& jy, int& jz) {
    const int r = s % gny, sp = s / gny;
    jy = (tiy - r + gny) % gny;
    jz = (tiz - sp + gnz) % gnz;
  };
  // DMA of the next step to issue (source row is, chunk ix): the row's slab
  // and chunk base change only at row starts, so the divisions stay off the
  // per-step path
  int is = sb0, ix = 0;
  int64_t islab = 0, iJ0 = 0;
  auto row_setup = [&]() {
    int jy, jz;
    src_row(is, jy, jz);
    islab = ((int64_t)(tiz - jz + gnz - 1) * (2 * gny - 1) + (tiy - jy + gny - 1)) * U + 2 * tx0;
    iJ0 = ((int64_t)jz * gny + jy) * cpr;
  };
  row_setup();
  auto issue = [&](int g) {  // g: the step number of (is, ix)
    const int e0 = cpr - 1 - ix;    // the step's lowest entry
    const int ne = ix < 2 ? 4 : 1;  // entries this step is the first to use
    for (int k = 0; k < ne; ++k) {
      const int e = e0 + 2 * k;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int f = wave * 3 + q;
        lds_dma16(Kt + ((islab + e) * 12 + f) * 64 + lane, lds_addr(&ent[e & 7][f][0]));
      }
    }
    const int64_t J = iJ0 + ix;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int f = wave * 3 + q;
      lds_dma16(Mp + (J * 12 + f) * 64 + lane, lds_addr(&bst[g & 1][f][0]));
    }
    if (++ix == cpr) {
      ix = 0;
      if (++is < sb1) row_setup();
    }
  };
  auto fetch = [&](int x1c, int g, uint4 (&ra)[12], uint4 (&rb)[12]) {
    const int slot = (cpr - 1 - x1c + 2 * wave) & 7;
#pragma unroll
    for (int q = 0; q < 12; ++q) ra[q] = ent[slot][q][lane];
#pragma unroll
    for (int q = 0; q < 12; ++q) rb[q] = bst[g & 1][q][lane];
  };
  floatx16 acc[2][3];
#pragma unroll
  for (int sb = 0; sb < 2; ++sb)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[sb][a][r] = 0.0f;
  auto mfmas = [&](const uint4 (&ra)[12], const uint4 (&rb)[12]) {
    const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int beta = 0; beta < 3; ++beta)
#pragma unroll
      for (int t = 0; t < 3; ++t)  // hi*hi, hi*lo, lo*hi
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int sb = 0; sb < 2; ++sb) {
            const int c = comp[a][beta];
            const bf16x8 kf = as_bf16x8(ra[2 * c + (t == 2 ? 1 : 0)]);
            const bf16x8 mf = as_bf16x8(rb[(beta * 2 + sb) * 2 + (t == 1 ? 1 : 0)]);
            acc[sb][a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(kf, mf, acc[sb][a], 0, 0, 0);
          }
  };
  // step g: wait for my step-g DMA and my step-(g-1) LDS reads, then for
  // everyone's; read step g's fragments, DMA step g+1 (into slots step g does
  // not read) and run step g's MFMAs (the compiler waits for each fragment
  // just before its first use)
  uint4 ra[12], rb[12];
  if (G > 0) {
    issue(0);
    int x1c = 0;  // chunk of step g within its source row
    for (int g = 0; g < G; ++g) {
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
      fetch(x1c, g, ra, rb);
      if (g + 1 < G) issue(g + 1);
      mfmas(ra, rb);
      if (++x1c == cpr) x1c = 0;
    }
    __builtin_amdgcn_s_waitcnt(0);  // no LDS write may outlive the workgroup
  }
  float* o = out + (int64_t)ks * split_stride;
  const int64_t I = I0 + wave;
#pragma unroll
  for (int sb = 0; sb < 2; ++sb) {
    const int col = 32 * sb + (lane & 31);
    if (col >= ncols) continue;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t row = I * kTileM + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
        o[col * col_stride + (int64_t)a * n + row] = acc[sb][a][r];
      }
  }
}

// The entry-reuse kernel applies to grids with nx % 128 == 0 (four x tiles per
// workgroup in one grid row, cpr % 8 == 0) and k-splits that divide the source
// rows.
bool table_window_ok(int nx, int ny, int nz, int ksplit) {
  return nx % 128 == 0 && ny >= 1 && nz >= 1 && ksplit >= 1 && (ny * nz) % ksplit == 0 && gemm_variant() == 0;
}

int gemm_ksplit_table(int nx, int ny, int nz) {
  const int n = nx * ny * nz;
  if (nx % 128 != 0 || gemm_variant() != 0) return gemm_ksplit_for(n, 2);
  // one resident workgroup per CU: the smallest k-split giving every CU a
  // workgroup, within the divisors of the source rows; and no partial sum
  // over more than kTableChainChunks chunks (the kernel's accumulators are
  // single-level FP32: long chains of additions lose accuracy on coherent
  // states; 8 at 256 x 64 x 2, where chains of 512 chunks left 1.07e-5 of the
  // FP64 field on a smooth in-plane state and chains of 256 leave 7.3e-6 at
  // the same speed)
  const int rows = ny * nz, groups = n / kTileM / 4, nJ = n / kTileK;
  int ks = 1;
  while ((groups * ks < 256 || nJ / ks > kTableChainChunks) && rows % (ks * 2) == 0) ks *= 2;
  if (forced_ksplit() > 0 && rows % forced_ksplit() == 0) ks = forced_ksplit();
  return ks;
}

namespace {
// Near-band correction of the split-BF16 product on a uniform grid (the
// terms and the table words are described with near_pack_offset).  A
// workgroup owns 64 x 4 receivers of one layer and one state, one receiver per
// thread.  Per source layer it stages the (64 + 2R) x (4 + 2R) neighbourhood
// in LDS as BF16 pairs in three planes (8, 8 and 4 bytes per cell, so that
// consecutive lanes read consecutive words): (bf(dm_x), mlo_x), (bf(dm_y),
// mlo_y) | (bf(dm_z), mlo_z), (mhi_x, mhi_y) | (mhi_z, 0), zero outside the
// grid.  Each offset then costs every receiver three LDS reads, one 16-word
// scalar load of the table (the same for the whole workgroup) and 15 BF16
// pair dot products with FP32 sums.  F64: the state comes from the LL
// descriptors (FP64 stage arguments, rounded to FP32 as the product's packing
// rounds them), else from FP32 states m32 + s * m_stride.  LLG (with F64):
// instead of storing the correction, each thread evaluates the LL right-hand
// side of its cell with the demag sum the LL kernel would form, the GEMM's
// partial-sum slots 0 .. nsplit-2 of the descriptor and then the correction
// (the value the correction slot would hold, pref = 1), in that order.
constexpr int kNearTX = 64, kNearTY = 4;
typedef __bf16 near_bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline float near_dot(uint32_t t, uint32_t v, float acc) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(near_bf16x2, t), __builtin_bit_cast(near_bf16x2, v), acc,
                                         false);
}

template <bool F64, int R, bool LLG>
__global__ void __launch_bounds__(256) near_corr_kernel(const uint32_t* __restrict__ table, int Rz, int nx,
                                                        int ny, int nz, const LlgDesc* __restrict__ descs,
                                                        const float* __restrict__ m32, int64_t m_stride,
                                                        float* __restrict__ out, int64_t col_stride, float pref,
                                                        int accumulate, const LlgGrid g) {
  constexpr int hx = kNearTX + 2 * R, hy = kNearTY + 2 * R, hc = hx * hy, w = 2 * R + 1;
  __shared__ uint2 near_lds[(hc * 20 + 7) / 8];
  uint2* p0 = near_lds;                                // (W_x, W_y)
  uint2* p1 = near_lds + hc;                           // (W_z, M_xy)
  uint32_t* p2 = reinterpret_cast<uint32_t*>(near_lds + 2 * hc);  // M_z
  const int n = nx * ny * nz;
  const int tilesx = (nx + kNearTX - 1) / kNearTX, tilesy = (ny + kNearTY - 1) / kNearTY;
  int b = blockIdx.x;
  const int tx0 = (b % tilesx) * kNearTX;
  b /= tilesx;
  const int ty0 = (b % tilesy) * kNearTY;
  const int iz = b / tilesy;
  const int s = blockIdx.y;
  const int lx = threadIdx.x % kNearTX, ly = threadIdx.x / kNearTX;
  const int ix = tx0 + lx, iy = ty0 + ly;
  const double* m64 = F64 ? descs[s].m : nullptr;
  const float* mf = F64 ? nullptr : m32 + (int64_t)s * m_stride;
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (int oz = -Rz; oz <= Rz; ++oz) {  // receiver - source layer
    const int jz = iz - oz;
    if (jz < 0 || jz >= nz) continue;  // uniform over the workgroup
    __syncthreads();                   // the previous layer's reads are done
    for (int h = threadIdx.x; h < hc; h += blockDim.x) {
      const int gx = tx0 - R + h % hx, gy = ty0 - R + h / hx;
      uint32_t wv[3] = {0u, 0u, 0u}, mh[3] = {0u, 0u, 0u};
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
        const int64_t j = gx + (int64_t)gy * nx + (int64_t)jz * nx * ny;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float f = F64 ? (float)m64[(int64_t)c * n + j] : mf[(int64_t)c * n + j];
          const uint32_t hi = (uint32_t)__builtin_bit_cast(uint16_t, __float2bfloat16(f));
          const float r = f - __uint_as_float(hi << 16);
          const uint32_t lo = (uint32_t)__builtin_bit_cast(uint16_t, __float2bfloat16(r));
          const float dm = r - __uint_as_float(lo << 16);
          wv[c] = (uint32_t)__builtin_bit_cast(uint16_t, __float2bfloat16(dm)) | lo << 16;
          mh[c] = hi;
        }
      }
      p0[h] = make_uint2(wv[0], wv[1]);
      p1[h] = make_uint2(wv[2], mh[0] | mh[1] << 16);
      p2[h] = mh[2];
    }
    __syncthreads();
    const uint32_t* tz = table + (int64_t)(oz + Rz) * w * w * kNearWords;
#pragma unroll
    for (int oy = -R; oy <= R; ++oy) {
      // source (ix - ox, iy - oy): halo cell (lx - ox + R, ly - oy + R)
      const int hrow = (ly - oy + R) * hx + lx + R;
      const uint32_t* ty = tz + (oy + R) * w * kNearWords;
#pragma unroll
      for (int ox = -R; ox <= R; ++ox) {
        const uint2 v0 = p0[hrow - ox], v1 = p1[hrow - ox];
        const uint32_t v2 = p2[hrow - ox];
        const uint32_t* t = ty + (ox + R) * kNearWords;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          float x = acc[a];
          x = near_dot(t[5 * a], v0.x, x);
          x = near_dot(t[5 * a + 1], v0.y, x);
          x = near_dot(t[5 * a + 2], v1.x, x);
          x = near_dot(t[5 * a + 3], v1.y, x);
          acc[a] = near_dot(t[5 * a + 4], v2, x);
        }
      }
    }
  }
  if (ix >= nx || iy >= ny) return;
  const int64_t i = ix + (int64_t)iy * nx + (int64_t)iz * nx * ny;
  if (LLG) {
    const LlgDesc& D = descs[s];
    llg_cell(g, D, (int)i, [&](int c) {
      float v = 0.0f;
      for (int k = 0; k + 1 < D.nsplit; ++k) v += D.hpart[(int64_t)k * D.split_stride + (int64_t)c * n + i];
      v += acc[c];
      return v;
    });
    return;
  }
  float* o = out + (int64_t)s * col_stride + i;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    o[(int64_t)a * n] = accumulate ? o[(int64_t)a * n] + pref * acc[a] : pref * acc[a];
}
}  // namespace

int near_band_rz(int R, int nz) { return R < nz - 1 ? R : nz - 1; }

int64_t near_table_words(int R, int nz) {
  const int rz = near_band_rz(R, nz);
  return (int64_t)(2 * R + 1) * (2 * R + 1) * (2 * rz + 1) * kNearWords;
}

int launch_near_correction(const uint32_t* table, int R, int nx, int ny, int nz, const LlgDesc* descs,
                           const float* m32, int64_t m_stride, int nstates, float* out, int64_t col_stride, float pref,
                           bool accumulate, hipStream_t stream, const LlgGrid* llg) {
  if (R < 1 || R > 4 || nx < 1 || ny < 1 || nz < 1 || nstates < 1 || nstates > 65535) {
    set_error("near-band correction: radius 1..4, a non-empty grid and 1..65535 states");
    return kErrArg;
  }
  if (llg && (!descs || pref != 1.0f || accumulate || llg->n != nx * ny * nz)) {
    set_error("near-band correction with the LL right-hand side: FP64 descriptors, pref 1, no accumulation");
    return kErrArg;
  }
  const int Rz = near_band_rz(R, nz);
  const int tiles = ((nx + kNearTX - 1) / kNearTX) * ((ny + kNearTY - 1) / kNearTY) * nz;
  const dim3 grid((unsigned)tiles, (unsigned)nstates);
  const int acc = accumulate ? 1 : 0;
  // the radius as a template argument: the offset loops unroll, the halo
  // index arithmetic and the table offsets are constants
  const LlgGrid g = llg ? *llg : LlgGrid{};
  auto go = [&](auto mode, auto r) {
    // mode: 0 FP32 states, 1 FP64 stage arguments, 2 those with the LL right-hand side
    constexpr int M = decltype(mode)::value;
    hipLaunchKernelGGL((near_corr_kernel<M != 0, decltype(r)::value, M == 2>), grid, dim3(256), 0, stream, table, Rz,
                       nx, ny, nz, descs, m32, m_stride, out, col_stride, pref, acc, g);
  };
  auto by_mode = [&](auto r) {
    if (llg) go(std::integral_constant<int, 2>{}, r);
    else if (descs) go(std::integral_constant<int, 1>{}, r);
    else go(std::integral_constant<int, 0>{}, r);
  };
  switch (R) {
    case 1: by_mode(std::integral_constant<int, 1>{}); break;
    case 2: by_mode(std::integral_constant<int, 2>{}); break;
    case 3: by_mode(std::integral_constant<int, 3>{}); break;
    default: by_mode(std::integral_constant<int, 4>{}); break;
  }
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_gemm_bsr(const uint8_t* frags, const BlockSparseView& bs, bool f32, const uint8_t* Mpack, int n, float* out,
                    int64_t col_stride, int64_t split_stride, int ncols, int ksplit, hipStream_t stream) {
  if (bs.ngroups == 0) {  // no kept entry anywhere: the partial sums are zero
    for (int k = 0; k < ksplit; ++k)
      for (int c = 0; c < ncols; ++c)
        MT_HIP(hipMemsetAsync(out + (int64_t)k * split_stride + (int64_t)c * col_stride, 0, 3 * (size_t)n * sizeof(float),
                              stream));
    return kOk;
  }
  if (f32)
    return launch_reg<4, false, true, true>(frags, Mpack, n, n, out, col_stride, split_stride, ncols, ksplit, stream, 0,
                                            1, 1, bs);
  return launch_reg<4, false, true>(frags, Mpack, n, n, out, col_stride, split_stride, ncols, ksplit, stream, 0, 1, 1,
                                    bs);
}

int launch_gemm_bf16x2_toeplitz(const uint8_t* Ktab, int nx, int ny, int nz, const uint8_t* Mpack, float* out,
                                int64_t col_stride, int64_t split_stride, int ncols, int ksplit, hipStream_t stream) {
  if (nx % 32 != 0 || nx < 32 || ny < 1 || nz < 1) {
    set_error("the translation-invariant GEMM needs nx % 32 == 0");
    return kErrArg;
  }
  const int n = nx * ny * nz;
  if (table_window_ok(nx, ny, nz, ksplit)) {
    const int groups = n / kTileM / 4;
    const int blocks = (groups * ksplit + 7) / 8 * 8;
    hipLaunchKernelGGL(gemm_bf16x2_tw_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       reinterpret_cast<const uint4*>(Ktab), reinterpret_cast<const uint4*>(Mpack), ksplit, out, n,
                       col_stride, split_stride, ncols, nx, ny, nz);
    MT_HIP(hipGetLastError());
    return kOk;
  }
  if (gemm_variant() == 4)
    return launch_reg<4, true>(Ktab, Mpack, n, n, out, col_stride, split_stride, ncols, ksplit, stream, nx, ny, nz);
  if (gemm_variant() != 2) {
    // default: single-stage four-tile kernel, two workgroups per CU (measured at
    // N = 32768, 64 states: 3.19 ms, against 3.31 ms for a two-stage version of
    // it and 3.50 ms for the two-tile four-stage kernel)
    const int groups4 = (n / kTileM + 3) / 4;
    const int blocks4 = (groups4 * ksplit + 7) / 8 * 8;
    hipLaunchKernelGGL(gemm_bf16x2_t4s_kernel, dim3((unsigned)blocks4), dim3(256), 0, stream,
                       reinterpret_cast<const uint4*>(Ktab), reinterpret_cast<const uint4*>(Mpack), (int64_t)n,
                       ksplit, out, n, col_stride, split_stride, ncols, nx, ny, nz);
    MT_HIP(hipGetLastError());
    return kOk;
  }
  // variant 2: the two-tile kernel with the four-stage LDS DMA ring
  const int groups = (n / kTileM + 1) / 2;
  const int blocks = (groups * ksplit + 7) / 8 * 8;  // a multiple of the XCD count (see the kernel's mapping)
  hipLaunchKernelGGL((gemm_bf16x2_lds_kernel<4, true>), dim3((unsigned)blocks), dim3(256), 0, stream,
                     reinterpret_cast<const uint4*>(Ktab), reinterpret_cast<const uint4*>(Mpack), (int64_t)n, ksplit,
                     out, n, col_stride, split_stride, ncols, nx, ny, nz);
  MT_HIP(hipGetLastError());
  return kOk;
}

}  // namespace mt
