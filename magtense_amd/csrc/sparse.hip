// Threshold-sparse demag tensor (dem_appr 'threshold' / 'threshold_fraction').
//
// Reference: ApplyThresholdDense and FindThresholdFraction
// (source/MagTenseMicroMag/DemagAuxFunctions.f90:20-122, 353-428),
// ConvertDenseToSparse_s (source/MagTenseMicroMag/util_micromag.f90:81-138),
// the sparse field product of updateDemagfield
// (LandauLifshitzEquationSolver.f90:581-624) and its CUDA form
// (MagTenseCudaBlas.cu:106-297).
//
// Layout: six CSR matrices (xx, xy, xz, yy, yz, zz), rows = receiver cells,
// columns ascending, 64-bit row pointers, 32-bit columns, FP32 values.  An
// entry is kept when |K| > threshold (FP32 comparison).  The threshold for a
// fraction is found by the reference's FP32 bisection over the positive
// entries, with the counts formed on the device.
//
// Product: H_a(i) = pref * sum_b sum_{j in row i of K_ab} K_ab(i,j) m_b(j) for
// every state.  With many states one wavefront owns a receiver row and its 64
// lanes are the states: each CSR entry is read once (uniform address) and
// multiplies a contiguous 64-state row of the transposed magnetisation.  With
// few states the lanes split the row's entries instead and reduce at the end.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_bf16.h>

#include "mt_common.h"
#include "mt_kernels.h"

namespace mt {
namespace {

// per-row count of |K| > thr, one block per (row, component)
__global__ void __launch_bounds__(256) csr_count_kernel(const float* __restrict__ K, int n, float thr,
                                                        int64_t* __restrict__ counts) {
  const int i = blockIdx.x, c = blockIdx.y;
  const float* row = K + (size_t)c * n * n + (size_t)i * n;
  int cnt = 0;
  for (int j = threadIdx.x; j < n; j += blockDim.x) cnt += fabsf(row[j]) > thr;
  __shared__ int red[256];
  red[threadIdx.x] = cnt;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[(size_t)c * (n + 1) + i + 1] = red[0];
}

// in-place inclusive scan of counts[c][1..n] (counts[c][0] = 0): one block per component
__global__ void __launch_bounds__(1024) csr_scan_kernel(int64_t* __restrict__ rowptr, int n) {
  int64_t* r = rowptr + (size_t)blockIdx.x * (n + 1);
  __shared__ int64_t part[1024];
  const int T = blockDim.x;
  const int per = (n + T - 1) / T;
  const int b = threadIdx.x * per, e = min(n, b + per);
  int64_t s = 0;
  for (int k = b; k < e; ++k) s += r[k + 1];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int t = 0; t < T; ++t) {
      int64_t v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int64_t acc = part[threadIdx.x];
  for (int k = b; k < e; ++k) {
    acc += r[k + 1];
    r[k + 1] = acc;
  }
  if (threadIdx.x == 0) r[0] = 0;
}

// ordered compaction of one row: tiles of 256 entries, block-wide exclusive scan
__global__ void __launch_bounds__(256) csr_fill_kernel(const float* __restrict__ K, int n, float thr,
                                                       const int64_t* __restrict__ rowptr, const int64_t* coff,
                                                       int* __restrict__ cols, float* __restrict__ vals) {
  const int i = blockIdx.x, c = blockIdx.y;
  const float* row = K + (size_t)c * n * n + (size_t)i * n;
  int64_t pos = coff[c] + rowptr[(size_t)c * (n + 1) + i];
  __shared__ int wsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j0 = 0; j0 < n; j0 += 256) {
    int j = j0 + threadIdx.x;
    float v = j < n ? row[j] : 0.0f;
    bool keep = j < n && fabsf(v) > thr;
    unsigned long long bal = __ballot(keep);
    int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(bal);
    __syncthreads();
    int off = 0, total = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wave) off += wsum[w];
      total += wsum[w];
    }
    if (keep) {
      cols[pos + off + before] = j;
      vals[pos + off + before] = v;
    }
    pos += total;
    __syncthreads();
  }
}

// positive-entry statistics for the fraction threshold: count, min, max
__global__ void __launch_bounds__(256) pos_stats_kernel(const float* __restrict__ K, int64_t total,
                                                        unsigned long long* __restrict__ cnt,
                                                        unsigned int* __restrict__ minbits,
                                                        unsigned int* __restrict__ maxbits) {
  unsigned long long c = 0;
  float lo = FLT_MAX, hi = 0.0f;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    float v = K[t];
    if (v > 0.0f) {
      c += 1;
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
  }
  // positive floats order like their bit patterns
  atomicAdd(cnt, c);
  if (c) {
    atomicMin(minbits, __float_as_uint(lo));
    atomicMax(maxbits, __float_as_uint(hi));
  }
}

// number of positive entries <= t
__global__ void __launch_bounds__(256) count_le_kernel(const float* __restrict__ K, int64_t total, float t,
                                                       unsigned long long* __restrict__ cnt) {
  unsigned long long c = 0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    float v = K[k];
    c += (v > 0.0f && v <= t);
  }
  atomicAdd(cnt, c);
}

// [S][3n] <-> [3n][S] FP32 transposes (32 x 32 LDS tiles)
__global__ void transpose_f32_kernel(const float* __restrict__ in, int rows, int cols, int64_t in_ld,
                                     float* __restrict__ out, int64_t out_ld) {
  __shared__ float tile[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    int r = r0 + k, c = c0 + threadIdx.x;
    if (r < rows && c < cols) tile[k][threadIdx.x] = in[(int64_t)r * in_ld + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    int c = c0 + k, r = r0 + threadIdx.x;
    if (r < rows && c < cols) out[(int64_t)c * out_ld + r] = tile[threadIdx.x][k];
  }
}

struct CsrView {
  const int64_t* rowptr;  // 6 x (n+1), relative to the component's base
  const int64_t* coff;    // 6 component bases into cols/vals
  const int* cols;
  const float* vals;
};

// many states: wave per receiver row, lane = state (mT, hT are [3n][S]).
// The row's entries come in 64 at a time with one coalesced load of columns
// and values (lane k holds entry k); each entry is then broadcast from its
// lane (v_readlane) and every lane gathers its state's component.  Sixteen
// entries per unrolled group keep sixteen independent gathers in flight
// (measured: eight 12.3 ms, sixteen 10.2 ms, thirty-two 10.6 ms per product on
// the 32768-cell film); the sums still run over the entries in CSR order.
__global__ void __launch_bounds__(256) spmm_rows_kernel(CsrView A, int n, int S, const float* __restrict__ mT,
                                                        float pref, float* __restrict__ hT) {
  const int lane = threadIdx.x & 63;
  // wave-uniform row: the CSR walk runs on scalar registers
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (i >= n) return;
  const int s = lane < S ? lane : 0;
  float acc[3] = {0.0f, 0.0f, 0.0f};
  // component c = (a,b): xx xy xz yy yz zz
  const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int a = ca[c], b = cb[c];
    const int64_t base = A.coff[c];
    const int64_t q0 = base + A.rowptr[(size_t)c * (n + 1) + i], q1 = base + A.rowptr[(size_t)c * (n + 1) + i + 1];
    const float* mb = mT + (size_t)b * n * S + s;
    const float* ma = mT + (size_t)a * n * S + s;
    float sa = 0.0f, sb = 0.0f;
    for (int64_t q = q0; q < q1; q += 64) {
      const int cnt = (int)(q1 - q < 64 ? q1 - q : 64);
      const int jl = lane < cnt ? A.cols[q + lane] : 0;
      const int vl = lane < cnt ? __float_as_int(A.vals[q + lane]) : 0;
      int k = 0;
      for (; k + 16 <= cnt; k += 16) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int j = __builtin_amdgcn_readlane(jl, k + u);
          const float v = __int_as_float(__builtin_amdgcn_readlane(vl, k + u));
          sa += v * mb[(size_t)j * S];
          if (a != b) sb += v * ma[(size_t)j * S];
        }
      }
      for (; k < cnt; ++k) {
        const int j = __builtin_amdgcn_readlane(jl, k);
        const float v = __int_as_float(__builtin_amdgcn_readlane(vl, k));
        sa += v * mb[(size_t)j * S];
        if (a != b) sb += v * ma[(size_t)j * S];
      }
    }
    acc[a] += sa;
    if (a != b) acc[b] += sb;
  }
  if (lane < S)
    for (int a = 0; a < 3; ++a) hT[((size_t)a * n + i) * S + lane] = pref * acc[a];
}

// few states (S <= 4): wave per receiver row, lanes split the entries; m, h are [S][3n]
template <int S>
__global__ void __launch_bounds__(256) spmv_rows_kernel(CsrView A, int n, const float* __restrict__ m, int64_t ld,
                                                        float pref, float* __restrict__ h) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  float acc[3][S];
  for (int a = 0; a < 3; ++a)
    for (int s = 0; s < S; ++s) acc[a][s] = 0.0f;
  const int ca[6] = {0, 0, 0, 1, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const int a = ca[c], b = cb[c];
    const int64_t base = A.coff[c];
    const int64_t q0 = base + A.rowptr[(size_t)c * (n + 1) + i], q1 = base + A.rowptr[(size_t)c * (n + 1) + i + 1];
    // four iterations unrolled: their loads are issued together (measured 1.19 ms ->
    // 0.95 ms per one-state product on the 32768-cell film); each lane still sums
    // its entries in order
#pragma unroll 4
    for (int64_t q = q0 + lane; q < q1; q += 64) {
      const int j = A.cols[q];
      const float v = A.vals[q];
      for (int s = 0; s < S; ++s) {
        acc[a][s] += v * m[s * ld + (size_t)b * n + j];
        if (a != b) acc[b][s] += v * m[s * ld + (size_t)a * n + j];
      }
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int s = 0; s < S; ++s) {
      float v = acc[a][s];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) h[s * ld + (size_t)a * n + i] = pref * v;
    }
}

__device__ inline uint16_t bf16_bits(float v) {
  __hip_bfloat16 h = __float2bfloat16(v);  // round to nearest even
  return *reinterpret_cast<uint16_t*>(&h);
}
__device__ inline float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// mask[I*nJ + J] = 1 when (row tile I, k-chunk J) holds an entry with |K| > thr
// (one block per receiver row and component; every writer stores the same 1)
__global__ void __launch_bounds__(256) bsr_mark_kernel(const float* __restrict__ K, int n, float thr, int64_t nJ,
                                                       uint8_t* __restrict__ mask) {
  const int i = blockIdx.x, c = blockIdx.y;
  const float* row = K + (size_t)c * n * n + (size_t)i * n;
  uint8_t* mrow = mask + (int64_t)(i / kTileM) * nJ;
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    if (fabsf(row[j]) > thr) mrow[j / kTileK] = 1;
}

// One block per list entry e of group egroup[e]: wave w writes the 12 split
// fragments of (row tile tiles[4g + w], k-chunk chunks[e]) with the entries at
// or below the threshold set to zero (hi = bf16(v), lo = bf16(v - hi), as the
// dense split).
__global__ void __launch_bounds__(256) bsr_fill_kernel(const float* __restrict__ K, int n, float thr,
                                                       BlockSparseView bs, const int* __restrict__ egroup,
                                                       uint8_t* __restrict__ frags) {
  const int64_t e = blockIdx.x;
  const int g = egroup[e];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int I = bs.tiles[4 * g + w];
  if (I < 0) return;
  const int64_t cnt = bs.goff[g + 1] - bs.goff[g];
  const int64_t set = 4 * bs.goff[g] + w * cnt + (e - bs.goff[g]);
  const int64_t i = (int64_t)I * kTileM + (lane % 32);
  const int64_t j0 = (int64_t)bs.chunks[e] * kTileK + 8 * (lane / 32);
  const size_t plane = (size_t)n * n;
  for (int c = 0; c < 6; ++c) {
    uint16_t hi[8], lo[8];
    for (int q = 0; q < 8; ++q) {
      const int64_t j = j0 + q;
      float v = (i < n && j < n) ? K[c * plane + (size_t)i * n + j] : 0.0f;
      if (!(fabsf(v) > thr)) v = 0.0f;
      hi[q] = bf16_bits(v);
      lo[q] = bf16_bits(v - bf16_value(hi[q]));
    }
    uint4 h4, l4;
    h4.x = hi[0] | ((uint32_t)hi[1] << 16); h4.y = hi[2] | ((uint32_t)hi[3] << 16);
    h4.z = hi[4] | ((uint32_t)hi[5] << 16); h4.w = hi[6] | ((uint32_t)hi[7] << 16);
    l4.x = lo[0] | ((uint32_t)lo[1] << 16); l4.y = lo[2] | ((uint32_t)lo[3] << 16);
    l4.z = lo[4] | ((uint32_t)lo[5] << 16); l4.w = lo[6] | ((uint32_t)lo[7] << 16);
    uint8_t* f = frags + ((set * 12 + c * 2) * kFragBytes) + (int64_t)lane * 16;
    *reinterpret_cast<uint4*>(f) = h4;
    *reinterpret_cast<uint4*>(f + kFragBytes) = l4;
  }
}

// The same list entries as FP32 fragments for v_mfma_f32_32x32x2f32: lane l
// of component c's fragments 2c + p holds K_c(row 32I + l % 32, k = 16J +
// 2s + l / 32) for the steps s = 4p .. 4p + 3 (entries at or below the
// threshold zero).
__global__ void __launch_bounds__(256) bsr_fill_f32_kernel(const float* __restrict__ K, int n, float thr,
                                                           BlockSparseView bs, const int* __restrict__ egroup,
                                                           uint8_t* __restrict__ frags) {
  const int64_t e = blockIdx.x;
  const int g = egroup[e];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int I = bs.tiles[4 * g + w];
  if (I < 0) return;
  const int64_t cnt = bs.goff[g + 1] - bs.goff[g];
  const int64_t set = 4 * bs.goff[g] + w * cnt + (e - bs.goff[g]);
  const int64_t i = (int64_t)I * kTileM + (lane % 32);
  const int64_t j0 = (int64_t)bs.chunks[e] * kTileK + (lane / 32);
  const size_t plane = (size_t)n * n;
  for (int c = 0; c < 6; ++c)
    for (int p = 0; p < 2; ++p) {
      float v[4];
      for (int q = 0; q < 4; ++q) {
        const int64_t j = j0 + 2 * (4 * p + q);
        float x = (i < n && j < n) ? K[c * plane + (size_t)i * n + j] : 0.0f;
        v[q] = fabsf(x) > thr ? x : 0.0f;
      }
      *reinterpret_cast<float4*>(frags + ((set * 12 + c * 2 + p) * kFragBytes) + (int64_t)lane * 16) =
          make_float4(v[0], v[1], v[2], v[3]);
    }
}

}  // namespace

// ---- host side -------------------------------------------------------------------

void BlockSparse::release() {
  if (tiles) (void)hipFree(tiles);
  if (goff) (void)hipFree(goff);
  if (chunks) (void)hipFree(chunks);
  if (egroup) (void)hipFree(egroup);
  if (frags) (void)hipFree(frags);
  if (wseg) (void)hipFree(wseg);
  if (segs) (void)hipFree(segs);
  tiles = nullptr; goff = nullptr; chunks = nullptr; egroup = nullptr; frags = nullptr; wseg = nullptr; segs = nullptr;
  n = 0; ngroups = 0; ksplit = 1; nwg = 0; seg_max = 0; max_list = 0; entries = 0; tiles_nonempty = 0;
  for (int c = 0; c < 6; ++c) nnz[c] = 0;
}

// Marks the non-empty tiles, groups the row tiles four to a workgroup (x-
// consecutive tiles, or on a uniform grid with grid_nx % 32 == 0 the tiles at
// the same x in four consecutive grid rows, whichever needs fewer list
// entries: neighbouring rows keep nearly the same sources), and writes the
// fragments of every list entry.
int BlockSparse::build(const float* K32, int n_, float thr, int grid_nx, hipStream_t st, bool fp32) {
  release();
  f32 = fp32;
  n = n_;
  threshold = thr;
  const int64_t np = round_up(n, kTileM);
  const int nI = (int)(np / kTileM);
  const int64_t nJ = np / kTileK;
  // kept entries per component (the CSR counts)
  {
    DevArr<int64_t> cnt;
    MT_HIP(cnt.alloc((size_t)6 * (n + 1)));
    MT_HIP(hipMemsetAsync(cnt.p, 0, (size_t)6 * (n + 1) * sizeof(int64_t), st));
    hipLaunchKernelGGL(csr_count_kernel, dim3((unsigned)n, 6), dim3(256), 0, st, K32, n, thr, cnt.p);
    MT_HIP(hipGetLastError());
    hipLaunchKernelGGL(csr_scan_kernel, dim3(6), dim3(1024), 0, st, cnt.p, n);
    MT_HIP(hipGetLastError());
    for (int c = 0; c < 6; ++c)
      MT_HIP(hipMemcpyAsync(&nnz[c], cnt.p + (size_t)c * (n + 1) + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MT_HIP(hipStreamSynchronize(st));
  }
  std::vector<uint8_t> mask((size_t)nI * nJ);
  {
    DevArr<uint8_t> dmask;
    MT_HIP(dmask.alloc(mask.size()));
    MT_HIP(hipMemsetAsync(dmask.p, 0, mask.size(), st));
    hipLaunchKernelGGL(bsr_mark_kernel, dim3((unsigned)n, 6), dim3(256), 0, st, K32, n, thr, nJ, dmask.p);
    MT_HIP(hipGetLastError());
    MT_HIP(hipMemcpyAsync(mask.data(), dmask.p, mask.size(), hipMemcpyDeviceToHost, st));
    MT_HIP(hipStreamSynchronize(st));
  }
  tiles_nonempty = 0;
  for (uint8_t v : mask) tiles_nonempty += v;
  // candidate groupings: row tiles of each group, first entry always a tile
  auto consecutive = [&]() {
    std::vector<int> t;
    for (int g = 0; g < (nI + 3) / 4; ++g)
      for (int w = 0; w < 4; ++w) t.push_back(4 * g + w < nI ? 4 * g + w : -1);
    return t;
  };
  auto by_rows = [&](int T) {  // T row tiles per grid row, R grid rows
    const int R = nI / T;
    std::vector<int> t;
    for (int rb = 0; rb < (R + 3) / 4; ++rb)
      for (int tx = 0; tx < T; ++tx)
        for (int w = 0; w < 4; ++w) t.push_back(4 * rb + w < R ? (4 * rb + w) * T + tx : -1);
    return t;
  };
  auto entries_of = [&](const std::vector<int>& t) {
    int64_t tot = 0;
    std::vector<uint8_t> u((size_t)nJ);
    for (size_t g = 0; g < t.size() / 4; ++g) {
      std::fill(u.begin(), u.end(), 0);
      for (int w = 0; w < 4; ++w)
        if (t[4 * g + w] >= 0) {
          const uint8_t* r = mask.data() + (size_t)t[4 * g + w] * nJ;
          for (int64_t J = 0; J < nJ; ++J) u[J] |= r[J];
        }
      for (int64_t J = 0; J < nJ; ++J) tot += u[J];
    }
    return tot;
  };
  std::vector<int> tl = consecutive();
  if (grid_nx >= kTileM && grid_nx % kTileM == 0 && n % grid_nx == 0 && nI % (grid_nx / kTileM) == 0) {
    std::vector<int> alt = by_rows(grid_nx / kTileM);
    if (entries_of(alt) < entries_of(tl)) tl.swap(alt);
  }
  ngroups = (int)(tl.size() / 4);
  std::vector<int64_t> go(ngroups + 1, 0);
  std::vector<int> ch, eg;
  std::vector<uint8_t> u((size_t)nJ);
  for (int g = 0; g < ngroups; ++g) {
    std::fill(u.begin(), u.end(), 0);
    for (int w = 0; w < 4; ++w)
      if (tl[4 * g + w] >= 0) {
        const uint8_t* r = mask.data() + (size_t)tl[4 * g + w] * nJ;
        for (int64_t J = 0; J < nJ; ++J) u[J] |= r[J];
      }
    for (int64_t J = 0; J < nJ; ++J)
      if (u[J]) {
        ch.push_back((int)J);
        eg.push_back(g);
      }
    go[g + 1] = (int64_t)ch.size();
  }
  entries = (int64_t)ch.size();
  MT_HIP(hipMalloc(&tiles, tl.size() * sizeof(int)));
  MT_HIP(hipMalloc(&goff, go.size() * sizeof(int64_t)));
  MT_HIP(hipMalloc(&chunks, (size_t)std::max<int64_t>(entries, 1) * sizeof(int)));
  MT_HIP(hipMalloc(&egroup, (size_t)std::max<int64_t>(entries, 1) * sizeof(int)));
  MT_HIP(hipMemcpyAsync(tiles, tl.data(), tl.size() * sizeof(int), hipMemcpyHostToDevice, st));
  MT_HIP(hipMemcpyAsync(goff, go.data(), go.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (entries > 0) {
    MT_HIP(hipMemcpyAsync(chunks, ch.data(), ch.size() * sizeof(int), hipMemcpyHostToDevice, st));
    MT_HIP(hipMemcpyAsync(egroup, eg.data(), eg.size() * sizeof(int), hipMemcpyHostToDevice, st));
    MT_HIP(hipMalloc(&frags, (size_t)bytes()));
    if (f32)
      hipLaunchKernelGGL(bsr_fill_f32_kernel, dim3((unsigned)entries), dim3(256), 0, st, K32, n, thr, view(), egroup,
                         frags);
    else
      hipLaunchKernelGGL(bsr_fill_kernel, dim3((unsigned)entries), dim3(256), 0, st, K32, n, thr, view(), egroup,
                         frags);
    MT_HIP(hipGetLastError());
  } else {
    ngroups = 0;  // nothing kept: the product is zero
  }
  if (ngroups > 0) {
    int cus = 0, dev = 0;
    MT_HIP(hipGetDevice(&dev));
    MT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int rc = schedule(go, cus > 0 ? cus : 256);
    if (rc) return rc;
  }
  MT_HIP(hipStreamSynchronize(st));
  return kOk;
}

// Balanced schedule over the group lists (list lengths vary between groups
// by a factor of two and more on a thin film; with one resident workgroup per
// CU the longest list would set the time).  The concatenated lists are cut
// into nwg ranges of about entries / cus entries (at least kMinRange), one per
// workgroup; a range covers the end of one group's list, whole lists and the
// start of the next, as segments.  The pieces of one group go to partial-sum
// slots 0, 1, ... in range order; ksplit = the most pieces of any group, and a
// group with fewer pieces gets segments of no entries for its remaining slots
// (they write zeros), in the workgroup that holds its last piece.
int BlockSparse::schedule(const std::vector<int64_t>& go, int cus) {
  // ranges of at least kMinRange entries, at most kMaxRange (their chunk
  // indices are staged in LDS beside the kernel's 24 KiB)
  constexpr int64_t kMinRange = 32, kMaxRange = 16384;
  const int64_t T = std::min(std::max<int64_t>((entries + cus - 1) / cus, kMinRange), kMaxRange);
  const int nwg_used = (int)((entries + T - 1) / T);
  // a multiple of 8 (ranges of no segments at the end): the kernel gives XCD x
  // the x-th eighth of the ranges
  nwg = (int)round_up(nwg_used, 8);
  std::vector<std::vector<int4>> per(nwg);
  std::vector<int> pieces(ngroups, 0), last_w(ngroups, 0);
  max_list = 0;
  for (int g = 0; g < ngroups; ++g) {
    max_list = std::max<int>(max_list, (int)(go[g + 1] - go[g]));
    for (int64_t e = go[g]; e < go[g + 1];) {
      const int w = (int)(e / T);
      const int64_t e1 = std::min(go[g + 1], (int64_t)(w + 1) * T);
      per[w].push_back(make_int4(g, pieces[g]++, (int)(e - go[g]), (int)(e1 - e)));
      last_w[g] = w;
      e = e1;
    }
    if (go[g + 1] == go[g]) last_w[g] = std::min<int>((int)(go[g] / T), nwg_used - 1);
  }
  ksplit = 1;
  for (int g = 0; g < ngroups; ++g) ksplit = std::max(ksplit, pieces[g]);
  seg_max = 0;
  for (const auto& pw : per)
    for (const int4& q : pw) seg_max = std::max(seg_max, q.w);
  for (int g = 0; g < ngroups; ++g)
    for (int k = pieces[g]; k < ksplit; ++k) per[last_w[g]].push_back(make_int4(g, k, 0, 0));
  std::vector<int> ws(nwg + 1, 0);
  std::vector<int4> sg;
  for (int w = 0; w < nwg; ++w) {
    sg.insert(sg.end(), per[w].begin(), per[w].end());
    ws[w + 1] = (int)sg.size();
  }
  MT_HIP(hipMalloc(&wseg, ws.size() * sizeof(int)));
  MT_HIP(hipMalloc(&segs, sg.size() * sizeof(int4)));
  MT_HIP(hipMemcpy(wseg, ws.data(), ws.size() * sizeof(int), hipMemcpyHostToDevice));
  MT_HIP(hipMemcpy(segs, sg.data(), sg.size() * sizeof(int4), hipMemcpyHostToDevice));
  return kOk;
}

void SparseDemag::release() {
  delete blocks;
  blocks = nullptr;
  if (ws) (void)hipFree(ws);
  ws = nullptr;
  ws_bytes = 0;
  if (rowptr) (void)hipFree(rowptr);
  if (cols) (void)hipFree(cols);
  if (vals) (void)hipFree(vals);
  if (coff) (void)hipFree(coff);
  if (mT) (void)hipFree(mT);
  if (hT) (void)hipFree(hT);
  rowptr = nullptr; cols = nullptr; vals = nullptr; coff = nullptr; mT = nullptr; hT = nullptr;
  n = 0; tcols = 0;
  for (int c = 0; c < 6; ++c) nnz[c] = 0;
}

// FindThresholdFraction (DemagAuxFunctions.f90:353-428) over the positive
// entries of the FP32 tensor, bisection in FP32 as the reference; the counts
// use 64-bit integers (the reference's 32-bit count overflows beyond 2^31).
int threshold_for_fraction(const float* K32, int n, double fraction, float* thr, hipStream_t st) {
  // problem%demag_threshold = sngl(dem_thres) (MagTenseMicroMagPyIO.f90:117-118)
  const float frac = (float)fraction;
  if (frac >= 1.0f) {
    *thr = FLT_MAX;  // huge(threshold_var): the demag field is zero
    return kOk;
  }
  const int64_t total = (int64_t)6 * n * n;
  DevArr<unsigned long long> dcnt_arr;
  DevArr<unsigned int> dbits_arr;
  MT_HIP(dcnt_arr.alloc(1));
  MT_HIP(dbits_arr.alloc(2));
  unsigned long long* dcnt = dcnt_arr.p;
  unsigned int* dbits = dbits_arr.p;
  unsigned int init[2] = {0x7F7FFFFFu, 0u};
  MT_HIP(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), st));
  MT_HIP(hipMemcpyAsync(dbits, init, sizeof(init), hipMemcpyHostToDevice, st));
  const unsigned blocks = 4096;
  hipLaunchKernelGGL(pos_stats_kernel, dim3(blocks), dim3(256), 0, st, K32, total, dcnt, dbits, dbits + 1);
  MT_HIP(hipGetLastError());
  unsigned long long npos = 0;
  unsigned int bits[2];
  MT_HIP(hipMemcpyAsync(&npos, dcnt, sizeof(npos), hipMemcpyDeviceToHost, st));
  MT_HIP(hipMemcpyAsync(bits, dbits, sizeof(bits), hipMemcpyDeviceToHost, st));
  MT_HIP(hipStreamSynchronize(st));
  int rc = kOk;
  if (npos == 0) {
    set_error("threshold fraction: the demag tensor has no positive entries");
    rc = kErrArg;
  } else {
    float f_small, f_large;
    std::memcpy(&f_small, &bits[0], 4);
    std::memcpy(&f_large, &bits[1], 4);
    const float nf = (float)npos;
    float f_middle = f_small;
    for (int k_do = 1;; ++k_do) {
      f_middle = (f_large - f_small) / 2.0f + f_small;
      MT_HIP(hipMemsetAsync(dcnt, 0, sizeof(unsigned long long), st));
      hipLaunchKernelGGL(count_le_kernel, dim3(blocks), dim3(256), 0, st, K32, total, f_middle, dcnt);
      MT_HIP(hipGetLastError());
      unsigned long long cm = 0;
      MT_HIP(hipMemcpyAsync(&cm, dcnt, sizeof(cm), hipMemcpyDeviceToHost, st));
      MT_HIP(hipStreamSynchronize(st));
      const float cmf = (float)cm;
      if (cmf > frac * nf) f_large = f_middle;
      else f_small = f_middle;
      if (cmf >= (frac - 0.01f) * nf && cmf <= (frac + 0.01f) * nf) break;
      if (k_do > 1000) break;
    }
    *thr = f_middle;
  }
  return rc;
}

int SparseDemag::build(const float* K32, int n_, float thr, hipStream_t st) {
  release();
  n = n_;
  threshold = thr;
  MT_HIP(hipMalloc(&rowptr, (size_t)6 * (n + 1) * sizeof(int64_t)));
  MT_HIP(hipMemsetAsync(rowptr, 0, (size_t)6 * (n + 1) * sizeof(int64_t), st));
  hipLaunchKernelGGL(csr_count_kernel, dim3((unsigned)n, 6), dim3(256), 0, st, K32, n, thr, rowptr);
  MT_HIP(hipGetLastError());
  hipLaunchKernelGGL(csr_scan_kernel, dim3(6), dim3(1024), 0, st, rowptr, n);
  MT_HIP(hipGetLastError());
  int64_t last[6];
  for (int c = 0; c < 6; ++c)
    MT_HIP(hipMemcpyAsync(&last[c], rowptr + (size_t)c * (n + 1) + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MT_HIP(hipStreamSynchronize(st));
  int64_t off[6], tot = 0;
  for (int c = 0; c < 6; ++c) {
    nnz[c] = last[c];
    off[c] = tot;
    tot += last[c];
  }
  MT_HIP(hipMalloc(&coff, sizeof(off)));
  MT_HIP(hipMemcpyAsync(coff, off, sizeof(off), hipMemcpyHostToDevice, st));
  MT_HIP(hipMalloc(&cols, (size_t)(tot > 0 ? tot : 1) * sizeof(int)));
  MT_HIP(hipMalloc(&vals, (size_t)(tot > 0 ? tot : 1) * sizeof(float)));
  hipLaunchKernelGGL(csr_fill_kernel, dim3((unsigned)n, 6), dim3(256), 0, st, K32, n, thr, rowptr, coff, cols, vals);
  MT_HIP(hipGetLastError());
  MT_HIP(hipStreamSynchronize(st));
  return kOk;
}

// Host CSR of one component (reference icl sparse ABI: 1-based columns, row
// starts of length n+1 whose last entry is nnz + rowInds[0]).
int SparseDemag::load_component(int c, int n_, int nnz_c, const float* values, const int* colInds,
                                const int* rowInds, hipStream_t st) {
  if (n == 0) {
    n = n_;
    MT_HIP(hipMalloc(&rowptr, (size_t)6 * (n + 1) * sizeof(int64_t)));
    MT_HIP(hipMemset(rowptr, 0, (size_t)6 * (n + 1) * sizeof(int64_t)));
    host_cols.assign(6, {});
    host_vals.assign(6, {});
  }
  MT_CHECK(n_ == n, "sparse component size differs from the first one loaded");
  MT_CHECK(c >= 0 && c < 6, "mat_no must be 1..6");
  std::vector<int64_t> rp(n + 1);
  const int base = rowInds[0];
  for (int i = 0; i <= n; ++i) rp[i] = (int64_t)rowInds[i] - base;
  MT_CHECK(rp[n] == nnz_c, "rowInds[n] must equal nnz + rowInds[0]");
  host_cols[c].resize(nnz_c);
  host_vals[c].assign(values, values + nnz_c);
  for (int q = 0; q < nnz_c; ++q) host_cols[c][q] = colInds[q] - base;
  nnz[c] = nnz_c;
  MT_HIP(hipMemcpyAsync(rowptr + (size_t)c * (n + 1), rp.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                        st));
  // rebuild the concatenated column/value arrays
  int64_t off[6], tot = 0;
  for (int k = 0; k < 6; ++k) {
    off[k] = tot;
    tot += nnz[k];
  }
  if (cols) (void)hipFree(cols);
  if (vals) (void)hipFree(vals);
  if (!coff) MT_HIP(hipMalloc(&coff, sizeof(off)));
  MT_HIP(hipMalloc(&cols, (size_t)(tot > 0 ? tot : 1) * sizeof(int)));
  MT_HIP(hipMalloc(&vals, (size_t)(tot > 0 ? tot : 1) * sizeof(float)));
  MT_HIP(hipMemcpyAsync(coff, off, sizeof(off), hipMemcpyHostToDevice, st));
  for (int k = 0; k < 6; ++k) {
    if (!nnz[k]) continue;
    MT_HIP(hipMemcpyAsync(cols + off[k], host_cols[k].data(), nnz[k] * sizeof(int), hipMemcpyHostToDevice, st));
    MT_HIP(hipMemcpyAsync(vals + off[k], host_vals[k].data(), nnz[k] * sizeof(float), hipMemcpyHostToDevice, st));
  }
  MT_HIP(hipStreamSynchronize(st));
  return kOk;
}

// h[s*ld + a*n + i] = pref * sum K_a.(i,:) m_s,.  for s < S (m, h: [S][3n], row stride ld)
int SparseDemag::apply(const float* m, int64_t ld, int S, float pref, float* h, hipStream_t st) {
  MT_CHECK(rowptr && cols && vals && coff, "sparse demag tensor not built");
  CsrView A{rowptr, coff, cols, vals};
  const unsigned blocks = (unsigned)((n + 3) / 4);
  const int n3 = 3 * n;
  // up to 16 states: the few-state kernel on groups of at most four (the
  // 32768-cell film at 8 % density: 0.95 / 1.23 / 1.76 ms for 1 / 2 / 4
  // states; the many-state kernel takes 8.4-9.8 ms from 5 to 32 states)
  if (S <= 16) {
    for (int s0 = 0; s0 < S; s0 += 4) {
      const int g = S - s0 < 4 ? S - s0 : 4;
      const float* mg = m + (int64_t)s0 * ld;
      float* hg = h + (int64_t)s0 * ld;
      switch (g) {
        case 1: hipLaunchKernelGGL(spmv_rows_kernel<1>, dim3(blocks), dim3(256), 0, st, A, n, mg, ld, pref, hg); break;
        case 2: hipLaunchKernelGGL(spmv_rows_kernel<2>, dim3(blocks), dim3(256), 0, st, A, n, mg, ld, pref, hg); break;
        case 3: hipLaunchKernelGGL(spmv_rows_kernel<3>, dim3(blocks), dim3(256), 0, st, A, n, mg, ld, pref, hg); break;
        default: hipLaunchKernelGGL(spmv_rows_kernel<4>, dim3(blocks), dim3(256), 0, st, A, n, mg, ld, pref, hg); break;
      }
      MT_HIP(hipGetLastError());
    }
    return kOk;
  }
  MT_CHECK(S <= kMaxCols, "at most 64 states per sparse product");
  MT_CHECK((n3 + 31) / 32 <= 65535, "sparse product supports up to 699040 cells");
  if (tcols < S) {
    if (mT) (void)hipFree(mT);
    if (hT) (void)hipFree(hT);
    mT = hT = nullptr;
    MT_HIP(hipMalloc(&mT, (size_t)n3 * kMaxCols * sizeof(float)));
    MT_HIP(hipMalloc(&hT, (size_t)n3 * kMaxCols * sizeof(float)));
    tcols = kMaxCols;
  }
  // [S][3n] -> [3n][S]
  hipLaunchKernelGGL(transpose_f32_kernel, dim3((unsigned)((n3 + 31) / 32), (unsigned)((S + 31) / 32)), dim3(32, 8), 0,
                     st, m, S, n3, ld, mT, (int64_t)S);
  MT_HIP(hipGetLastError());
  hipLaunchKernelGGL(spmm_rows_kernel, dim3(blocks), dim3(256), 0, st, A, n, S, mT, pref, hT);
  MT_HIP(hipGetLastError());
  // [3n][S] -> [S][3n]
  hipLaunchKernelGGL(transpose_f32_kernel, dim3((unsigned)((S + 31) / 32), (unsigned)((n3 + 31) / 32)), dim3(32, 8), 0,
                     st, hT, n3, S, (int64_t)S, h, ld);
  MT_HIP(hipGetLastError());
  return kOk;
}

}  // namespace mt
