// (a) Analytical prism demagnetization tensor assembly on gfx950.
//
// Math: the Smith et al. (2010) point tensor of a uniformly magnetized prism as
// used by MagTense, including its guards and sign convention:
//   source/TileDemagTensor/TileDemagTensor/TileNComponents.f90:629-724
//   source/TileDemagTensor/TileDemagTensor/TileRectangularPrismTensor.f90:7-149
// and the per-pair loop of LandauLifshitzEquationSolver.f90:935-984 /
// DemagFieldGetSolution.f90:685-730 (identity rotation: offset = p_i - c_j).
//
// GPU form: one thread evaluates eight consecutive source columns j of one
// receiver row i, entirely in FP64 (the eight-term arctangent sums and the
// logarithm ratios cancel catastrophically in the far field, so FP64 is kept
// for every pair; the FP64 pipe of gfx950 makes this cheap relative to one
// integration step).  The eight corner distances are shared by all six
// components.  Results are written in up to three layouts in the same pass:
//   * FP32 receiver-major planes K[c][i*N + j]   (the plain per-state path),
//   * FP64 receiver-major planes                 (validation),
//   * split BF16 hi/lo MFMA A-fragments           (the batched MFMA path).
// Contraction into FMA is disabled here so that every rounding matches the
// serial restatement of the reference formula.

#include <hip/hip_bf16.h>

#include "mt_common.h"
#include "mt_kernels.h"

#pragma clang fp contract(off)

namespace mt {
namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kSclH = 1.0001, kSclL = 0.9999;

__device__ inline double sq(double v) { return v * v; }

// One guarded arctangent argument f/g/h: numerator p*q, denominator d*R, with
// the "d == 0" limit rule of TileRectangularPrismTensor.f90:53-61 (the
// coordinate that zeroes d is scaled by 1.0001 and 0.9999 and averaged).
// d = half_edge - s, where s is the (signed) coordinate along that axis; pos
// is the position of d's square in the reference's sum of squares (0,1,2).
__device__ inline double guarded_ratio(double half_edge, double s, double p, double q, double d, double R, int pos) {
  if (d == 0.0) {
    double dh = half_edge - s * kSclH;
    double dl = half_edge - s * kSclL;
    double rh = pos == 0 ? sq(dh) + sq(p) + sq(q) : (pos == 1 ? sq(p) + sq(dh) + sq(q) : sq(p) + sq(q) + sq(dh));
    double rl = pos == 0 ? sq(dl) + sq(p) + sq(q) : (pos == 1 ? sq(p) + sq(dl) + sq(q) : sq(p) + sq(q) + sq(dl));
    double fh = p * q / (dh * sqrt(rh));
    double fl = p * q / (dl * sqrt(rl));
    return 0.5 * (fh + fl);
  }
  return p * q / (d * R);
}

// The eight factors of nom/denom (TileNComponents.f90:665-666, 684-685, 702-703)
// for one log term.  which: 0 -> FF (xy), 1 -> GG (yz), 2 -> HH (xz).  hx[0] =
// a/2-x, hx[1] = a/2+x (likewise hy, hz); R[i][j][k] the matching corner
// distances.  A factor with edge sign -1 uses -(half edge + coordinate).
__device__ inline void log_products(const double hx[2], const double hy[2], const double hz[2],
                                    const double R[2][2][2], int which, double& nom, double& den) {
  // (sa,sb,sc) -> index 0 for +1, 1 for -1, reference factor order
  const int sn[4][3] = {{0, 0, 0}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
  const int sd[4][3] = {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {1, 1, 1}};
  double fn[4], fd[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const int* s = pass == 0 ? sn[k] : sd[k];
      double lead = which == 0 ? (s[2] ? -hz[1] : hz[0]) : (which == 1 ? (s[0] ? -hx[1] : hx[0]) : (s[1] ? -hy[1] : hy[0]));
      (pass == 0 ? fn : fd)[k] = lead + R[s[0]][s[1]][s[2]];
    }
  }
  nom = fn[0] * fn[1] * fn[2] * fn[3];
  den = fd[0] * fd[1] * fd[2] * fd[3];
}

__device__ inline void corner_terms(double a, double b, double c, double x, double y, double z, double hx[2],
                                    double hy[2], double hz[2], double R[2][2][2]) {
  hx[0] = a / 2. - x; hx[1] = a / 2. - (-x);
  hy[0] = b / 2. - y; hy[1] = b / 2. - (-y);
  hz[0] = c / 2. - z; hz[1] = c / 2. - (-z);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) R[i][j][k] = sqrt(sq(hx[i]) + sq(hy[j]) + sq(hz[k]));
}

__device__ inline double off_diagonal(double a, double b, double c, double x, double y, double z, int which,
                                      double nom, double den) {
  double v;
  if (den == 0.0 || nom == 0.0) {
    // getF_limit (TileRectangularPrismTensor.f90:7-38): all three coordinates scaled
    double hx[2], hy[2], hz[2], R[2][2][2];
    double nl, dl, nh, dh;
    corner_terms(a, b, c, x * kSclL, y * kSclL, z * kSclL, hx, hy, hz, R);
    log_products(hx, hy, hz, R, which, nl, dl);
    corner_terms(a, b, c, x * kSclH, y * kSclH, z * kSclH, hx, hy, hz, R);
    log_products(hx, hy, hz, R, which, nh, dh);
    double lim = (0.5 * (nl + nh)) / (0.5 * (dl + dh));
    v = -1. / (4. * kPi) * log(lim);
  } else {
    v = -1. / (4. * kPi) * log(nom / den);
  }
  if (fabs((nom - den) / den) < 10. * 2.220446049250313e-16) v = 0.0;
  return v;
}

}  // namespace

// K = -N_standard for a prism with full edges (a,b,c) at offset (x,y,z) from
// its centre.  out: xx, xy, xz, yy, yz, zz.
__device__ void prism_k6(double a, double b, double c, double x, double y, double z, double out[6]) {
  double hx[2], hy[2], hz[2], R[2][2][2];
  corner_terms(a, b, c, x, y, z, hx, hy, hz, R);
  double sx[2] = {x, -x}, sy[2] = {y, -y}, sz[2] = {z, -z};
  // sign combinations in the order of TileNComponents.f90:647-649:
  // (x,y,z) (-x,y,z) (x,-y,z) (x,y,-z) (-x,-y,z) (x,-y,-z) (-x,y,-z) (-x,-y,-z)
  const int cmb[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
  double sxx = 0.0, syy = 0.0, szz = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int i = cmb[q][0], j = cmb[q][1], k = cmb[q][2];
    double r = R[i][j][k];
    double fx = guarded_ratio(a / 2., sx[i], hy[j], hz[k], hx[i], r, 0);
    double fy = guarded_ratio(b / 2., sy[j], hx[i], hz[k], hy[j], r, 1);
    double fz = guarded_ratio(c / 2., sz[k], hx[i], hy[j], hz[k], r, 2);
    sxx = (q == 0) ? atan(fx) : sxx + atan(fx);
    syy = (q == 0) ? atan(fy) : syy + atan(fy);
    szz = (q == 0) ? atan(fz) : szz + atan(fz);
  }
  double nxx = 1. / (4. * kPi) * sxx;
  double nyy = 1. / (4. * kPi) * syy;
  double nzz = 1. / (4. * kPi) * szz;
  double nm, dn;
  log_products(hx, hy, hz, R, 0, nm, dn);
  double nxy = off_diagonal(a, b, c, x, y, z, 0, nm, dn);
  log_products(hx, hy, hz, R, 1, nm, dn);
  double nyz = off_diagonal(a, b, c, x, y, z, 1, nm, dn);
  log_products(hx, hy, hz, R, 2, nm, dn);
  double nxz = off_diagonal(a, b, c, x, y, z, 2, nm, dn);
  out[0] = -nxx;
  out[1] = -nxy;
  out[2] = -nxz;
  out[3] = -nyy;
  out[4] = -nyz;
  out[5] = -nzz;
}

namespace {

__device__ inline uint16_t bf16_bits(float v) {
  __hip_bfloat16 h = __float2bfloat16(v);  // round to nearest even
  return *reinterpret_cast<uint16_t*>(&h);
}
__device__ inline float bf16_value(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  return __uint_as_float(u);
}

// Byte offset of the 16-byte lane chunk holding (row i, cols j..j+7), j % 8 == 0,
// of component c, part p (0 hi, 1 lo) in the fragment-tiled tensor of padded size np.
__device__ inline int64_t frag_offset(int64_t np, int64_t i, int64_t j, int c, int p) {
  int64_t I = i / kTileM, J = j / kTileK;
  int64_t nJ = np / kTileK;
  int lane = (int)(i % kTileM) + kTileM * (int)((j % kTileK) / 8);
  int64_t frag = ((I * nJ + J) * 6 + c) * 2 + p;
  return frag * kFragBytes + (int64_t)lane * 16;
}

__global__ void __launch_bounds__(256) assemble_kernel(const double* __restrict__ src, const double* __restrict__ abc,
                                                       int abc_stride, const double* __restrict__ rot,
                                                       const double* __restrict__ pts, int n_src,
                                                       int n_pts, int64_t np, float* __restrict__ out32,
                                                       double* __restrict__ out64, uint8_t* __restrict__ split) {
  // Thread -> (row i, 8 columns starting at j0); for the split layout the grid
  // covers the padded rows x padded sources (np = round_up(n_src, 32)).
  int64_t groups_per_row = (split ? np : round_up(n_src, 8)) / 8;
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t i = tid / groups_per_row;
  int64_t j0 = (tid % groups_per_row) * 8;
  int64_t rows = split ? round_up(n_pts, kTileM) : n_pts;
  if (i >= rows) return;
  float v32[6][8];
  bool row_ok = i < n_pts;
  double px = 0, py = 0, pz = 0;
  if (row_ok) {
    px = pts[i * 3 + 0];
    py = pts[i * 3 + 1];
    pz = pts[i * 3 + 2];
  }
  for (int e = 0; e < 8; ++e) {
    int64_t j = j0 + e;
    double k6[6] = {0, 0, 0, 0, 0, 0};
    bool ok = row_ok && j < n_src;
    if (ok) {
      const double* s = src + j * 3;
      const double* d = abc + j * abc_stride;
      double dx = px - s[0], dy = py - s[1], dz = pz - s[2];
      if (rot) {  // diffPos = rotMat (p - offset), DemagFieldGetSolution.f90:705-709
        const double* R = rot + j * 9;
        double rx = R[0] * dx + R[1] * dy + R[2] * dz;
        double ry = R[3] * dx + R[4] * dy + R[5] * dz;
        double rz = R[6] * dx + R[7] * dy + R[8] * dz;
        dx = rx; dy = ry; dz = rz;
      }
      prism_k6(d[0], d[1], d[2], dx, dy, dz, k6);
      if (out64) {
        size_t plane = (size_t)n_pts * n_src;
        for (int c = 0; c < 6; ++c) out64[c * plane + (size_t)i * n_src + j] = k6[c];
      }
    }
    for (int c = 0; c < 6; ++c) v32[c][e] = (float)k6[c];
  }
  if (out32 && row_ok) {
    size_t plane = (size_t)n_pts * n_src;
    for (int c = 0; c < 6; ++c)
      for (int e = 0; e < 8; ++e)
        if (j0 + e < n_src) out32[c * plane + (size_t)i * n_src + j0 + e] = v32[c][e];
  }
  if (split) {
    for (int c = 0; c < 6; ++c) {
      uint16_t hi[8], lo[8];
      for (int e = 0; e < 8; ++e) {
        hi[e] = bf16_bits(v32[c][e]);
        lo[e] = bf16_bits(v32[c][e] - bf16_value(hi[e]));
      }
      uint4 h4, l4;
      h4.x = hi[0] | ((uint32_t)hi[1] << 16); h4.y = hi[2] | ((uint32_t)hi[3] << 16);
      h4.z = hi[4] | ((uint32_t)hi[5] << 16); h4.w = hi[6] | ((uint32_t)hi[7] << 16);
      l4.x = lo[0] | ((uint32_t)lo[1] << 16); l4.y = lo[2] | ((uint32_t)lo[3] << 16);
      l4.z = lo[4] | ((uint32_t)lo[5] << 16); l4.w = lo[6] | ((uint32_t)lo[7] << 16);
      *reinterpret_cast<uint4*>(split + frag_offset(np, i, j0, c, 0)) = h4;
      *reinterpret_cast<uint4*>(split + frag_offset(np, i, j0, c, 1)) = l4;
    }
  }
}

// Translation-invariant split-BF16 fragment table for a uniform grid with
// nx % 32 == 0.  Fragment set (slab = (dz, dy), u) holds, for rows r < 32 and
// k = e < 16, the tensor at cell offset (D + r - e, dy, dz) with D = 16 u -
// (nx - 16): every (row tile, k-chunk) pair of the dense tensor whose
// x0 - x1 = D and (iy - jy, iz - jz) = (dy, dz) has exactly these values.
// One thread per (slab, u, lane): eight columns, six components, hi and lo.
__global__ void __launch_bounds__(256) toeplitz_assemble_kernel(int nx, int ny, int nz, double Lx, double Ly,
                                                                double Lz, uint8_t* __restrict__ out) {
  // setupGrid (LandauLifshitzEquationSolver.f90:815-872): d = L/n, or d = L when n = 1
  const double dx = nx > 1 ? Lx / nx : Lx, dy = ny > 1 ? Ly / ny : Ly, dz = nz > 1 ? Lz / nz : Lz;
  const int U = nx / 8 - 2;
  const int64_t sets = (int64_t)(2 * nz - 1) * (2 * ny - 1) * U;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= sets * 64) return;
  const int lane = (int)(tid % 64);
  const int64_t set = tid / 64;
  const int u = (int)(set % U);
  const int64_t slab = set / U;
  const int dyi = (int)(slab % (2 * ny - 1)) - (ny - 1);
  const int dzi = (int)(slab / (2 * ny - 1)) - (nz - 1);
  const int D = 16 * u - (nx - 16);
  const int r = lane % 32, e0 = 8 * (lane / 32);
  // the offset is formed as the dense assembly forms p_i - c_j, from the cell
  // centres of one representative pair (the first cells along each axis)
  auto centre = [](int m, double L, double d, int i) { return m > 1 ? -L / 2 + i * d + d / 2 : 0.0; };
  float v32[6][8];
  for (int e = 0; e < 8; ++e) {
    const int ddx = D + r - (e0 + e);
    const int ix = ddx > 0 ? ddx : 0, jx = ddx > 0 ? 0 : -ddx;
    const int iy = dyi > 0 ? dyi : 0, jy = dyi > 0 ? 0 : -dyi;
    const int iz = dzi > 0 ? dzi : 0, jz = dzi > 0 ? 0 : -dzi;
    double k6[6];
    prism_k6(dx, dy, dz, centre(nx, Lx, dx, ix) - centre(nx, Lx, dx, jx), centre(ny, Ly, dy, iy) - centre(ny, Ly, dy, jy),
             centre(nz, Lz, dz, iz) - centre(nz, Lz, dz, jz), k6);
    for (int c = 0; c < 6; ++c) v32[c][e] = (float)k6[c];
  }
  for (int c = 0; c < 6; ++c) {
    uint16_t hi[8], lo[8];
    for (int e = 0; e < 8; ++e) {
      hi[e] = bf16_bits(v32[c][e]);
      lo[e] = bf16_bits(v32[c][e] - bf16_value(hi[e]));
    }
    uint4 h4, l4;
    h4.x = hi[0] | ((uint32_t)hi[1] << 16); h4.y = hi[2] | ((uint32_t)hi[3] << 16);
    h4.z = hi[4] | ((uint32_t)hi[5] << 16); h4.w = hi[6] | ((uint32_t)hi[7] << 16);
    l4.x = lo[0] | ((uint32_t)lo[1] << 16); l4.y = lo[2] | ((uint32_t)lo[3] << 16);
    l4.z = lo[4] | ((uint32_t)lo[5] << 16); l4.w = lo[6] | ((uint32_t)lo[7] << 16);
    const int64_t frag = (set * 6 + c) * 2;
    *reinterpret_cast<uint4*>(out + frag * kFragBytes + lane * 16) = h4;
    *reinterpret_cast<uint4*>(out + (frag + 1) * kFragBytes + lane * 16) = l4;
  }
}

// FP32 receiver-major tensor (6 x n x n) -> fragment-tiled split BF16 (padded).
__global__ void __launch_bounds__(256) split_kernel(const float* __restrict__ K, int n, int64_t np,
                                                    uint8_t* __restrict__ split) {
  int64_t groups_per_row = np / 8;
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t i = tid / groups_per_row;
  int64_t j0 = (tid % groups_per_row) * 8;
  if (i >= np) return;
  size_t plane = (size_t)n * n;
  for (int c = 0; c < 6; ++c) {
    uint16_t hi[8], lo[8];
    for (int e = 0; e < 8; ++e) {
      int64_t j = j0 + e;
      float v = (i < n && j < n) ? K[c * plane + (size_t)i * n + j] : 0.0f;
      hi[e] = bf16_bits(v);
      lo[e] = bf16_bits(v - bf16_value(hi[e]));
    }
    uint4 h4, l4;
    h4.x = hi[0] | ((uint32_t)hi[1] << 16); h4.y = hi[2] | ((uint32_t)hi[3] << 16);
    h4.z = hi[4] | ((uint32_t)hi[5] << 16); h4.w = hi[6] | ((uint32_t)hi[7] << 16);
    l4.x = lo[0] | ((uint32_t)lo[1] << 16); l4.y = lo[2] | ((uint32_t)lo[3] << 16);
    l4.z = lo[4] | ((uint32_t)lo[5] << 16); l4.w = lo[6] | ((uint32_t)lo[7] << 16);
    *reinterpret_cast<uint4*>(split + frag_offset(np, i, j0, c, 0)) = h4;
    *reinterpret_cast<uint4*>(split + frag_offset(np, i, j0, c, 1)) = l4;
  }
}

// FP32 planes -> FP32 MFMA A-fragments of the dense tensor (v_mfma_f32_32x32x2f32):
// the fragment sets of the split tensor's order, component c's fragments 2c +
// p giving lane l the k-steps s = 4p .. 4p + 3, K_c(row 32I + l % 32, k = 16J +
// 2s + l / 32); rows and sources past n are zero.  One wave per (row tile I,
// chunk J), lane l of the wave writing lane l of each of the twelve fragments,
// so every fragment is one contiguous 1 KiB store; consecutive waves take
// consecutive chunks of the same 32 rows.
__global__ void __launch_bounds__(256) f32_frag_kernel(const float* __restrict__ K, int n, int64_t np,
                                                       uint8_t* __restrict__ frags) {
  const int64_t nJ = np / kTileK;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = tid / 64;
  const int lane = (int)(tid % 64);
  const int64_t I = w / nJ, J = w % nJ;
  if (I >= np / kTileM) return;
  const int64_t i = I * kTileM + lane % 32;
  const int half = lane / 32;
  const size_t plane = (size_t)n * n;
  for (int c = 0; c < 6; ++c)
    for (int p = 0; p < 2; ++p) {
      float v[4];
      for (int q = 0; q < 4; ++q) {
        const int64_t j = J * kTileK + 2 * (4 * p + q) + half;
        v[q] = (i < n && j < n) ? K[c * plane + (size_t)i * n + j] : 0.0f;
      }
      const int64_t frag = ((I * nJ + J) * 6 + c) * 2 + p;
      *reinterpret_cast<float4*>(frags + frag * kFragBytes + (int64_t)lane * 16) = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// ---- tetrahedron (tileType 5) ---------------------------------------------------
// Sum over the four faces of the right-triangle face tensor
// (getN_tensor_tetrahedron, TileNComponents.f90:944-961; getN_Triangle and the
// Nxz / Nyz / Nzz integrals, TileTriangle.f90:26-286), FP64 as the reference
// builds it (default real of 8 bytes).  Returns false for a degenerate face
// (the reference stops); the caller writes NaN then.
namespace tet {
constexpr double kPi = 3.14159265359;  // TileTriangle.f90:6, as written

__device__ double F_Nxz(const double r[3], double yp, double l, double h) {
  return h / sqrt(h * h + l * l) *
         atanh((l * l - l * r[0] + h * r[1] - h * yp * (1 + l * l / (h * h))) /
               (sqrt(h * h + l * l) * sqrt(l * l - 2 * l * r[0] + r[0] * r[0] + r[1] * r[1] -
                                           2 * yp * (l * l - l * r[0] + h * r[1]) / h +
                                           yp * yp * (1. + l * l / (h * h)) + r[2] * r[2])));
}
__device__ double G_Nxz(const double r[3], double yp) {
  const double g = atanh((r[1] - yp) / (sqrt(r[0] * r[0] + (r[1] - yp) * (r[1] - yp) + r[2] * r[2])));
  return isfinite(g) ? g : 0.0;
}
__device__ double K_Nyz(const double r[3], double xp, double l, double h) {
  return l / sqrt(h * h + l * l) *
         atanh((h * h - h * r[1] + l * r[0] - l * xp * (1 + h * h / (l * l))) /
               (sqrt(h * h + l * l) * sqrt(h * h - 2 * h * r[1] + r[0] * r[0] + r[1] * r[1] -
                                           2 * xp * (h * h + l * r[0] - h * r[1]) / l +
                                           xp * xp * (1 + h * h / (l * l)) + r[2] * r[2])));
}
__device__ double L_Nyz(const double r[3], double xp) {
  const double g = atanh((r[0] - xp) / (sqrt((r[0] - xp) * (r[0] - xp) + r[1] * r[1] + r[2] * r[2])));
  return isfinite(g) ? g : 0.0;
}
__device__ double P_Nzz(const double r[3], double xp, double l, double h) {
  return atan((r[0] * (h - r[1]) - xp * (h * (1 - r[0] / l) - r[1]) - h * (r[0] * r[0] + r[2] * r[2]) / l) /
              (r[2] * sqrt(h * h + r[0] * r[0] + xp * xp * (1 + h * h / (l * l)) - 2 * h * r[1] + r[1] * r[1] -
                           2 * xp * (h * h + l * r[0] - h * r[1]) / l + r[2] * r[2])));
}
__device__ double Q_Nzz(const double r[3], double xp) {
  return -atan((r[0] - xp) * r[1] / (r[2] * sqrt((r[0] - xp) * (r[0] - xp) + r[1] * r[1] + r[2] * r[2])));
}

__device__ inline double nrm(const double a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ inline double dt(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void sb(const double a[3], const double b[3], double c[3]) {
  c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2];
}
__device__ inline void cr(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline void cp(double d[3], const double s[3]) { d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; }

// getN_Triangle: face v[0..2], v[3] behind it; N (row-major 3x3) += face tensor
__device__ bool triangle_add(const double vin[4][3], const double r[3], double N[9]) {
  const double numErr = 1e-25, threshold = 1e-20;
  double v[4][3];
  for (int k = 0; k < 4; ++k) cp(v[k], vin[k]);
  double a[3], b[3], c[3];
  sb(v[0], v[1], a); sb(v[1], v[2], b); sb(v[0], v[2], c);
  if (fabs(nrm(c) - (nrm(a) + nrm(b))) < numErr) return false;
  double ang[3];
  sb(v[0], v[1], a); sb(v[0], v[2], b); ang[0] = acos(dt(a, b) / (nrm(a) * nrm(b)));
  sb(v[1], v[0], a); sb(v[1], v[2], b); ang[1] = acos(dt(a, b) / (nrm(a) * nrm(b)));
  sb(v[2], v[1], a); sb(v[2], v[0], b); ang[2] = acos(dt(a, b) / (nrm(a) * nrm(b)));
  // simple_sort (SpecialFunctions.f90:13-34): repeated MINLOC, first of equals
  int ind[3];
  bool mk[3] = {true, true, true};
  double srt[3];
  for (int i = 0; i < 3; ++i) {
    int best = -1;
    for (int k = 0; k < 3; ++k)
      if (mk[k] && (best < 0 || ang[k] < ang[best])) best = k;
    srt[i] = ang[best]; ind[i] = best; mk[best] = false;
  }
  double w[3][3];
  for (int i = 0; i < 3; ++i) cp(w[i], v[ind[i]]);
  // sorted order with the largest angle in the middle: v = (w0, w2, w1)
  cp(v[0], w[0]); cp(v[1], w[2]); cp(v[2], w[1]);
  ang[0] = srt[0]; ang[1] = srt[2]; ang[2] = srt[1];
  double e1[3], e2[3], e3[3], d32[3], d41[3];
  sb(v[0], v[2], e1);
  double n1 = nrm(e1);
  for (int k = 0; k < 3; ++k) e1[k] = e1[k] / n1;
  sb(v[1], v[2], d32);
  cr(e1, d32, e3);
  double n3 = nrm(e3);
  for (int k = 0; k < 3; ++k) e3[k] = e3[k] / n3;
  sb(v[3], v[0], d41);
  if (fabs(dt(d41, e3)) <= numErr) return false;
  if (dt(e3, d41) > 0.) {
    double t[3];
    cp(t, v[0]); cp(v[0], v[2]); cp(v[2], t);
    const double ta = ang[0]; ang[0] = ang[2]; ang[2] = ta;
    sb(v[0], v[2], e1);
    n1 = nrm(e1);
    for (int k = 0; k < 3; ++k) e1[k] = e1[k] / n1;
    sb(v[1], v[2], d32);
    cr(e1, d32, e3);
    n3 = nrm(e3);
    for (int k = 0; k < 3; ++k) e3[k] = e3[k] / n3;
  }
  cr(e3, e1, e2);
  // P = [e1 e2 e3] (columns); local coordinates q' = P^T q
  sb(v[1], v[2], d32);
  const double cd = cos(ang[2]) * nrm(d32);
  double D[3], q[3];
  for (int k = 0; k < 3; ++k) D[k] = cd * e1[k] + v[2][k];
  const double* E[3] = {e1, e2, e3};
  double vp00, vp11, vp20, rt[3];
  sb(v[0], D, q); vp00 = dt(E[0], q);   // vp(1,1)
  sb(v[1], D, q); vp11 = dt(E[1], q);   // vp(2,2)
  sb(v[2], D, q); vp20 = dt(E[0], q);   // vp(1,3)
  sb(r, D, q);
  for (int k = 0; k < 3; ++k) rt[k] = dt(E[k], q);  // Pinv (r - D), Pinv = P^T
  for (int k = 0; k < 3; ++k)
    if (fabs(rt[k]) < threshold) rt[k] = rt[k] < 0 ? -threshold : threshold;
  const double pre = -1. / (4. * kPi);
  const double n13 = pre * (F_Nxz(rt, vp11, vp00, vp11) - F_Nxz(rt, 0., vp00, vp11) - (G_Nxz(rt, vp11) - G_Nxz(rt, 0.))) -
                     pre * (F_Nxz(rt, vp11, vp20, vp11) - F_Nxz(rt, 0., vp20, vp11) - (G_Nxz(rt, vp11) - G_Nxz(rt, 0.)));
  const double n23 = pre * (K_Nyz(rt, vp00, vp00, vp11) - K_Nyz(rt, 0., vp00, vp11) - (L_Nyz(rt, vp00) - L_Nyz(rt, 0.))) -
                     pre * (K_Nyz(rt, vp20, vp20, vp11) - K_Nyz(rt, 0., vp20, vp11) - (L_Nyz(rt, vp20) - L_Nyz(rt, 0.)));
  const double n33 = pre * (P_Nzz(rt, vp00, vp00, vp11) - P_Nzz(rt, 0., vp00, vp11) - (Q_Nzz(rt, vp00) - Q_Nzz(rt, 0.))) -
                     pre * (P_Nzz(rt, vp20, vp20, vp11) - P_Nzz(rt, 0., vp20, vp11) - (Q_Nzz(rt, vp20) - Q_Nzz(rt, 0.)));
  // N += (P N_loc) P^T with N_loc(:,3) = (n13, n23, n33)
  double pn[3];
  for (int a2 = 0; a2 < 3; ++a2) pn[a2] = e1[a2] * n13 + e2[a2] * n23 + e3[a2] * n33;
  for (int a2 = 0; a2 < 3; ++a2)
    for (int b2 = 0; b2 < 3; ++b2) N[a2 * 3 + b2] = N[a2 * 3 + b2] + pn[a2] * e3[b2];
  return true;
}
}  // namespace tet

// getN_tensor_tetrahedron: faces cshift(vert, i-1, 2), i = 1..4
__device__ bool tetra_n(const double* __restrict__ vert, const double r[3], double N[9]) {
  for (int q = 0; q < 9; ++q) N[q] = 0.0;
  for (int i = 0; i < 4; ++i) {
    double v[4][3];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) v[k][c] = vert[((k + i) & 3) * 3 + c];
    if (!tet::triangle_add(v, r, N)) return false;
  }
  return true;
}

// ---- magnetostatic tiles (getNFromTiles / getHFromTiles) ---------------------
// Tile record (kTileRec = 48 doubles): abc[3] off[3] R[9] Rinv[9] RM[3] ops[3]
// sym type vert[12], with R/Rinv of getRotationMatrices
// (DemagFieldGetSolution.f90:925-948), RM = R M (the rotated magnetisation of
// getFieldFromTile, :718), type 2 (prism) or 5 (tetrahedron, vertices v_k at
// vert[3k..3k+2]).

// Tensor of one tile at point (px,py,pz): getFieldFromTile (:685-730), or the
// eight mirrored evaluations of getFieldFromTile_symm (:744-822).  The symmetry
// matrices of getSymmOpMatrices (:833-896) are a signed identity (for M) and a
// diagonal mirror (for H and the point), applied here as sign products.
__device__ void tile_tensor(const double* __restrict__ T, double px, double py, double pz, double N[9]) {
  const double* R = T + 6;
  const bool tetra = T[31] == 5.0;
  int sym = T[30] != 0.0;
  int cases = sym ? 8 : 1;
  double o0 = T[27], o1 = T[28], o2 = T[29];
  // c[j]: sign applied to M; (sx,sy,sz)[j]: mirror applied to the point and rows
  const double cs[8] = {1.0, o2, o1, o0, o1 * o2, o0 * o2, o0 * o1, o2 * (o1 * o0)};
  const signed char mir[8][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1},
                                 {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
  for (int q = 0; q < 9; ++q) N[q] = 0.0;
  for (int j = 0; j < cases; ++j) {
    double s0 = mir[j][0], s1 = mir[j][1], s2 = mir[j][2];
    double dx = (sym ? s0 * px : px) - T[3];
    double dy = (sym ? s1 * py : py) - T[4];
    double dz = (sym ? s2 * pz : pz) - T[5];
    double rx = R[0] * dx + R[1] * dy + R[2] * dz;
    double ry = R[3] * dx + R[4] * dy + R[5] * dz;
    double rz = R[6] * dx + R[7] * dy + R[8] * dz;
    double full[9];
    if (tetra) {
      const double r3[3] = {rx, ry, rz};
      if (!tetra_n(T + 32, r3, full))
        for (int q = 0; q < 9; ++q) full[q] = __builtin_nan("");
    } else {
      double k[6];
      prism_k6(T[0], T[1], T[2], rx, ry, rz, k);
      full[0] = k[0]; full[1] = k[1]; full[2] = k[2];
      full[3] = k[1]; full[4] = k[3]; full[5] = k[4];
      full[6] = k[2]; full[7] = k[4]; full[8] = k[5];
    }
    if (!sym) {
      for (int q = 0; q < 9; ++q) N[q] = full[q];
    } else {
      double srow[3] = {s0, s1, s2};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) N[a * 3 + b] = srow[a] * (full[a * 3 + b] * cs[j]) + N[a * 3 + b];
    }
  }
}

// out: N(n_tiles, n_pts, 3, 3) in Fortran order, (i,p,a,b) at i + nt*(p + np*(a + 3b)).
__global__ void __launch_bounds__(256) tiles_tensor_kernel(const double* __restrict__ tiles, int nt,
                                                           const double* __restrict__ pts, int npt,
                                                           double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const double* T = tiles + (size_t)i * kTileRec;
  for (int p = blockIdx.y; p < npt; p += gridDim.y) {
    double N[9];
    tile_tensor(T, pts[p], pts[p + (size_t)npt], pts[p + 2 * (size_t)npt], N);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) out[i + (size_t)nt * (p + (size_t)npt * (a + 3 * b))] = N[a * 3 + b];
  }
}

// H(p) = sum_i Rinv_i (N_ip (R_i M_i)) (getDotProd, :909-918, and the
// back-rotation, :721-726).  One block per point, threads stride over tiles,
// fixed-order tree reduction.  Nin (Fortran-ordered as above) replaces the
// tensor evaluation when given (useStoredN).
__global__ void __launch_bounds__(256) tiles_field_kernel(const double* __restrict__ tiles, int nt,
                                                          const double* __restrict__ pts, int npt,
                                                          const double* __restrict__ Nin, double* __restrict__ H) {
  __shared__ double red[3][256];
  for (int p = blockIdx.x; p < npt; p += gridDim.x) {
    double px = pts[p], py = pts[p + (size_t)npt], pz = pts[p + 2 * (size_t)npt];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
      const double* T = tiles + (size_t)i * kTileRec;
      double N[9];
      if (Nin) {
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) N[a * 3 + b] = Nin[i + (size_t)nt * (p + (size_t)npt * (a + 3 * b))];
      } else {
        tile_tensor(T, px, py, pz, N);
      }
      const double* Ri = T + 15;
      const double* RM = T + 24;
      double d0 = N[0] * RM[0] + N[1] * RM[1] + N[2] * RM[2];
      double d1 = N[3] * RM[0] + N[4] * RM[1] + N[5] * RM[2];
      double d2 = N[6] * RM[0] + N[7] * RM[1] + N[8] * RM[2];
      acc[0] += Ri[0] * d0 + Ri[1] * d1 + Ri[2] * d2;
      acc[1] += Ri[3] * d0 + Ri[4] * d1 + Ri[5] * d2;
      acc[2] += Ri[6] * d0 + Ri[7] * d1 + Ri[8] * d2;
    }
    for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      for (int a = 0; a < 3; ++a) H[p + (size_t)npt * a] = red[a][0];
    __syncthreads();
  }
}

}  // namespace

// Host launchers -------------------------------------------------------------

int launch_tiles_tensor(const double* tiles, int nt, const double* pts, int npt, double* out, hipStream_t stream) {
  if (nt <= 0 || npt <= 0) return kOk;
  dim3 grid((unsigned)((nt + 255) / 256), (unsigned)(npt < 65535 ? npt : 65535));
  hipLaunchKernelGGL(tiles_tensor_kernel, grid, dim3(256), 0, stream, tiles, nt, pts, npt, out);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_tiles_field(const double* tiles, int nt, const double* pts, int npt, const double* Nin, double* H,
                       hipStream_t stream) {
  if (npt <= 0) return kOk;
  unsigned blocks = (unsigned)(npt < 65535 ? npt : 65535);
  hipLaunchKernelGGL(tiles_field_kernel, dim3(blocks), dim3(256), 0, stream, tiles, nt, pts, npt, Nin, H);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_assemble(const double* src, const double* abc, int abc_stride, const double* rot, const double* pts,
                    int n_src, int n_pts, float* out32, double* out64, uint8_t* split, hipStream_t stream) {
  int64_t np = round_up(n_src, kTileM);
  int64_t groups_per_row = (split ? np : round_up(n_src, 8)) / 8;
  int64_t rows = split ? round_up(n_pts, kTileM) : n_pts;
  int64_t threads = rows * groups_per_row;
  int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(assemble_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, abc, abc_stride, rot, pts,
                     n_src, n_pts, np, out32, out64, split);
  MT_HIP(hipGetLastError());
  return kOk;
}

namespace {
// Unstructured prisms (LandauLifshitzEquationSolver.f90:1044-1133): prism j
// (edges abc_j, centre p_j) at the nave sub-points (p_i - abc_i/2) +
// (abc_i/(nave+1)) * (ia, ja, ka) of receiver i, summed in FP64 (ka outer, ia
// inner) and divided by their number; out32[c*n*n + i*n + j].  One thread per
// pair, consecutive threads along j (coalesced stores).
__global__ void __launch_bounds__(256) assemble_avg_kernel(const double* __restrict__ pts,
                                                           const double* __restrict__ abc, int n, int nax, int nay,
                                                           int naz, float* __restrict__ out32,
                                                           double* __restrict__ out64) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  if (tid >= nn) return;
  const int i = (int)(tid / n), j = (int)(tid % n);
  const double ax = abc[3 * i], ay = abc[3 * i + 1], az = abc[3 * i + 2];
  const double dx = ax / (nax + 1), dy = ay / (nay + 1), dz = az / (naz + 1);
  const double bx = pts[3 * i] - ax / 2, by = pts[3 * i + 1] - ay / 2, bz = pts[3 * i + 2] - az / 2;
  const double cx = pts[3 * j], cy = pts[3 * j + 1], cz = pts[3 * j + 2];
  const double* a = abc + 3 * j;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int ka = 1; ka <= naz; ++ka)
    for (int ja = 1; ja <= nay; ++ja)
      for (int ia = 1; ia <= nax; ++ia) {
        const double px = bx + dx * ia, py = by + dy * ja, pz = bz + dz * ka;
        double k6[6];
        prism_k6(a[0], a[1], a[2], px - cx, py - cy, pz - cz, k6);
        for (int c = 0; c < 6; ++c) acc[c] += k6[c];
      }
  const double cnt = (double)(nax * nay * naz);
  for (int c = 0; c < 6; ++c) {
    const double v = acc[c] / cnt;
    if (out32) out32[c * nn + tid] = (float)v;
    if (out64) out64[c * nn + tid] = v;
  }
}
}  // namespace

namespace {
// Tetrahedral grid (LandauLifshitzEquationSolver.f90:988-1041): element j
// (nodes elem[4j..4j+3], 0-based here) at evaluation point i, no offset or
// rotation; stored components N(1,1) N(1,2) N(1,3) N(2,2) N(2,3) N(3,3).
__global__ void __launch_bounds__(256) assemble_tetra_kernel(const double* __restrict__ pts,
                                                             const double* __restrict__ nodes,
                                                             const int* __restrict__ elem, int n,
                                                             float* __restrict__ out32, double* __restrict__ out64) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  if (tid >= nn) return;
  const int i = (int)(tid / n), j = (int)(tid % n);
  double v[12];
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) v[3 * k + c] = nodes[(int64_t)elem[4 * j + k] * 3 + c];
  const double r[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  double N[9];
  if (!tetra_n(v, r, N))
    for (int q = 0; q < 9; ++q) N[q] = __builtin_nan("");
  const double k6[6] = {N[0], N[1], N[2], N[4], N[5], N[8]};
  for (int c = 0; c < 6; ++c) {
    if (out32) out32[c * nn + tid] = (float)k6[c];
    if (out64) out64[c * nn + tid] = k6[c];
  }
}
}  // namespace

int launch_assemble_tetra(const double* pts, const double* nodes, const int* elem, int n, float* out32,
                          double* out64, hipStream_t stream) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(assemble_tetra_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, stream, pts, nodes,
                     elem, n, out32, out64);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_assemble_avg(const double* pts, const double* abc, int n, const int nave[3], float* out32, double* out64,
                        hipStream_t stream) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(assemble_avg_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, stream, pts, abc, n,
                     nave[0], nave[1], nave[2], out32, out64);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_toeplitz_assemble(int nx, int ny, int nz, double Lx, double Ly, double Lz, uint8_t* out,
                             hipStream_t stream) {
  if (nx % 32 != 0) {
    set_error("the translation-invariant tensor needs nx % 32 == 0");
    return kErrArg;
  }
  const int64_t threads = toeplitz_fragments(nx, ny, nz) / 12 * 64;
  hipLaunchKernelGGL(toeplitz_assemble_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream, nx, ny,
                     nz, Lx, Ly, Lz, out);
  MT_HIP(hipGetLastError());
  return kOk;
}

namespace {
// Near-band table of a uniform grid (one thread per source offset): for the
// offset (ox, oy, oz) = receiver - source cell index, |ox|, |oy| <= R, |oz| <=
// Rz, the FP32 tensor value K32 of every component, packed by
// near_pack_offset into the split as the batched product uses it (khi =
// bf16(K32), klo = bf16(K32 - khi)) and the BF16 rounding of the rest dK.
// Values with |K32| <= thr (a threshold-sparse tensor's dropped entries) are
// zero.  The offset is formed from the centres of one representative pair, as
// the translation-invariant table forms it.
__global__ void near_table_kernel(int nx, int ny, int nz, double Lx, double Ly, double Lz, int R, int Rz, float thr,
                                  uint32_t* __restrict__ out) {
  const int w = 2 * R + 1;
  const int no = w * w * (2 * Rz + 1);
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= no) return;
  const int ox = o % w - R, oy = (o / w) % w - R, oz = o / (w * w) - Rz;
  const double dx = nx > 1 ? Lx / nx : Lx, dy = ny > 1 ? Ly / ny : Ly, dz = nz > 1 ? Lz / nz : Lz;
  auto centre = [](int m, double L, double d, int i) { return m > 1 ? -L / 2 + i * d + d / 2 : 0.0; };
  const int ix = ox > 0 ? ox : 0, jx = ox > 0 ? 0 : -ox;
  const int iy = oy > 0 ? oy : 0, jy = oy > 0 ? 0 : -oy;
  const int iz = oz > 0 ? oz : 0, jz = oz > 0 ? 0 : -oz;
  double k6[6];
  prism_k6(dx, dy, dz, centre(nx, Lx, dx, ix) - centre(nx, Lx, dx, jx), centre(ny, Ly, dy, iy) - centre(ny, Ly, dy, jy),
           centre(nz, Lz, dz, iz) - centre(nz, Lz, dz, jz), k6);
  float k32[6];
  for (int c = 0; c < 6; ++c) {
    k32[c] = (float)k6[c];
    if (!(fabsf(k32[c]) > thr)) k32[c] = 0.0f;  // an entry the threshold drops stays dropped
  }
  near_pack_offset(k32, out + (int64_t)o * kNearWords);
}
}  // namespace

int launch_near_table(int nx, int ny, int nz, double Lx, double Ly, double Lz, int R, int Rz, float thr,
                      uint32_t* table, hipStream_t stream) {
  if (R < 0 || Rz < 0) {
    set_error("near-band radius must be >= 0");
    return kErrArg;
  }
  const int no = (2 * R + 1) * (2 * R + 1) * (2 * Rz + 1);
  hipLaunchKernelGGL(near_table_kernel, dim3((unsigned)((no + 63) / 64)), dim3(64), 0, stream, nx, ny, nz, Lx, Ly, Lz,
                     R, Rz, thr, table);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_f32_fragments(const float* K, int n, uint8_t* frags, hipStream_t stream) {
  const int64_t np = round_up(n, kTileM);
  const int64_t threads = (np / kTileM) * (np / kTileK) * 64;  // one wave per (row tile, chunk)
  hipLaunchKernelGGL(f32_frag_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream, K, n, np, frags);
  MT_HIP(hipGetLastError());
  return kOk;
}

int launch_split(const float* K, int n, uint8_t* split, hipStream_t stream) {
  int64_t np = round_up(n, kTileM);
  int64_t threads = np * (np / 8);
  int64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(split_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, K, n, np, split);
  MT_HIP(hipGetLastError());
  return kOk;
}

}  // namespace mt
