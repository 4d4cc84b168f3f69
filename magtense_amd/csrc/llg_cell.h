// The Landau-Lifshitz right-hand side of one cell, shared by the LL kernel and
// the kernels that fuse it (llg.hip: FP32 demag rows + LL; demag.hip: the
// split product's near-band correction + LL), so that all of them form dm/dt
// with the same arithmetic in the same order.
//
// Reference: dmdt_fct, LandauLifshitzEquationSolver.f90:220-273; exchange
// :440-474 (Laplacian with the Neumann rows of :1194-1522); external field
// :485-514; anisotropy :523-548; demag scaling :749-754.  All effective fields
// are kept negated as in the reference; the API's alpha is alpha_LL * gamma.
#pragma once

#include "mt_common.h"
#include "mt_kernels.h"

namespace mt {

// LL right-hand side of cell i for the state of descriptor D.  hsum(c)
// returns the FP32 demag sum S_c = sum of the partials of (K m)_c at i (used
// unless D.hpart64 holds FP64 sums).
template <typename HS>
__device__ inline void llg_cell(const LlgGrid& g, const LlgDesc& D, int i, HS hsum) {
  const int n = g.n;
  const double* m = D.m;
  const int ix = i % g.nx;
  const int iy = (i / g.nx) % g.ny;
  const int iz = i / (g.nx * g.ny);
  const int sy = g.nx, sz = g.nx * g.ny;
  double mc[3], hj[3], hm[3], hk[3];
  if (g.xs) {
    // loaded operator: alpha = -2 times the row's stored entries in order
    // (updateExchangeTerms, LandauLifshitzEquationSolver.f90:440-474)
    const int kb = g.xs[i], ke = g.xe[i];
    for (int c = 0; c < 3; ++c) {
      const double* v = m + (size_t)c * n;
      mc[c] = v[i];
      double sum = 0.0;
      for (int k = kb; k < ke; ++k) sum += g.xv[k] * v[g.xc[k]];
      hj[c] = (-2.0 * sum) * g.jfact[i];
    }
  } else {
    // Laplacian with the reference's boundary rows: (-1, 1) and (1, -1).  The
    // neighbour indices are clamped to the cell itself at a boundary, so every
    // load is issued unconditionally before the per-row formula is selected.
    const int xm = ix > 0 ? i - 1 : i, xp = ix < g.nx - 1 ? i + 1 : i;
    const int ym = iy > 0 ? i - sy : i, yp = iy < g.ny - 1 ? i + sy : i;
    const bool z3 = g.nz > 1;
    const int zm = z3 && iz > 0 ? i - sz : i, zp = z3 && iz < g.nz - 1 ? i + sz : i;
    double c0[3], cxm[3], cxp[3], cym[3], cyp[3], czm[3], czp[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double* v = m + (size_t)c * n;
      c0[c] = v[i];
      cxm[c] = v[xm]; cxp[c] = v[xp];
      cym[c] = v[ym]; cyp[c] = v[yp];
      czm[c] = v[zm]; czp[c] = v[zp];
    }
    const double jf = g.jfact[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double vi = c0[c];
      mc[c] = vi;
      double lx = ix == 0 ? (cxp[c] - vi) : (ix == g.nx - 1 ? (cxm[c] - vi) : (cxm[c] - 2.0 * vi + cxp[c]));
      double ly = iy == 0 ? (cyp[c] - vi) : (iy == g.ny - 1 ? (cym[c] - vi) : (cym[c] - 2.0 * vi + cyp[c]));
      double lap = lx * g.idx2 + ly * g.idy2;
      if (z3) {
        double lz = iz == 0 ? (czp[c] - vi) : (iz == g.nz - 1 ? (czm[c] - vi) : (czm[c] - 2.0 * vi + czp[c]));
        lap += lz * g.idz2;
      }
      hj[c] = (-2.0 * lap) * jf;
    }
  }
  // demag: S = sum_k partial (K m); field = fl32(fl32(-S) * Ms_i)
  if (D.hpart64) {
    for (int c = 0; c < 3; ++c) hm[c] = (-D.hpart64[(int64_t)c * n + i]) * g.ms64[i];
  } else {
    for (int c = 0; c < 3; ++c) {
      const float s = hsum(c);
      float h = (-s) * g.ms32[i];
      hm[c] = (double)h;
    }
  }
  if (g.aniso) {
    const double* A = g.aniso + (size_t)i * 6;  // xx xy xz yy yz zz
    double f = -2.0 * g.kfact[i];
    hk[0] = f * (A[0] * mc[0] + A[1] * mc[1] + A[2] * mc[2]);
    hk[1] = f * (A[1] * mc[0] + A[3] * mc[1] + A[4] * mc[2]);
    hk[2] = f * (A[2] * mc[0] + A[4] * mc[1] + A[5] * mc[2]);
  } else {
    hk[0] = hk[1] = hk[2] = 0.0;
  }
  double hh[3] = {-D.hext[0], -D.hext[1], -D.hext[2]};
  double H[3];
  for (int c = 0; c < 3; ++c) H[c] = hh[c] + hj[c] + hm[c] + hk[c];
  double cx = -1.0 * (mc[1] * H[2] - mc[2] * H[1]);
  double cy = -1.0 * (mc[2] * H[0] - mc[0] * H[2]);
  double cz = -1.0 * (mc[0] * H[1] - mc[1] * H[0]);
  double dx = mc[1] * cz - mc[2] * cy;
  double dy = mc[2] * cx - mc[0] * cz;
  double dz = mc[0] * cy - mc[1] * cx;
  if (D.out) {
    D.out[i] = -D.gamma * cx - D.alpha * dx;
    D.out[n + i] = -D.gamma * cy - D.alpha * dy;
    D.out[2 * n + i] = -D.gamma * cz - D.alpha * dz;
  }
  if (D.fields) {
    double* F = D.fields;
    for (int c = 0; c < 3; ++c) {
      F[(size_t)c * n + i] = hj[c];
      F[(size_t)(3 + c) * n + i] = hh[c];
      F[(size_t)(6 + c) * n + i] = hm[c];
      F[(size_t)(9 + c) * n + i] = hk[c];
    }
  }
}

}  // namespace mt
