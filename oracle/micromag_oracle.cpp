// FP64 CPU oracle for the MagTense micromagnetic hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain serial restatement of the
// reference algorithm and is used as the checker: only tests/, the smoke test
// and the CPU-baseline leg of the benchmark may load it.  The product library
// (magtense_amd/csrc) never links or calls it.
//
// What it restates (reference paths relative to the MagTense source tree):
//   * prism tensor      source/TileDemagTensor/TileDemagTensor/TileNComponents.f90:629-724
//                       source/TileDemagTensor/TileDemagTensor/TileRectangularPrismTensor.f90:7-149
//   * tile/point loop   source/DemagField/DemagField/DemagFieldGetSolution.f90:685-730
//   * uniform grid      source/MagTenseMicroMag/LandauLifshitzEquationSolver.f90:815-872
//   * tensor assembly   LandauLifshitzEquationSolver.f90:883-986 (uniform), :1044-1133 (prisms)
//   * exchange CSR      LandauLifshitzEquationSolver.f90:1194-1522
//   * anisotropy        LandauLifshitzEquationSolver.f90:1566-1587
//   * field factors     LandauLifshitzEquationSolver.f90:359-430
//   * dm/dt             LandauLifshitzEquationSolver.f90:220-273, 440-548, 558-779
//   * interp1 / locate  source/NumericalIntegration/NumericalIntegration/util.f90:39-85
//   * ODE driver        source/NumericalIntegration/NumericalIntegration/ODE_Solvers.f90:124-237
//   * RKSUITE 'M'       source/NumericalIntegration/NumericalIntegration/RKSuitef90.f90
//                       (method_const :159-593, setup :596-856, range_integrate :942-1183,
//                        step_integrate :1185-1587, step :1726-1978, stiff :1981-2367,
//                        interpolate :2655-2997)
//   * field storage     LandauLifshitzEquationSolver.f90:306-349
//
// Conventions: every vector is SoA [mx(0:N); my; mz]; demag blocks are stored
// receiver-major: K[c][i*N + j] is component c (xx,xy,xz,yy,yz,zz) of the
// tensor at receiver i due to source prism j (the transpose of the reference's
// column-major Kxx(:,j) storage, identical for symmetric grids).
//
// Independent checks of every restated function (tests/test_oracle*.py,
// tests/test_adjoint_oracle.py; none retypes these formulas):
//   * prism tensor: cube self term -I/3, centre closed form, trace -1 / 0,
//     symmetries, far-field dipole, surface-charge quadrature (scipy), the
//     reference's COMSOL line profiles through the tile rotation;
//   * grid / layout: the reference's centre formula, block symmetry;
//   * exchange: zero row sums, symmetry, semi-definiteness, linear fields;
//   * anisotropy, field factors, dm/dt: closed forms, m . dm/dt = 0, the
//     damping and precession invariants;
//   * interp1 / locate: numpy interp / searchsorted on non-uniform tables,
//     endpoint and left-of-table rules;
//   * RKSUITE 'M': order conditions of the coefficient tables, convergence
//     orders, the interpolant, the work-limit return, the tolerance check;
//   * RK4 adjoint: central finite differences.
//
// Deviations (documented, all outside normal operation):
//   * interp1 clamps to the last table value when t lies right of the table
//     (the reference reads one element past the end there, util.f90:58).
//   * fatal RKSUITE conditions return an error code instead of STOPping.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

const double kPi = 3.141592653589793;      // TileTensorHelperFunctions.f90:5
const double kMu0 = 4.0 * kPi * 1e-7;

// ----------------------------------------------------------------------------
// Prism tensor (Smith et al. 2010, as used by MagTense)
// ----------------------------------------------------------------------------
const double kSclH = 1.0001, kSclL = 0.9999;  // TileRectangularPrismTensor.f90:11,48

double f3(double a, double b, double c, double x, double y, double z) {
  // TileRectangularPrismTensor.f90:45-70
  if (a / 2. - x == 0) {
    double xh = x * kSclH, xl = x * kSclL;
    double fh = (b / 2. - y) * (c / 2. - z) /
                ((a / 2. - xh) * std::sqrt((a / 2. - xh) * (a / 2. - xh) + (b / 2. - y) * (b / 2. - y) +
                                           (c / 2. - z) * (c / 2. - z)));
    double fl = (b / 2. - y) * (c / 2. - z) /
                ((a / 2. - xl) * std::sqrt((a / 2. - xl) * (a / 2. - xl) + (b / 2. - y) * (b / 2. - y) +
                                           (c / 2. - z) * (c / 2. - z)));
    return 0.5 * (fh + fl);
  }
  return (b / 2. - y) * (c / 2. - z) /
         ((a / 2. - x) *
          std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - y) * (b / 2. - y) + (c / 2. - z) * (c / 2. - z)));
}

double g3(double a, double b, double c, double x, double y, double z) {
  // TileRectangularPrismTensor.f90:74-93
  if (b / 2. - y == 0) {
    double yh = y * kSclH, yl = y * kSclL;
    double gh = (a / 2. - x) * (c / 2. - z) /
                ((b / 2. - yh) * std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - yh) * (b / 2. - yh) +
                                           (c / 2. - z) * (c / 2. - z)));
    double gl = (a / 2. - x) * (c / 2. - z) /
                ((b / 2. - yl) * std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - yl) * (b / 2. - yl) +
                                           (c / 2. - z) * (c / 2. - z)));
    return 0.5 * (gh + gl);
  }
  return (a / 2. - x) * (c / 2. - z) /
         ((b / 2. - y) *
          std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - y) * (b / 2. - y) + (c / 2. - z) * (c / 2. - z)));
}

double h3(double a, double b, double c, double x, double y, double z) {
  // TileRectangularPrismTensor.f90:97-116
  if (c / 2. - z == 0) {
    double zh = z * kSclH, zl = z * kSclL;
    double hh = (a / 2. - x) * (b / 2. - y) /
                ((c / 2. - zh) * std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - y) * (b / 2. - y) +
                                           (c / 2. - zh) * (c / 2. - zh)));
    double hl = (a / 2. - x) * (b / 2. - y) /
                ((c / 2. - zl) * std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - y) * (b / 2. - y) +
                                           (c / 2. - zl) * (c / 2. - zl)));
    return 0.5 * (hl + hh);
  }
  return (a / 2. - x) * (b / 2. - y) /
         ((c / 2. - z) *
          std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - y) * (b / 2. - y) + (c / 2. - z) * (c / 2. - z)));
}

double rr(double a, double b, double c, double x, double y, double z) {
  return std::sqrt((a / 2. - x) * (a / 2. - x) + (b / 2. - y) * (b / 2. - y) + (c / 2. - z) * (c / 2. - z));
}
// TileRectangularPrismTensor.f90:120-149
double FF3(double a, double b, double c, double x, double y, double z) { return (c / 2. - z) + rr(a, b, c, x, y, z); }
double GG3(double a, double b, double c, double x, double y, double z) { return (a / 2. - x) + rr(a, b, c, x, y, z); }
double HH3(double a, double b, double c, double x, double y, double z) { return (b / 2. - y) + rr(a, b, c, x, y, z); }

typedef double (*LogFn)(double, double, double, double, double, double);

double nom_of(LogFn F, double a, double b, double c, double x, double y, double z) {
  return F(a, b, c, x, y, z) * F(-a, -b, c, x, y, z) * F(a, -b, -c, x, y, z) * F(-a, b, -c, x, y, z);
}
double denom_of(LogFn F, double a, double b, double c, double x, double y, double z) {
  return F(a, -b, c, x, y, z) * F(-a, b, c, x, y, z) * F(a, b, -c, x, y, z) * F(-a, -b, -c, x, y, z);
}

double getF_limit(double a, double b, double c, double x, double y, double z, LogFn F) {
  // TileRectangularPrismTensor.f90:7-38
  double xh = x * kSclH, yh = y * kSclH, zh = z * kSclH;
  double xl = x * kSclL, yl = y * kSclL, zl = z * kSclL;
  double nom_l = nom_of(F, a, b, c, xl, yl, zl);
  double nom_h = nom_of(F, a, b, c, xh, yh, zh);
  double den_l = denom_of(F, a, b, c, xl, yl, zl);
  double den_h = denom_of(F, a, b, c, xh, yh, zh);
  double nom = 0.5 * (nom_l + nom_h);
  double den = 0.5 * (den_l + den_h);
  return nom / den;
}

double offdiag(LogFn F, double a, double b, double c, double x, double y, double z) {
  // TileNComponents.f90:665-716 (one off-diagonal component, before the final sign flip)
  double nom = nom_of(F, a, b, c, x, y, z);
  double den = denom_of(F, a, b, c, x, y, z);
  double v;
  if (den == 0 || nom == 0) {
    v = -1. / (4. * kPi) * std::log(getF_limit(a, b, c, x, y, z, F));
  } else {
    v = -1. / (4. * kPi) * std::log(nom / den);
  }
  if (std::fabs((nom - den) / den) < 10. * DBL_EPSILON) v = 0;
  return v;
}

typedef double (*AtanFn)(double, double, double, double, double, double);
double diag(AtanFn F, double a, double b, double c, double x, double y, double z) {
  // TileNComponents.f90:647-661 (the eight-term arctangent sum, summed in the reference's order)
  return 1. / (4. * kPi) *
         (std::atan(F(a, b, c, x, y, z)) + std::atan(F(a, b, c, -x, y, z)) + std::atan(F(a, b, c, x, -y, z)) +
          std::atan(F(a, b, c, x, y, -z)) + std::atan(F(a, b, c, -x, -y, z)) + std::atan(F(a, b, c, x, -y, -z)) +
          std::atan(F(a, b, c, -x, y, -z)) + std::atan(F(a, b, c, -x, -y, -z)));
}

// Returns the MagTense tensor K (= -N_standard) for a prism of full edges a,b,c
// evaluated at offset (x,y,z) from its centre.  Order: xx,xy,xz,yy,yz,zz.
void prism_k6(double a, double b, double c, double x, double y, double z, double out[6]) {
  double nxx = diag(f3, a, b, c, x, y, z);
  double nyy = diag(g3, a, b, c, x, y, z);
  double nzz = diag(h3, a, b, c, x, y, z);
  double nxy = offdiag(FF3, a, b, c, x, y, z);
  double nyz = offdiag(GG3, a, b, c, x, y, z);
  double nxz = offdiag(HH3, a, b, c, x, y, z);
  out[0] = -nxx;  // TileNComponents.f90:722
  out[1] = -nxy;
  out[2] = -nxz;
  out[3] = -nyy;
  out[4] = -nyz;
  out[5] = -nzz;
}

// ----------------------------------------------------------------------------
// Grid
// ----------------------------------------------------------------------------
struct Grid {
  int nx, ny, nz, n;
  double dx, dy, dz;
  std::vector<double> pts;  // n*3, row-major
};

void setup_grid(int nx, int ny, int nz, double Lx, double Ly, double Lz, Grid& g) {
  // LandauLifshitzEquationSolver.f90:815-872
  g.nx = nx; g.ny = ny; g.nz = nz; g.n = nx * ny * nz;
  g.pts.assign((size_t)g.n * 3, 0.0);
  std::vector<double> xs(nx), ys(ny), zs(nz);
  if (nx > 1) { g.dx = Lx / nx; for (int i = 0; i < nx; ++i) xs[i] = -Lx / 2 + i * g.dx + g.dx / 2; }
  else { g.dx = Lx; xs[0] = 0.; }
  if (ny > 1) { g.dy = Ly / ny; for (int i = 0; i < ny; ++i) ys[i] = -Ly / 2 + i * g.dy + g.dy / 2; }
  else { g.dy = Ly; ys[0] = 0.; }
  if (nz > 1) { g.dz = Lz / nz; for (int i = 0; i < nz; ++i) zs[i] = -Lz / 2 + i * g.dz + g.dz / 2; }
  else { g.dz = Lz; zs[0] = 0.; }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t ind = (size_t)i + (size_t)j * nx + (size_t)k * nx * ny;
        g.pts[ind * 3 + 0] = xs[i];
        g.pts[ind * 3 + 1] = ys[j];
        g.pts[ind * 3 + 2] = zs[k];
      }
}

// ----------------------------------------------------------------------------
// Exchange operator: the reference's CSR Laplacian with Neumann rows
// (LandauLifshitzEquationSolver.f90:1194-1522).  Stored row-wise with columns
// ascending, values already divided by d^2 and summed over directions.
// ----------------------------------------------------------------------------
struct Csr {
  int n;
  std::vector<int> rowptr, col;
  std::vector<double> val;
};

void build_exchange(const Grid& g, Csr& A) {
  int nx = g.nx, ny = g.ny, nz = g.nz, n = g.n;
  A.n = n;
  A.rowptr.assign(n + 1, 0);
  A.col.clear(); A.val.clear();
  double ix2 = 1. / (g.dx * g.dx), iy2 = 1. / (g.dy * g.dy), iz2 = 1. / (g.dz * g.dz);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int r = i + j * nx + k * nx * ny;
        // collect (col, val) per direction, then merge by column (mkl_sparse_d_add)
        std::vector<std::pair<int, double>> e;
        // x: boundary rows (-1, 1) / (1, -1), interior (1, -2, 1)
        if (i == 0) { e.push_back({r, -1. * ix2}); e.push_back({r + 1, 1. * ix2}); }
        else if (i == nx - 1) { e.push_back({r - 1, 1. * ix2}); e.push_back({r, -1. * ix2}); }
        else { e.push_back({r - 1, 1. * ix2}); e.push_back({r, -2. * ix2}); e.push_back({r + 1, 1. * ix2}); }
        // y
        if (j == 0) { e.push_back({r, -1. * iy2}); e.push_back({r + nx, 1. * iy2}); }
        else if (j == ny - 1) { e.push_back({r - nx, 1. * iy2}); e.push_back({r, -1. * iy2}); }
        else { e.push_back({r - nx, 1. * iy2}); e.push_back({r, -2. * iy2}); e.push_back({r + nx, 1. * iy2}); }
        // z only when nz > 1 (LandauLifshitzEquationSolver.f90:1386)
        if (nz > 1) {
          int s = nx * ny;
          if (k == 0) { e.push_back({r, -1. * iz2}); e.push_back({r + s, 1. * iz2}); }
          else if (k == nz - 1) { e.push_back({r - s, 1. * iz2}); e.push_back({r, -1. * iz2}); }
          else { e.push_back({r - s, 1. * iz2}); e.push_back({r, -2. * iz2}); e.push_back({r + s, 1. * iz2}); }
        }
        std::stable_sort(e.begin(), e.end(),
                         [](const std::pair<int, double>& p, const std::pair<int, double>& q) { return p.first < q.first; });
        for (size_t q = 0; q < e.size(); ++q) {
          if (!A.col.empty() && (int)A.col.size() > A.rowptr[r] && A.col.back() == e[q].first) {
            A.val.back() += e[q].second;
          } else {
            A.col.push_back(e[q].first);
            A.val.push_back(e[q].second);
          }
        }
        A.rowptr[r + 1] = (int)A.col.size();
      }
}

// ----------------------------------------------------------------------------
// interp1_MagTense / locate  (util.f90:39-85)
// ----------------------------------------------------------------------------
int locate(const double* xx, int n, double x) {
  int jl = 0, ju = n + 1;
  while (ju - jl > 1) {
    int jm = (ju + jl) / 2;
    if ((xx[n - 1] >= xx[0]) == (x >= xx[jm - 1])) jl = jm; else ju = jm;
  }
  if (x == xx[0]) return 1;
  if (x == xx[n - 1]) return n - 1;
  return jl;
}

// Contiguous-abscissa version used by the solver.
double interp1c(const std::vector<double>& xs, const std::vector<double>& ys, double x) {
  int n = (int)xs.size();
  int j = locate(xs.data(), n, x);
  if (j == 0) return 0.0;
  if (j >= n) return ys[n - 1];  // deviation: clamp (reference reads past the table end)
  double xl = (x - xs[j - 1]) / (xs[j] - xs[j - 1]);
  return (1 - xl) * ys[j - 1] + xl * ys[j];
}

// ----------------------------------------------------------------------------
// Micromagnetic problem and dm/dt
// ----------------------------------------------------------------------------
struct Problem {
  Grid g;
  Csr Aex;
  int n;
  std::vector<float> K;      // 6*N*N FP32 (receiver-major)
  std::vector<double> Kd;    // 6*N*N FP64 when demag_fp64
  bool demag_fp64 = false;
  std::vector<double> Axx, Axy, Axz, Ayy, Ayz, Azz;
  std::vector<double> Jfact, Kfact;
  std::vector<float> Mfact;
  std::vector<double> Ms64;  // FP64 Ms: the demag scaling of the FP64 mode
  double gamma, alpha0;
  std::vector<double> alpha_t, alpha_v;
  int solver;
  std::vector<double> hext_t, hext_x, hext_y, hext_z;  // dynamic table
  int hext_ind = 0;                                      // explicit mode row
  // scratch
  std::vector<double> Hj, Hh, Hk;
  std::vector<float> Hm, ms;
  long long rhs_calls = 0;
  float demag_threshold = 0.0f;  // threshold applied to K (sparse modes)
};

double alpha_at(const Problem& P, double t) {
  // LandauLifshitzEquationSolver.f90:281-294
  if (P.alpha0 == 0) return interp1c(P.alpha_t, P.alpha_v, t);
  return P.alpha0;
}

void update_fields(Problem& P, const double* m, double t) {
  int n = P.n;
  const double* mx = m; const double* my = m + n; const double* mz = m + 2 * n;
  // exchange: LandauLifshitzEquationSolver.f90:440-474
  for (int c = 0; c < 3; ++c) {
    const double* v = m + (size_t)c * n;
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int q = P.Aex.rowptr[r]; q < P.Aex.rowptr[r + 1]; ++q) s += P.Aex.val[q] * v[P.Aex.col[q]];
      double temp = -2. * s;
      P.Hj[(size_t)c * n + r] = temp * P.Jfact[r];
    }
  }
  // external field: LandauLifshitzEquationSolver.f90:485-514
  double hx, hy, hz;
  if (P.solver == 1) {
    hx = P.hext_x[P.hext_ind]; hy = P.hext_y[P.hext_ind]; hz = P.hext_z[P.hext_ind];
  } else {
    hx = interp1c(P.hext_t, P.hext_x, t);
    hy = interp1c(P.hext_t, P.hext_y, t);
    hz = interp1c(P.hext_t, P.hext_z, t);
  }
  for (int r = 0; r < n; ++r) { P.Hh[r] = -hx; P.Hh[n + r] = -hy; P.Hh[2 * n + r] = -hz; }
  // anisotropy: LandauLifshitzEquationSolver.f90:523-548
  for (int r = 0; r < n; ++r) {
    double f = -2. * P.Kfact[r];
    P.Hk[r] = f * (P.Axx[r] * mx[r] + P.Axy[r] * my[r] + P.Axz[r] * mz[r]);
    P.Hk[n + r] = f * (P.Axy[r] * mx[r] + P.Ayy[r] * my[r] + P.Ayz[r] * mz[r]);
    P.Hk[2 * n + r] = f * (P.Axz[r] * mx[r] + P.Ayz[r] * my[r] + P.Azz[r] * mz[r]);
  }
  // demag: LandauLifshitzEquationSolver.f90:558-579, 706-755.
  // FP32 semantics: m cast to FP32; each of the nine gemv steps is the
  // correctly rounded FP32 value of (beta*H - K m) (exact products, FP64 sums);
  // the result is scaled by the receiver cell's FP32 Ms.
  for (int r = 0; r < 3 * n; ++r) P.ms[r] = (float)m[r];
  static const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  size_t nn = (size_t)n * n;
  if (P.demag_fp64) return;  // FP64 product is formed by demag_fp64()
  for (int a = 0; a < 3; ++a) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      float h = 0.0f;
      for (int b = 0; b < 3; ++b) {
        const float* Krow = P.K.data() + comp[a][b] * nn + (size_t)i * n;
        const float* v = P.ms.data() + (size_t)b * n;
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += (double)Krow[j] * (double)v[j];
        h = (float)((double)h - s);
      }
      P.Hm[(size_t)a * n + i] = h * P.Mfact[i];
    }
  }
}

// FP64 demag field (used when demag_fp64): H = -Ms_i * sum_j K m, all in FP64
void demag_fp64(const Problem& P, const double* m, double* H) {
  int n = P.n;
  size_t nn = (size_t)n * n;
  static const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  for (int a = 0; a < 3; ++a)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int b = 0; b < 3; ++b) {
        const double* Krow = P.Kd.data() + comp[a][b] * nn + (size_t)i * n;
        const double* v = m + (size_t)b * n;
        for (int j = 0; j < n; ++j) s += Krow[j] * v[j];
      }
      H[(size_t)a * n + i] = -s * P.Ms64[i];
    }
}

void dmdt(Problem& P, double t, const double* m, double* out, std::vector<double>& Hm64) {
  // LandauLifshitzEquationSolver.f90:220-273
  int n = P.n;
  ++P.rhs_calls;
  update_fields(P, m, t);
  if (P.demag_fp64) demag_fp64(P, m, Hm64.data());
  double al = alpha_at(P, t);
  for (int r = 0; r < n; ++r) {
    double mx = m[r], my = m[n + r], mz = m[2 * n + r];
    double hmx = P.demag_fp64 ? Hm64[r] : (double)P.Hm[r];
    double hmy = P.demag_fp64 ? Hm64[n + r] : (double)P.Hm[n + r];
    double hmz = P.demag_fp64 ? Hm64[2 * n + r] : (double)P.Hm[2 * n + r];
    double Hx = P.Hh[r] + P.Hj[r] + hmx + P.Hk[r];
    double Hy = P.Hh[n + r] + P.Hj[n + r] + hmy + P.Hk[n + r];
    double Hz = P.Hh[2 * n + r] + P.Hj[2 * n + r] + hmz + P.Hk[2 * n + r];
    double cx = -1. * (my * Hz - mz * Hy);
    double cy = -1. * (mz * Hx - mx * Hz);
    double cz = -1. * (mx * Hy - my * Hx);
    double dx = my * cz - mz * cy;
    double dy = mz * cx - mx * cz;
    double dz = mx * cy - my * cx;
    out[r] = -P.gamma * cx - al * dx;
    out[n + r] = -P.gamma * cy - al * dy;
    out[2 * n + r] = -P.gamma * cz - al * dz;
  }
}

// ----------------------------------------------------------------------------
// RKSUITE method 'M' (Bogacki-Shampine 4(5) FSAL pair) in range mode.
// ----------------------------------------------------------------------------
typedef void (*RhsFn)(void* ctx, double t, const double* y, double* dydt);

struct RkStats {
  long long f_calls;        // total RHS evaluations (primary integration)
  long long steps;          // accepted steps
  long long bad_steps;      // rejected attempts after the first step
  long long flag3;          // "too much work" returns
  long long flag4;          // stiffness returns
  double h_last;
};

struct RkComm {
  int n;
  RhsFn f; void* ctx;
  double t, t_old, t_start, t_end, dir;
  double h, h_old, h_start, h_average;
  double tol;
  long long f_count, full_f_count, step_count, bad_step_count;
  bool at_t_start, at_t_end;
  std::vector<double> thresh, weights, ymax;
  std::vector<double> scratch, y, yp, y_old, err_estimates;  // y_new == yp_old == vtemp == scratch
  std::vector<double> stages;  // 7 * n (ptr(2..8) = 1..7)
  double a[13][13], b[13], c[13], r[11][6], e[7];
  double cost, safety, expon, stability_radius, tan_angle, rs, rs1, rs2, rs3, rs4, toosml;
  int order, last_stage, max_stiff_iters;
  double round_off, sqrrmc, cubrmc, sqtiny;
  double tlast, range_t_end;
  std::vector<double> xstage, ytemp, p;  // p: 5 * n
  long long stiff_bad_step_count, hit_t_end_count;
  double errold;
  bool chkeff, phase2;
  int st_range, st_step, st_interp;  // saved states (usable = -2)
  long long flag3, flag4;
  double* stage(int k) { return stages.data() + (size_t)(k - 1) * n; }  // k = 1..7
};

const int kUsable = -2, kFatal = 911;

void method_const_M(RkComm& C) {
  // RKSuitef90.f90:225-412 (METHD = 2) and :560-593
  std::memset(C.a, 0, sizeof(C.a)); std::memset(C.b, 0, sizeof(C.b)); std::memset(C.c, 0, sizeof(C.c));
  std::memset(C.r, 0, sizeof(C.r)); std::memset(C.e, 0, sizeof(C.e));
  double (*a)[13] = C.a;
  // a(i,j) stored at a[i-1][j-1]
  a[1][0] = 1.0 / 6.0;
  a[2][0] = 2.0 / 27.0; a[2][1] = 4.0 / 27.0;
  a[3][0] = 183.0 / 1372.0; a[3][1] = -162.0 / 343.0; a[3][2] = 1053.0 / 1372.0;
  a[4][0] = 68.0 / 297.0; a[4][1] = -4.0 / 11.0; a[4][2] = 42.0 / 143.0; a[4][3] = 1960.0 / 3861.0;
  a[5][0] = 597.0 / 22528.0; a[5][1] = 81.0 / 352.0; a[5][2] = 63099.0 / 585728.0; a[5][3] = 58653.0 / 366080.0;
  a[5][4] = 4617.0 / 20480.0;
  a[6][0] = 174197.0 / 959244.0; a[6][1] = -30942.0 / 79937.0; a[6][2] = 8152137.0 / 19744439.0;
  a[6][3] = 666106.0 / 1039181.0; a[6][4] = -29421.0 / 29068.0; a[6][5] = 482048.0 / 414219.0;
  a[7][0] = 587.0 / 8064.0; a[7][1] = 0.0; a[7][2] = 4440339.0 / 15491840.0; a[7][3] = 24353.0 / 124800.0;
  a[7][4] = 387.0 / 44800.0; a[7][5] = 2152.0 / 5985.0; a[7][6] = 7267.0 / 94080.0;
  double* b = C.b;
  b[0] = 2479.0 / 34992.0; b[1] = 0.0; b[2] = 123.0 / 416.0; b[3] = 612941.0 / 3411720.0; b[4] = 43.0 / 1440.0;
  b[5] = 2272.0 / 6561.0; b[6] = 79937.0 / 1113912.0; b[7] = 3293.0 / 556956.0;
  double* e = C.e;
  e[0] = -3.0 / 1280.0; e[1] = 0.0; e[2] = 6561.0 / 632320.0; e[3] = -343.0 / 20800.0; e[4] = 243.0 / 12800.0;
  e[5] = -1.0 / 95.0; e[6] = 0.0;
  double* c = C.c;
  c[0] = 0.0; c[1] = 1.0 / 6.0; c[2] = 2.0 / 9.0; c[3] = 3.0 / 7.0; c[4] = 2.0 / 3.0; c[5] = 3.0 / 4.0; c[6] = 1.0;
  c[7] = 1.0; c[8] = 1.0 / 2.0; c[9] = 5.0 / 6.0; c[10] = 1.0 / 9.0;
  a[8][0] = 455.0 / 6144.0; a[9][0] = -837888343715.0 / 13176988637184.0; a[10][0] = 98719073263.0 / 1551965184000.0;
  a[8][1] = 0.0; a[9][1] = 30409415.0 / 52955362.0; a[10][1] = 1307.0 / 123552.0;
  a[8][2] = 10256301.0 / 35409920.0; a[9][2] = -48321525963.0 / 759168069632.0;
  a[10][2] = 4632066559387.0 / 70181753241600.0;
  a[8][3] = 2307361.0 / 17971200.0; a[9][3] = 8530738453321.0 / 197654829557760.0;
  a[10][3] = 7828594302389.0 / 382182512025600.0;
  a[8][4] = -387.0 / 102400.0; a[9][4] = 1361640523001.0 / 1626788720640.0; a[10][4] = 40763687.0 / 11070259200.0;
  a[8][5] = 73.0 / 5130.0; a[9][5] = -13143060689.0 / 38604458898.0; a[10][5] = 34872732407.0 / 224610586200.0;
  a[8][6] = -7267.0 / 215040.0; a[9][6] = 18700221969.0 / 379584034816.0; a[10][6] = -2561897.0 / 30105600.0;
  a[8][7] = 1.0 / 32.0; a[9][7] = -5831595.0 / 847285792.0; a[10][7] = 1.0 / 10.0;
  a[9][8] = -5183640.0 / 26477681.0; a[10][8] = -1.0 / 10.0;
  a[10][9] = -1403317093.0 / 11371610250.0;
  // r(i,j) at r[i-1][j-1]
  double (*r)[6] = C.r;
  r[0][5] = -12134338393.0 / 1050809760.0; r[0][4] = -1620741229.0 / 50038560.0;
  r[0][3] = -2048058893.0 / 59875200.0; r[0][2] = -87098480009.0 / 5254048800.0;
  r[0][1] = -11513270273.0 / 3502699200.0;
  r[2][5] = -33197340367.0 / 1218433216.0; r[2][4] = -539868024987.0 / 6092166080.0;
  r[2][3] = -39991188681.0 / 374902528.0; r[2][2] = -69509738227.0 / 1218433216.0;
  r[2][1] = -29327744613.0 / 2436866432.0;
  r[3][5] = -284800997201.0 / 19905339168.0; r[3][4] = -7896875450471.0 / 165877826400.0;
  r[3][3] = -333945812879.0 / 5671036800.0; r[3][2] = -16209923456237.0 / 497633479200.0;
  r[3][1] = -2382590741699.0 / 331755652800.0;
  r[4][5] = -540919.0 / 741312.0; r[4][4] = -103626067.0 / 43243200.0; r[4][3] = -633779.0 / 211200.0;
  r[4][2] = -32406787.0 / 18532800.0; r[4][1] = -36591193.0 / 86486400.0;
  r[5][5] = 7157998304.0 / 374350977.0; r[5][4] = 30405842464.0 / 623918295.0; r[5][3] = 183022264.0 / 5332635.0;
  r[5][2] = -3357024032.0 / 1871754885.0; r[5][1] = -611586736.0 / 89131185.0;
  r[6][5] = -138073.0 / 9408.0; r[6][4] = -719433.0 / 15680.0; r[6][3] = -1620541.0 / 31360.0;
  r[6][2] = -385151.0 / 15680.0; r[6][1] = -65403.0 / 15680.0;
  r[7][5] = 1245.0 / 64.0; r[7][4] = 3991.0 / 64.0; r[7][3] = 4715.0 / 64.0; r[7][2] = 2501.0 / 64.0;
  r[7][1] = 149.0 / 16.0; r[7][0] = 1.0;
  r[8][5] = 55.0 / 3.0; r[8][4] = 71.0; r[8][3] = 103.0; r[8][2] = 199.0 / 3.0; r[8][1] = 16.0;
  r[9][5] = -1774004627.0 / 75810735.0; r[9][4] = -1774004627.0 / 25270245.0; r[9][3] = -26477681.0 / 359975.0;
  r[9][2] = -11411880511.0 / 379053675.0; r[9][1] = -423642896.0 / 126351225.0;
  r[10][5] = 35.0; r[10][4] = 105.0; r[10][3] = 117.0; r[10][2] = 59.0; r[10][1] = 12.0;

  C.order = 4; C.tan_angle = 5.2; C.stability_radius = 3.9; C.safety = 0.8;
  C.last_stage = 7;                          // ptr(8)
  C.cost = 7.0;                              // fsal: no_of_stages - 1
  C.max_stiff_iters = std::min(8, std::max(1, (int)(0.05 * C.cost * 50.0)));
  C.expon = 1.0 / (C.order + 1.0);
  double cdiff = 1.0;
  for (int i = 0; i < 7; ++i)
    for (int j = i + 1; j < 8; ++j) {
      double d = c[i] - c[j];
      if (d != 0) cdiff = std::min(cdiff, std::fabs(d));
    }
  C.rs = 2.0; C.rs1 = 1.0 / C.rs; C.rs2 = C.rs * C.rs; C.rs3 = C.rs * C.rs2; C.rs4 = 1.0 / C.rs3;
  C.round_off = 10.0 * DBL_EPSILON;
  C.sqrrmc = std::sqrt(DBL_EPSILON);
  C.cubrmc = std::pow(DBL_EPSILON, 1.0 / 3.0);
  C.sqtiny = std::sqrt(DBL_MIN);
  C.toosml = C.round_off / cdiff;
}

// returns 0 or a negative error code
int rk_setup(RkComm& C, int n, RhsFn f, void* ctx, double t_start, const double* y_start, double t_end, double tol,
             const double* thres) {
  // RKSuitef90.f90:596-856 with method 'M', task 'R', errass false, h_start 0
  C.n = n; C.f = f; C.ctx = ctx;
  C.h_start = 0.0; C.h_old = 0.0; C.f_count = 0; C.full_f_count = 0; C.step_count = 0; C.bad_step_count = 0;
  C.at_t_start = true; C.at_t_end = false; C.stiff_bad_step_count = 0; C.hit_t_end_count = 0;
  C.errold = 0.0; C.h_average = 0.0; C.chkeff = false; C.phase2 = true;
  C.st_range = -1; C.st_step = -1; C.st_interp = -1;
  C.flag3 = 0; C.flag4 = 0;
  method_const_M(C);
  if (n <= 0) return -1;
  if (t_start == t_end) return -2;
  C.t_end = t_end; C.t_start = t_start; C.t_old = t_start; C.t = t_start;
  C.dir = (t_end - t_start) >= 0 ? 1.0 : -1.0;
  if (tol > 0.01 || tol < C.round_off) return -3;
  C.tol = tol;
  double mn = thres[0];
  for (int i = 1; i < n; ++i) mn = std::min(mn, thres[i]);
  if (mn < C.sqtiny) return -4;
  double hmin = std::max(C.sqtiny, C.toosml * std::max(std::fabs(t_start), std::fabs(t_end)));
  if (std::fabs(t_end - t_start) < hmin) return -5;
  C.h = std::max(std::fabs(t_end - t_start) / C.rs3, hmin);
  C.thresh.assign(thres, thres + n);
  C.err_estimates.assign(n, 0.0); C.weights.assign(n, 0.0); C.y_old.assign(n, 0.0);
  C.scratch.assign(n, 0.0); C.y.assign(y_start, y_start + n); C.yp.assign(n, 0.0);
  C.stages.assign((size_t)7 * n, 0.0); C.ymax.resize(n);
  for (int i = 0; i < n; ++i) C.ymax[i] = std::fabs(y_start[i]);
  C.p.assign((size_t)5 * n, 0.0); C.ytemp.assign(n, 0.0); C.xstage.assign(n, 0.0);
  C.st_range = C.st_range;  // SETUP ok
  return 0;
}

inline void call_f(RkComm& C, double t, const double* y, double* out) { C.f(C.ctx, t, y, out); }

// step_r1 (RKSuitef90.f90:1726-1978) for the primary integration.
void rk_step(RkComm& C, double tnow, double& htry, double& err, double hmin) {
  int n = C.n;
  double* y = C.y.data();
  double* yp = C.yp.data();
  double* y_new = C.scratch.data();
  double* errest = C.err_estimates.data();
  double* w = C.weights.data();
  const double* th = C.thresh.data();
  for (;;) {  // attempt_step
    if (C.phase2)
      for (int k = 0; k < n; ++k) w[k] = std::max(th[k], std::fabs(y[k]));
    bool restart = false;
    for (int i = 2; i <= 8; ++i) {
      double ai1 = C.a[i - 1][0];
      for (int k = 0; k < n; ++k) errest[k] = ai1 * yp[k];
      for (int j = 2; j <= i - 1; ++j) {
        double aij = C.a[i - 1][j - 1];
        if (aij != 0.0) {
          const double* s = C.stage(j - 1);
          for (int k = 0; k < n; ++k) errest[k] = errest[k] + aij * s[k];
        }
      }
      for (int k = 0; k < n; ++k) y_new[k] = y[k] + htry * errest[k];
      if (i == 7) {
        // stepb (RKSuitef90.f90:1957-1976); note it scales by |comm%h|, not htry
        double m = 0.0;
        const double* s3 = C.stage(2); const double* s4 = C.stage(3);
        const double* s5 = C.stage(4); const double* s6 = C.stage(5);
        for (int k = 0; k < n; ++k) {
          double v = C.e[0] * yp[k] + C.e[2] * s3[k] + C.e[3] * s4[k] + C.e[4] * s5[k] + C.e[5] * s6[k];
          double wt = std::max(0.5 * (std::fabs(y[k]) + std::fabs(y_new[k])), th[k]);
          m = std::max(m, std::fabs(v) / wt);
        }
        err = std::fabs(C.h) * m;
        if (err > C.tol) return;
      }
      double tstg = tnow + C.c[i - 1] * htry;
      if (C.at_t_end && C.c[i - 1] == 1.0) tstg = C.t_end;
      double* out = C.stage(i - 1);
      call_f(C, tstg, y_new, out);
      C.f_count += 1;
      if (C.phase2 && i <= 3 && std::fabs(htry) > hmin) {
        // stepa (RKSuitef90.f90:1876-1955)
        double ynrm = 0, ystgnm = 0, argdif = 0, fdiff = 0;
        for (int k = 0; k < n; ++k) {
          w[k] = std::max(w[k], std::fabs(y_new[k]));
          ynrm = std::max(ynrm, std::fabs(y[k]) / w[k]);
          ystgnm = std::max(ystgnm, std::fabs(y_new[k]) / w[k]);
          argdif = std::max(argdif, std::fabs(y_new[k] - y[k]) / w[k]);
          fdiff = std::max(fdiff, std::fabs(out[k] - yp[k]) / w[k]);
        }
        double tdiff = tstg - tnow;
        double twt = std::fabs(C.t_end - tnow);
        ynrm = std::max(ynrm, std::fabs(tnow) / twt);
        ystgnm = std::max(ystgnm, std::fabs(tstg) / twt);
        argdif = std::max(argdif, std::fabs(tdiff) / twt);
        bool cutbak = false;
        if (argdif > C.round_off * std::max(ynrm, ystgnm)) {
          if (std::fabs(htry) * fdiff > C.stability_radius * argdif) {
            double scl = std::max(C.rs4, std::min(C.rs1, (C.stability_radius * argdif) / (std::fabs(htry) * fdiff)));
            htry = scl * htry;
            cutbak = true;
          }
        }
        if (cutbak) {
          C.at_t_end = false;
          if (std::fabs(htry) <= hmin) {
            htry = std::copysign(hmin, htry);
            C.phase2 = false;
          }
          restart = true;
          break;
        }
      }
    }
    if (restart) continue;
    // error estimate (RKSuitef90.f90:1861-1874)
    if (C.b[0] != 0.0)
      for (int k = 0; k < n; ++k) errest[k] = errest[k] - C.b[0] * yp[k];
    for (int i = 2; i <= 8; ++i) {
      double bi = C.b[i - 1];
      if (bi != 0.0) {
        const double* s = C.stage(i - 1);
        for (int k = 0; k < n; ++k) errest[k] = errest[k] - bi * s[k];
      }
    }
    for (int k = 0; k < n; ++k) errest[k] = htry * errest[k];
    double m = 0.0;
    for (int k = 0; k < n; ++k) {
      w[k] = std::max(0.5 * (std::fabs(y[k]) + std::fabs(y_new[k])), th[k]);
      m = std::max(m, std::fabs(errest[k] / w[k]));
    }
    err = m;
    return;
  }
}

// stiff_r1 (RKSuitef90.f90:1981-2367)
void rk_stiff(RkComm& C, bool toomch, bool& sure_stiff) {
  int n = C.n;
  sure_stiff = false;
  bool lots_of_fails = false;
  if ((C.step_count - 10) % 40 == 0) {
    lots_of_fails = C.stiff_bad_step_count >= 10;
    C.stiff_bad_step_count = 0;
  }
  bool maybe_stiff = toomch || lots_of_fails;
  double* w = C.weights.data();
  double* y = C.y.data();
  double* v0 = C.err_estimates.data();
  double* v1 = C.stage(1); double* v2 = C.stage(2); double* v3 = C.stage(3);
  auto wdot = [&](const double* u, const double* v) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += (u[k] / w[k]) * (v[k] / w[k]);
    return s;
  };
  double v0v0 = 0, scale = 0;
  if (maybe_stiff) {
    for (int k = 0; k < n; ++k) w[k] = std::max(0.5 * (std::fabs(y[k]) + std::fabs(C.y_old[k])), C.thresh[k]);
    double q = std::fabs(C.h / C.h_average);
    maybe_stiff = 0.2 < q && q < 5.0;
    if (maybe_stiff) {
      maybe_stiff = C.cost * std::fabs((C.t_end - C.t) / C.h_average) > 5000.0;
      if (maybe_stiff) {
        v0v0 = wdot(v0, v0);
        double yy = wdot(y, y);
        double ynrm = std::sqrt(yy);
        scale = ynrm * C.sqrrmc;
        if (scale == 0.0) {
          scale = v0v0 * C.sqrrmc;
          maybe_stiff = scale > 0.0;
        }
      }
    }
  }
  if (!maybe_stiff) return;
  if (v0v0 == 0.0) {
    for (int k = 0; k < n; ++k) v0[k] = 1.0;
    v0v0 = wdot(v0, v0);
  }
  double v0nrm = std::sqrt(v0v0);
  for (int k = 0; k < n; ++k) v0[k] = v0[k] / v0nrm;
  v0v0 = 1.0;
  typedef std::complex<double> cd;
  cd root_pair[2], prv_root_pair[2];
  double rold = 0, rho = 0;
  double* vtemp = C.scratch.data();
  auto approx_jacobian = [&](const double* v, double vdotv, double* out) {
    double temp1 = scale / std::sqrt(vdotv);
    for (int k = 0; k < n; ++k) vtemp[k] = y[k] + temp1 * v[k];
    call_f(C, C.t, vtemp, out);
    C.f_count += 1;
    for (int k = 0; k < n; ++k) out[k] = (C.h_average / temp1) * (out[k] - C.yp[k]);
  };
  auto dominant = [&](double a11, double a10, double a00) {
    double ratio = a10 / a00;
    rho = std::fabs(ratio);
    double det = a00 * a11 - a10 * a10;
    double res = std::fabs(det / a00);
    bool big = det == 0.0 || (res <= std::fabs(a11) * 0.001 * 0.001 && std::fabs(ratio - rold) <= 0.001 * rho);
    if (big) { root_pair[0] = cd(ratio, 0.0); root_pair[1] = cd(0.0, 0.0); }
    rold = ratio;
    return big;
  };
  auto qroots = [&](double alpha, double beta, cd out[2]) {
    cd temp = alpha / 2.0;
    cd sqdisc = std::sqrt(temp * temp - beta);
    if (sqdisc == cd(0.0, 0.0)) { out[0] = -temp; out[1] = -temp; }
    else {
      cd r1 = -temp + sqdisc, r2 = -temp + sqdisc;  // as written in RKSuitef90.f90:2308
      if (std::abs(r1) > std::abs(r2)) { out[0] = r1; out[1] = r2; } else { out[0] = r2; out[1] = r1; }
    }
  };
  int ntry = 1;
  for (;;) {
    approx_jacobian(v0, v0v0, v1);
    double v1v1 = wdot(v1, v1);
    if (std::sqrt(v1v1) > 1.0e10 * std::sqrt(v0v0)) return;
    double v1v0 = wdot(v1, v0);
    if (ntry == 1) {
      rold = v1v0 / v0v0;
      if (std::fabs(rold) < C.cubrmc) return;
    } else {
      if (dominant(v1v1, v1v0, v0v0)) break;
    }
    approx_jacobian(v1, v1v1, v2);
    double v2v2 = wdot(v2, v2), v2v0 = wdot(v2, v0), v2v1 = wdot(v2, v1);
    if (dominant(v2v2, v2v1, v1v1)) break;
    double det1 = v0v0 * v1v1 - v1v0 * v1v0;
    double alpha1 = (-v0v0 * v2v1 + v1v0 * v2v0) / det1;
    double beta1 = (v1v0 * v2v1 - v1v1 * v2v0) / det1;
    approx_jacobian(v2, v2v2, v3);
    double v3v3 = wdot(v3, v3), v3v1 = wdot(v3, v1), v3v2 = wdot(v3, v2);
    if (dominant(v3v3, v3v2, v2v2)) break;
    double det2 = v1v1 * v2v2 - v2v1 * v2v1;
    double alpha2 = (-v1v1 * v3v2 + v2v1 * v3v1) / det2;
    double beta2 = (v2v1 * v3v2 - v2v2 * v3v1) / det2;
    double res2 = std::fabs(v3v3 + alpha2 * v3v2 + beta2 * v3v1 + alpha2 * v3v2 + alpha2 * alpha2 * v2v2 +
                            alpha2 * beta2 * v2v1 + beta2 * v3v1 + beta2 * alpha2 * v2v1 + beta2 * beta2 * v1v1);
    if (res2 <= std::fabs(v3v3) * 0.001 * 0.001) {
      qroots(alpha1, beta1, prv_root_pair);
      qroots(alpha2, beta2, root_pair);
      rho = std::abs(prv_root_pair[0]);
      double dist = std::min(std::abs(root_pair[0] - prv_root_pair[0]), std::abs(root_pair[0] - prv_root_pair[1]));
      if (dist <= 0.001 * rho) break;
    }
    ntry += 1;
    if (ntry > C.max_stiff_iters) return;
    double v3nrm = std::sqrt(v3v3);
    for (int k = 0; k < n; ++k) v0[k] = v3[k] / v3nrm;
    v0v0 = 1.0;
  }
  if (root_pair[0].real() < 0.0) {
    if (!(std::abs(root_pair[1]) >= 0.9 * rho && root_pair[1].real() > 0.0)) {
      if (std::fabs(root_pair[0].imag()) <= std::fabs(root_pair[0].real()) * C.tan_angle) {
        sure_stiff = rho >= 0.9 * C.stability_radius;
      }
    }
  }
}

// step_integrate_r1 (RKSuitef90.f90:1185-1587).  Returns the flag (1 ok, 3 work, 4 stiff, 5 tiny h, <0 fatal).
int rk_step_integrate(RkComm& C) {
  int n = C.n;
  int ier = 1;
  double hmin = 0;
  if (C.at_t_start) {
    call_f(C, C.t, C.y.data(), C.yp.data());
    C.f_count += 1;
    for (int k = 0; k < n; ++k) C.weights[k] = std::max(std::fabs(C.y[k]), C.thresh[k]);
    bool phase1 = C.h_start == 0.0;
    C.phase2 = phase1;
    if (phase1) {
      C.h = std::fabs(C.h);
      double ypnorm = 0.0;
      for (int k = 0; k < n; ++k)
        if (C.y[k] != 0.0) ypnorm = std::max(ypnorm, std::fabs(C.yp[k]) / C.weights[k]);
      double tau = std::pow(C.tol, C.expon);
      if (C.h * ypnorm > tau) C.h = tau / ypnorm;
      hmin = std::max(C.sqtiny, C.toosml * std::max(std::fabs(C.t_start), std::fabs(C.t_end)));
      C.h = C.dir * std::max(C.h, hmin);
    }
  } else {
    if (C.at_t_end) return -10;  // fatal: already at t_end
  }
  bool phase3 = true;
  bool failed = false;
  for (;;) {  // take_step
    C.h = std::copysign(std::fabs(C.h), C.dir);
    C.at_t_end = C.dir * ((C.t + C.h) - C.t_end) >= 0.0;
    if (C.at_t_end) C.h = C.t_end - C.t;
    else if (C.dir * ((C.t + 2.0 * C.h) - C.t_end) >= 0.0) C.h = (C.t_end - C.t) / 2.0;
    hmin = std::max(C.sqtiny, C.toosml * std::max(std::fabs(C.t), std::fabs(C.t + C.h)));
    if (std::fabs(C.h) < hmin) { ier = 5; break; }
    if (C.chkeff) {
      C.hit_t_end_count += 1;
      if (C.hit_t_end_count >= 100 && C.hit_t_end_count >= C.step_count / 3) {
        ier = 2; C.hit_t_end_count = 0; break;
      }
    }
    if (!failed) {
      bool toomch = C.f_count > 5000;
      if (toomch) {
        ier = 3;
        C.full_f_count += C.f_count;
        C.f_count = 0;
      }
      bool sure_stiff = false;
      rk_stiff(C, toomch, sure_stiff);
      if (sure_stiff) ier = 4;
      if (ier != 1) break;
    }
    double htry = C.h;
    double err = 0.0;
    rk_step(C, C.t, htry, err, hmin);
    C.h = htry;
    if (err > C.tol) {
      double alpha;
      if (C.at_t_start) { phase3 = false; alpha = C.rs1; }
      else {
        C.bad_step_count += 1;
        C.stiff_bad_step_count += 1;
        if (failed) alpha = C.rs1;
        else { alpha = C.safety * std::pow(C.tol / err, C.expon); alpha = std::max(alpha, C.rs1); }
      }
      C.h = alpha * C.h;
      failed = true;
      continue;
    }
    double beta = std::pow(err / C.tol, C.expon);
    if (!C.at_t_start) {
      double t1 = std::pow(err, C.expon) / C.h;
      double t2 = std::pow(C.errold, C.expon) / C.h_old;
      if (t1 < t2 * 100.0 && t2 < t1 * 100.0) beta = beta * (t1 / t2);
    }
    double alpha = C.rs3;
    if (C.safety < beta * alpha) alpha = C.safety / beta;
    if (C.at_t_start) {
      C.phase2 = false;
      phase3 = phase3 && !C.at_t_end && (alpha > C.rs2);
      if (phase3) { C.h = alpha * C.h; continue; }
    }
    alpha = std::min(alpha, C.rs);
    if (failed) alpha = std::min(alpha, 1.0);
    alpha = std::max(alpha, C.rs1);
    C.h_old = C.h;
    C.h = alpha * C.h;
    if (C.at_t_start) C.h_average = C.h_old;
    else C.h_average = 0.9 * C.h_average + 0.1 * C.h_old;
    C.at_t_start = false;
    C.errold = err;
    C.t_old = C.t;
    if (C.at_t_end) C.t = C.t_end; else C.t = C.t + C.h_old;
    C.step_count += 1;
    // y_old = y; y = y_new; yp_old = yp (yp_old shares storage with y_new); yp = last stage
    C.y_old = C.y;
    C.y = C.scratch;
    C.scratch = C.yp;
    std::memcpy(C.yp.data(), C.stage(C.last_stage), sizeof(double) * n);
    break;
  }
  if (ier < kFatal) {
    C.at_t_end = C.t == C.t_end;
    C.chkeff = C.at_t_end;
    for (int k = 0; k < n; ++k) C.ymax[k] = std::max(std::fabs(C.y[k]), C.ymax[k]);
    if (ier == 1) C.st_interp = kUsable;
  }
  C.st_step = ier;
  if (ier == 3) C.flag3 += 1;
  if (ier == 4) C.flag4 += 1;
  return ier;
}

// interpolate_r1 with METHOD 'M' (RKSuitef90.f90:2655-2997); yp_old lives in scratch.
void rk_interpolate(RkComm& C, double t_want, double* y_want) {
  int n = C.n;
  bool initialised = C.st_interp != kUsable;
  double* p1 = C.p.data(); double* p2 = p1 + n; double* p3 = p2 + n; double* p4 = p3 + n; double* p5 = p4 + n;
  const double* yp = C.yp.data();
  const double* yp_old = C.scratch.data();
  if (!initialised) {
    // extra_stages (RKSuitef90.f90:2929-2995)
    double* ytemp = C.ytemp.data();
    for (int i = 9; i <= 11; ++i) {
      for (int j = 1; j <= i - 1; ++j) {
        double aij = C.a[i - 1][j - 1];
        if (j == 1) for (int k = 0; k < n; ++k) ytemp[k] = aij * yp_old[k];
        else if (j <= 7) { const double* s = C.stage(j - 1); for (int k = 0; k < n; ++k) ytemp[k] = ytemp[k] + aij * s[k]; }
        else if (j == 8) for (int k = 0; k < n; ++k) ytemp[k] = ytemp[k] + aij * yp[k];
        else if (j == 9) { const double* s = C.stage(7); for (int k = 0; k < n; ++k) ytemp[k] = ytemp[k] + aij * s[k]; }
        else if (j == 10) for (int k = 0; k < n; ++k) ytemp[k] = ytemp[k] + aij * C.xstage[k];
      }
      for (int k = 0; k < n; ++k) ytemp[k] = C.y_old[k] + C.h_old * ytemp[k];
      double ts = C.t_old + C.c[i - 1] * C.h_old;
      if (i == 9) call_f(C, ts, ytemp, C.stage(7));
      else if (i == 10) call_f(C, ts, ytemp, C.xstage.data());
      else call_f(C, ts, ytemp, C.stage(1));
      C.f_count += 1;
    }
    // form_intrp (RKSuitef90.f90:2826-2873); r(i,j) -> C.r[i-1][j-1]
    double (*r)[6] = C.r;
    const double *s1 = C.stage(1), *s2 = C.stage(2), *s3 = C.stage(3), *s4 = C.stage(4), *s5 = C.stage(5),
                 *s6 = C.stage(6), *s7 = C.stage(7);
    const double* xs = C.xstage.data();
    for (int k = 0; k < n; ++k) {
      p5[k] = r[4][5] * s4[k] + ((r[9][5] * xs[k] + r[7][5] * yp[k]) + (r[6][5] * s6[k] + r[5][5] * s5[k])) +
              ((r[3][5] * s3[k] + r[8][5] * s7[k]) + (r[2][5] * s2[k] + r[10][5] * s1[k]) + r[0][5] * yp_old[k]);
      p4[k] = (r[9][4] * xs[k] + r[8][4] * s7[k]) + ((r[6][4] * s6[k] + r[5][4] * s5[k]) + r[4][4] * s4[k]) +
              ((r[3][4] * s3[k] + r[7][4] * yp[k]) + (r[2][4] * s2[k] + r[10][4] * s1[k]) + r[0][4] * yp_old[k]);
      p3[k] = ((r[3][3] * s3[k] + r[7][3] * yp[k]) + (r[6][3] * s6[k] + r[5][3] * s5[k]) + r[4][3] * s4[k]) +
              ((r[9][3] * xs[k] + r[8][3] * s7[k]) + (r[2][3] * s2[k] + r[10][3] * s1[k]) + r[0][3] * yp_old[k]);
      p2[k] = r[4][2] * s4[k] + r[5][2] * s5[k] +
              ((r[2][2] * s2[k] + r[8][2] * s7[k]) + (r[9][2] * xs[k] + r[7][2] * yp[k]) + r[0][2] * yp_old[k]) +
              ((r[3][2] * s3[k] + r[10][2] * s1[k]) + r[6][2] * s6[k]);
      p1[k] = r[4][1] * s4[k] + ((r[5][1] * s5[k] + r[7][1] * yp[k]) + r[0][1] * yp_old[k]) +
              ((r[2][1] * s2[k] + r[8][1] * s7[k]) + r[9][1] * xs[k]) +
              ((r[3][1] * s3[k] + r[10][1] * s1[k]) + r[6][1] * s6[k]);
    }
    for (size_t k = 0; k < (size_t)5 * n; ++k) C.p[k] = C.h_old * C.p[k];
    C.st_interp = 1;
  }
  // evaluate_intrp (RKSuitef90.f90:2876-2899), intrp_degree = 6
  double sigma = (t_want - C.t) / C.h_old;
  for (int k = 0; k < n; ++k) {
    double v = p5[k] * sigma;
    v = (v + p4[k]) * sigma;
    v = (v + p3[k]) * sigma;
    v = (v + p2[k]) * sigma;
    v = (v + p1[k]) * sigma;
    y_want[k] = (v + C.h_old * yp[k]) * sigma + C.y[k];
  }
}

// range_integrate_r1 (RKSuitef90.f90:942-1183).  Returns flag; y_got/t_got set.
int rk_range_integrate(RkComm& C, double t_want, double& t_got, double* y_got) {
  int n = C.n;
  if (C.st_range == 5 || C.st_range == 6) return -20;
  if (C.at_t_start) {
    C.range_t_end = C.t_end;
    C.tlast = C.t_start;
    t_got = C.t_start;
    if (C.h_start == 0.0) {
      C.h = std::min(std::fabs(C.h), std::fabs(t_want - C.t_start));
      double hmin = std::max(C.sqtiny, C.toosml * std::max(std::fabs(C.t_start), std::fabs(C.t_end)));
      C.h = std::max(C.h, hmin);
    }
  } else {
    if (C.tlast == C.range_t_end) return -21;
  }
  if (C.dir * (t_want - C.tlast) <= 0.0) return -22;
  if (C.dir * (t_want - C.range_t_end) > 0.0) return -23;
  int ier = 1;
  for (;;) {
    bool goback = C.dir * (C.t - t_want) >= 0.0;
    if (goback) {
      rk_interpolate(C, t_want, y_got);
      t_got = t_want;
      break;
    }
    int flag = rk_step_integrate(C);
    ier = flag;
    if (flag == 1) continue;
    if (flag < 0) { C.st_range = kFatal; return flag; }
    t_got = C.t;
    std::memcpy(y_got, C.y.data(), sizeof(double) * n);
    break;
  }
  C.tlast = t_got;
  C.st_range = ier;
  return ier;
}

// ----------------------------------------------------------------------------
// Tensor assembly
// ----------------------------------------------------------------------------
void assemble_uniform(const Grid& g, std::vector<double>& Kd) {
  // LandauLifshitzEquationSolver.f90:935-984: tile j (source, centred at its
  // cell centre, edges dx,dy,dz) evaluated at every cell centre i.
  int n = g.n;
  size_t nn = (size_t)n * n;
  Kd.assign(6 * nn, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) {
      double x = g.pts[i * 3 + 0] - g.pts[j * 3 + 0];
      double y = g.pts[i * 3 + 1] - g.pts[j * 3 + 1];
      double z = g.pts[i * 3 + 2] - g.pts[j * 3 + 2];
      double k6[6];
      prism_k6(g.dx, g.dy, g.dz, x, y, z, k6);
      for (int c = 0; c < 6; ++c) Kd[c * nn + (size_t)i * n + j] = k6[c];
    }
  }
}

struct RhsCtx {
  Problem* P;
  std::vector<double> Hm64;
};

void micromag_rhs(void* ctx, double t, const double* y, double* out) {
  RhsCtx* R = (RhsCtx*)ctx;
  dmdt(*R->P, t, y, out, R->Hm64);
}

struct GenericCtx {
  void (*cb)(double, const double*, double*, void*);
  void* user;
};

void generic_rhs(void* ctx, double t, const double* y, double* out) {
  GenericCtx* g = (GenericCtx*)ctx;
  g->cb(t, y, out, g->user);
}

}  // namespace

// ============================================================================
// C ABI (ctypes)
// ============================================================================
extern "C" {

// K (MagTense sign) of one prism, full edges a,b,c, at offset pos from its centre.  out: xx,xy,xz,yy,yz,zz
void or_prism_k6(double a, double b, double c, const double* pos, double* out) {
  prism_k6(a, b, c, pos[0], pos[1], pos[2], out);
}

// Uniform-grid cell centres (row-major n*3) and spacings.
int or_setup_grid(int nx, int ny, int nz, double Lx, double Ly, double Lz, double* pts, double* dxyz) {
  Grid g;
  setup_grid(nx, ny, nz, Lx, Ly, Lz, g);
  std::memcpy(pts, g.pts.data(), sizeof(double) * g.pts.size());
  dxyz[0] = g.dx; dxyz[1] = g.dy; dxyz[2] = g.dz;
  return 0;
}

// Full FP64 tensor on a uniform grid: K6[c*N*N + i*N + j].
int or_tensor_uniform(int nx, int ny, int nz, double Lx, double Ly, double Lz, double* K6) {
  Grid g;
  setup_grid(nx, ny, nz, Lx, Ly, Lz, g);
  std::vector<double> Kd;
  assemble_uniform(g, Kd);
  std::memcpy(K6, Kd.data(), sizeof(double) * Kd.size());
  return 0;
}

// Tensor for arbitrary prisms (sources) at arbitrary points (receivers):
// out[c*n_pts*n_src + i*n_src + j] for source j (centre src[j*3..], edges abc[j*3..])
// at point pts[i*3..].  Only source columns [j0, j1) are filled (the rest is untouched).
int or_tensor_points(int n_src, const double* src, const double* abc, int n_pts, const double* pts, int j0, int j1,
                     double* out) {
  size_t plane = (size_t)n_pts * n_src;
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = j0; j < j1; ++j)
    for (int i = 0; i < n_pts; ++i) {
      double k6[6];
      prism_k6(abc[j * 3], abc[j * 3 + 1], abc[j * 3 + 2], pts[i * 3] - src[j * 3], pts[i * 3 + 1] - src[j * 3 + 1],
               pts[i * 3 + 2] - src[j * 3 + 2], k6);
      for (int c = 0; c < 6; ++c) out[c * plane + (size_t)i * n_src + j] = k6[c];
    }
  return 0;
}

// Exchange operator in CSR (0-based, columns ascending); returns nnz.  Pass null arrays to query nnz.
int or_exchange_csr(int nx, int ny, int nz, double Lx, double Ly, double Lz, int* rowptr, int* col, double* val) {
  if (nx < 2 || ny < 2) return -1;
  Grid g;
  setup_grid(nx, ny, nz, Lx, Ly, Lz, g);
  Csr A;
  build_exchange(g, A);
  if (rowptr) {
    std::memcpy(rowptr, A.rowptr.data(), sizeof(int) * A.rowptr.size());
    std::memcpy(col, A.col.data(), sizeof(int) * A.col.size());
    std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
  }
  return (int)A.col.size();
}

// Problem description shared by or_dmdt and or_run_micromag.
typedef struct {
  int nx, ny, nz;
  double Lx, Ly, Lz;
  const double* u_ea;    // n*3 row-major (may be null: no anisotropy axis)
  const double* Ms;      // n
  const double* K0;      // n
  double A0, gamma, alpha0;
  int solver;            // 1 explicit, 2 dynamic
  int nt_hext;
  const double* hext;    // nt_hext*4 row-major [t, hx, hy, hz] (A/m)
  int nt_alpha;
  const double* alphat;  // nt_alpha*2 row-major
  int nt;
  const double* t;
  const double* m0;      // 3n SoA
  double tol, thres;
  int nt_conv;
  const double* t_conv;
  double conv_tol;
  int demag_fp64;        // 0: FP32 demag product as the reference; 1: FP64 product from FP64 tensor
  const float* K_in;     // optional 6*n*n FP32 tensor (receiver-major); null: assemble
  int return_fields;     // 1: fill H_exc/H_ext/H_dem/H_ani
  int demag_approx;      // 0/1 dense, 2 threshold, 4 threshold fraction
  double demag_threshold;
  // grid of unstructured prisms (gridTypeUnstructuredPrisms = 3,
  // MagTenseMicroMagPyIO.f90:76-85,163-180); grid_type 0 or 1: uniform
  int grid_type;
  const double* grid_pts;       // n*3 row-major cell centres
  const double* grid_abc;       // n*3 row-major full edges
  int n_ave[3];                 // receiver sub-points per axis (N_ave)
  int exch_nnz;
  const int* exch_rows_start;   // n, 1-based (mkl_sparse_d_create_csr rows_start)
  const int* exch_rows_end;     // n, 1-based
  const int* exch_cols;         // exch_nnz, 1-based
  const double* exch_vals;      // exch_nnz
  // tetrahedral grid (gridTypeTetrahedron = 2): grid_pts are the evaluation
  // points, element i has the 1-based nodes grid_elements[4i .. 4i+3]
  int grid_nnod;
  const double* grid_nodes;     // grid_nnod*3 row-major
  const int* grid_elements;     // 4n
} or_problem;

typedef struct {
  long long f_calls, steps, bad_steps, flag3, flag4;
  double h_last;
} or_stats;

}  // extern "C"

namespace {

// FindThresholdFraction (DemagAuxFunctions.f90:353-428) over the positive
// entries of the FP32 tensor (ApplyThresholdDense packs K > 0, :35-105):
// FP32 bisection; counts in 64-bit integers.
float threshold_fraction(const float* K, size_t total, double fraction) {
  const float thr = (float)fraction;  // sngl(dem_thres), MagTenseMicroMagPyIO.f90:117-118
  if (thr >= 1.0f) return FLT_MAX;
  long long npos = 0;
  float f_large = 0.0f, f_small = FLT_MAX;
  for (size_t k = 0; k < total; ++k)
    if (K[k] > 0.0f) {
      ++npos;
      f_large = std::max(f_large, K[k]);
      f_small = std::min(f_small, K[k]);
    }
  if (npos == 0) return -1.0f;
  const float nf = (float)npos;
  float f_middle = f_small;
  for (int k_do = 1;; ++k_do) {
    f_middle = (f_large - f_small) / 2.0f + f_small;
    long long cm = 0;
#pragma omp parallel for reduction(+ : cm)
    for (long long k = 0; k < (long long)total; ++k) cm += (K[k] > 0.0f && K[k] <= f_middle);
    const float cmf = (float)cm;
    if (cmf > thr * nf) f_large = f_middle;
    else f_small = f_middle;
    if (cmf >= (thr - 0.01f) * nf && cmf <= (thr + 0.01f) * nf) break;
    if (k_do > 1000) break;
  }
  return f_middle;
}

// ConvertDenseToSparse_s (util_micromag.f90:81-138): entries with |K| <= thr
// are dropped; the dense product over the remaining entries equals the CSR one.
void apply_threshold(std::vector<float>& K, float thr) {
  for (auto& v : K)
    if (!(std::fabs(v) > thr)) v = 0.0f;
}

// ComputeDemagfieldTensor, unstructured-prism branch
// (LandauLifshitzEquationSolver.f90:1044-1133): tile j (edges abc_j, centred at
// p_j) evaluated at the nx_ave x ny_ave x nz_ave sub-points
// (p_i - abc_i/2) + (abc_i/(n_ave+1)) * (i_a, j_a, k_a) of every receiver cell
// i, summed in FP64 (k_a outer, i_a inner) and divided by their number.
void assemble_unstructured(int n, const double* pts, const double* abc, const int nave[3], std::vector<double>& Kd) {
  size_t nn = (size_t)n * n;
  Kd.assign(6 * nn, 0.0);
  const double cnt = (double)(nave[0] * nave[1] * nave[2]);
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n; ++j) {
    const double* aj = abc + (size_t)j * 3;
    const double* cj = pts + (size_t)j * 3;
    for (int i = 0; i < n; ++i) {
      const double* pi = pts + (size_t)i * 3;
      const double* ai = abc + (size_t)i * 3;
      const double d[3] = {ai[0] / (nave[0] + 1), ai[1] / (nave[1] + 1), ai[2] / (nave[2] + 1)};
      double acc[6] = {0, 0, 0, 0, 0, 0};
      for (int ka = 1; ka <= nave[2]; ++ka)
        for (int ja = 1; ja <= nave[1]; ++ja)
          for (int ia = 1; ia <= nave[0]; ++ia) {
            const double px = (pi[0] - ai[0] / 2) + d[0] * ia;
            const double py = (pi[1] - ai[1] / 2) + d[1] * ja;
            const double pz = (pi[2] - ai[2] / 2) + d[2] * ka;
            double k6[6];
            prism_k6(aj[0], aj[1], aj[2], px - cj[0], py - cj[1], pz - cj[2], k6);
            for (int c = 0; c < 6; ++c) acc[c] += k6[c];
          }
      for (int c = 0; c < 6; ++c) Kd[c * nn + (size_t)i * n + j] = acc[c] / cnt;
    }
  }
}

// ComputeDemagfieldTensor, tetrahedron branch (LandauLifshitzEquationSolver.f90:988-1041):
// element j (vertices = its four nodes, offset 0, no rotation) at every
// evaluation point i; the upper triangle N(1,1) N(1,2) N(1,3) N(2,2) N(2,3)
// N(3,3) is stored (:1026-1037).  N_ave is not used for tetrahedra (:990-992).
int tetra_tensor_at(const double* vert12, const double r[3], double N9[9]);
int assemble_tetra(int n, const double* pts, int nnod, const double* nodes, const int* elem, std::vector<double>& Kd) {
  size_t nn = (size_t)n * n;
  Kd.assign(6 * nn, 0.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < 4; ++k)
      if (elem[4 * j + k] < 1 || elem[4 * j + k] > nnod) return -1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad)
  for (int j = 0; j < n; ++j) {
    double v[12];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) v[3 * k + c] = nodes[(size_t)(elem[4 * j + k] - 1) * 3 + c];
    for (int i = 0; i < n; ++i) {
      double N[9];
      if (tetra_tensor_at(v, pts + (size_t)i * 3, N)) { ++bad; continue; }
      const double k6[6] = {N[0], N[1], N[2], N[4], N[5], N[8]};
      for (int c = 0; c < 6; ++c) Kd[c * nn + (size_t)i * n + j] = k6[c];
    }
  }
  return bad ? -1 : 0;
}

// ConvertExchangeTerm3D_NonUniform (LandauLifshitzEquationSolver.f90:1530-1544):
// the loaded 1-based CSR (rows_start / rows_end) as it is; the product sums each
// row's entries in stored order.
int loaded_exchange(const or_problem* q, int n, Csr& A) {
  if (!q->exch_rows_start || !q->exch_rows_end || !q->exch_cols || !q->exch_vals) return -1;
  A.n = n;
  A.rowptr.assign(n + 1, 0);
  A.col.clear(); A.val.clear();
  for (int r = 0; r < n; ++r) {
    for (int k = q->exch_rows_start[r] - 1; k < q->exch_rows_end[r] - 1; ++k) {
      if (k < 0 || k >= q->exch_nnz) return -1;
      const int c = q->exch_cols[k] - 1;
      if (c < 0 || c >= n) return -1;
      A.col.push_back(c);
      A.val.push_back(q->exch_vals[k]);
    }
    A.rowptr[r + 1] = (int)A.col.size();
  }
  return 0;
}

int build_problem(const or_problem* q, Problem& P) {
  const bool tetra = q->grid_type == 2;
  const bool unstructured = q->grid_type == 3 || tetra;
  if (q->grid_type != 0 && q->grid_type != 1 && !unstructured) return -3;
  if (!unstructured && (q->nx < 2 || q->ny < 2)) return -1;
  int n;
  if (unstructured) {
    n = q->nx * q->ny * q->nz;
    if (!q->grid_pts) return -1;
    if (tetra && (!q->grid_nodes || !q->grid_elements || q->grid_nnod < 4)) return -1;
    if (!tetra && (!q->grid_abc || q->n_ave[0] < 1 || q->n_ave[1] < 1 || q->n_ave[2] < 1)) return -1;
    P.g.nx = q->nx; P.g.ny = q->ny; P.g.nz = q->nz; P.g.n = n;
    P.g.pts.assign(q->grid_pts, q->grid_pts + (size_t)n * 3);
    if (loaded_exchange(q, n, P.Aex)) return -1;
  } else {
    setup_grid(q->nx, q->ny, q->nz, q->Lx, q->Ly, q->Lz, P.g);
    n = P.g.n;
    build_exchange(P.g, P.Aex);
  }
  P.n = n;
  size_t nn = (size_t)n * n;
  P.demag_fp64 = q->demag_fp64 != 0;
  if (P.demag_fp64 || !q->K_in) {
    if (tetra) {
      if (assemble_tetra(n, q->grid_pts, q->grid_nnod, q->grid_nodes, q->grid_elements, P.Kd)) return -1;
    } else if (unstructured) {
      assemble_unstructured(n, q->grid_pts, q->grid_abc, q->n_ave, P.Kd);
    } else {
      assemble_uniform(P.g, P.Kd);
    }
    if (!P.demag_fp64) {
      P.K.resize(6 * nn);
      for (size_t k = 0; k < 6 * nn; ++k) P.K[k] = (float)P.Kd[k];
      std::vector<double>().swap(P.Kd);
    }
  } else {
    P.K.assign(q->K_in, q->K_in + 6 * nn);
  }
  if (q->demag_approx == 2 || q->demag_approx == 4) {
    if (P.demag_fp64) return -3;
    float thr = (float)q->demag_threshold;
    if (q->demag_approx == 4) thr = threshold_fraction(P.K.data(), P.K.size(), q->demag_threshold);
    if (thr < 0.0f) return -3;
    apply_threshold(P.K, thr);
    P.demag_threshold = thr;
  } else if (q->demag_approx != 0 && q->demag_approx != 1) {
    return -3;
  }
  P.Axx.resize(n); P.Axy.resize(n); P.Axz.resize(n); P.Ayy.resize(n); P.Ayz.resize(n); P.Azz.resize(n);
  for (int i = 0; i < n; ++i) {
    double ux = q->u_ea ? q->u_ea[i * 3] : 0, uy = q->u_ea ? q->u_ea[i * 3 + 1] : 0, uz = q->u_ea ? q->u_ea[i * 3 + 2] : 0;
    P.Axx[i] = ux * ux; P.Axy[i] = ux * uy; P.Axz[i] = ux * uz; P.Ayy[i] = uy * uy; P.Ayz[i] = uy * uz; P.Azz[i] = uz * uz;
  }
  P.Jfact.resize(n); P.Kfact.resize(n); P.Mfact.resize(n); P.Ms64.assign(q->Ms, q->Ms + n);
  for (int i = 0; i < n; ++i) {
    P.Jfact[i] = q->A0 / (kMu0 * q->Ms[i]);
    P.Mfact[i] = (float)q->Ms[i];
    P.Kfact[i] = q->K0[i] / (kMu0 * q->Ms[i]);
  }
  P.gamma = q->gamma; P.alpha0 = q->alpha0; P.solver = q->solver;
  for (int k = 0; k < q->nt_alpha; ++k) { P.alpha_t.push_back(q->alphat[k * 2]); P.alpha_v.push_back(q->alphat[k * 2 + 1]); }
  for (int k = 0; k < q->nt_hext; ++k) {
    P.hext_t.push_back(q->hext[k * 4]); P.hext_x.push_back(q->hext[k * 4 + 1]);
    P.hext_y.push_back(q->hext[k * 4 + 2]); P.hext_z.push_back(q->hext[k * 4 + 3]);
  }
  P.Hj.assign(3 * (size_t)n, 0.0); P.Hh.assign(3 * (size_t)n, 0.0); P.Hk.assign(3 * (size_t)n, 0.0);
  P.Hm.assign(3 * (size_t)n, 0.0f); P.ms.assign(3 * (size_t)n, 0.0f);
  return 0;
}

}  // namespace

extern "C" {

// One right-hand-side evaluation dm/dt(t, m); optionally returns the four
// (negated) field terms, each 3n SoA (pass nulls to skip).
int or_dmdt(const or_problem* q, double t, const double* m, double* out, double* Hj, double* Hh, double* Hm,
            double* Hk) {
  Problem P;
  int rc = build_problem(q, P);
  if (rc) return rc;
  std::vector<double> Hm64(3 * (size_t)P.n);
  dmdt(P, t, m, out, Hm64);
  size_t n3 = 3 * (size_t)P.n;
  if (Hj) std::memcpy(Hj, P.Hj.data(), sizeof(double) * n3);
  if (Hh) std::memcpy(Hh, P.Hh.data(), sizeof(double) * n3);
  if (Hk) std::memcpy(Hk, P.Hk.data(), sizeof(double) * n3);
  if (Hm) for (size_t k = 0; k < n3; ++k) Hm[k] = P.demag_fp64 ? Hm64[k] : (double)P.Hm[k];
  return 0;
}

// Whole simulation: MagTense_ODE_RKSuite driver (ODE_Solvers.f90:124-237) per
// applied-field row (explicit) or once (dynamic), then the field storage of
// LandauLifshitzEquationSolver.f90:306-349.
// Outputs (n_h = nt_hext for explicit, 1 for dynamic):
//   t_out[n_h*nt], y_out[n_h*nt*3n] (SoA per time), fields[4][n_h*nt*3n] when requested.
int or_run_micromag(const or_problem* q, double* t_out, double* y_out, double* H_exc, double* H_ext, double* H_dem,
                    double* H_ani, or_stats* stats) {
  Problem P;
  int rc = build_problem(q, P);
  if (rc) return rc;
  int n = P.n, n3 = 3 * n, nt = q->nt;
  int n_h = q->solver == 1 ? q->nt_hext : 1;
  std::vector<double> m0(q->m0, q->m0 + n3);
  RhsCtx R; R.P = &P; R.Hm64.assign(n3, 0.0);
  std::vector<double> thres(n3, q->thres);
  if (stats) std::memset(stats, 0, sizeof(or_stats));
  // Combined output/convergence times (simple_unique then simple_sort)
  std::vector<double> tc(q->t, q->t + nt);
  for (int k = 0; k < q->nt_conv; ++k) tc.push_back(q->t_conv[k]);
  std::vector<double> uniq;
  for (double v : tc) if (std::find(uniq.begin(), uniq.end(), v) == uniq.end()) uniq.push_back(v);
  std::vector<double> sorted = uniq;
  std::stable_sort(sorted.begin(), sorted.end());
  std::vector<double> y_step(n3), y_last(n3);
  for (int hi = 0; hi < n_h; ++hi) {
    P.hext_ind = hi;
    double* to = t_out + (size_t)hi * nt;
    double* yo = y_out + (size_t)hi * nt * n3;
    for (int k = 0; k < nt; ++k) to[k] = 0.0;
    std::memset(yo, 0, sizeof(double) * (size_t)nt * n3);
    RkComm C;
    rc = rk_setup(C, n3, micromag_rhs, &R, q->t[0], m0.data(), q->t[nt - 1], q->tol, thres.data());
    if (rc) return rc;
    y_last = m0;
    to[0] = q->t[0];
    std::memcpy(yo, m0.data(), sizeof(double) * n3);
    int k = 1;
    for (size_t i = 1; i < sorted.size(); ++i) {
      double tw = sorted[i], t_step = 0;
      int flag = rk_range_integrate(C, tw, t_step, y_step.data());
      if (flag < 0) return flag;
      bool is_conv = false;
      for (int c = 0; c < q->nt_conv; ++c) if (q->t_conv[c] == tw) is_conv = true;
      if (is_conv) {
        double ce = 0.0;
        for (int r = 0; r < n3; ++r) ce = std::max(ce, std::fabs(y_step[r] - y_last[r]));
        if (ce < q->conv_tol) {
          if (k < nt) { std::memcpy(yo + (size_t)k * n3, y_step.data(), sizeof(double) * n3); to[k] = t_step; }
          break;
        }
        y_last = y_step;
      }
      bool is_out = false;
      for (int c = 0; c < nt; ++c) if (q->t[c] == tw) is_out = true;
      if (is_out && k < nt) {
        std::memcpy(yo + (size_t)k * n3, y_step.data(), sizeof(double) * n3);
        to[k] = t_step;
        k += 1;
      }
    }
    if (stats) {
      stats->f_calls += C.full_f_count + C.f_count;
      stats->steps += C.step_count;
      stats->bad_steps += C.bad_step_count;
      stats->flag3 += C.flag3;
      stats->flag4 += C.flag4;
      stats->h_last = C.h;
    }
    // next field starts from the last stored state (LandauLifshitzEquationSolver.f90:177)
    std::memcpy(m0.data(), yo + (size_t)(nt - 1) * n3, sizeof(double) * n3);
    if (q->return_fields && H_exc) {
      std::vector<double> tmp(n3);
      for (int j = 0; j < nt; ++j) {
        const double* m = yo + (size_t)j * n3;
        dmdt(P, q->t[j], m, tmp.data(), R.Hm64);
        size_t off = ((size_t)hi * nt + j) * n3;
        std::memcpy(H_exc + off, P.Hj.data(), sizeof(double) * n3);
        std::memcpy(H_ext + off, P.Hh.data(), sizeof(double) * n3);
        std::memcpy(H_ani + off, P.Hk.data(), sizeof(double) * n3);
        for (int r = 0; r < n3; ++r) H_dem[off + r] = P.demag_fp64 ? R.Hm64[r] : (double)P.Hm[r];
      }
    }
  }
  return 0;
}

// Generic RKSUITE 'M' integration of y' = cb(t, y) with outputs at t[0..nt) (t[0] = start).
int or_rk_integrate(int neq, void (*cb)(double, const double*, double*, void*), void* user, int nt, const double* t,
                    const double* y0, double tol, double thres, double* t_out, double* y_out, or_stats* stats) {
  GenericCtx g; g.cb = cb; g.user = user;
  RkComm C;
  std::vector<double> th(neq, thres);
  int rc = rk_setup(C, neq, generic_rhs, &g, t[0], y0, t[nt - 1], tol, th.data());
  if (rc) return rc;
  t_out[0] = t[0];
  std::memcpy(y_out, y0, sizeof(double) * neq);
  for (int i = 1; i < nt; ++i) {
    double tg = 0;
    int flag = rk_range_integrate(C, t[i], tg, y_out + (size_t)i * neq);
    if (flag < 0) return flag;
    t_out[i] = tg;
  }
  if (stats) {
    stats->f_calls = C.full_f_count + C.f_count;
    stats->steps = C.step_count;
    stats->bad_steps = C.bad_step_count;
    stats->flag3 = C.flag3;
    stats->flag4 = C.flag4;
    stats->h_last = C.h;
  }
  return 0;
}

// The RKSUITE 'M' coefficient tables (for order-condition tests): a[11*11], b[8], c[11], e[7], r[11*6].
void or_rk_tables(double* a, double* b, double* c, double* e, double* r) {
  RkComm C;
  method_const_M(C);
  for (int i = 0; i < 11; ++i) for (int j = 0; j < 11; ++j) a[i * 11 + j] = C.a[i][j];
  for (int i = 0; i < 8; ++i) b[i] = C.b[i];
  for (int i = 0; i < 11; ++i) c[i] = C.c[i];
  for (int i = 0; i < 7; ++i) e[i] = C.e[i];
  for (int i = 0; i < 11; ++i) for (int j = 0; j < 6; ++j) r[i * 6 + j] = C.r[i][j];
}

}  // extern "C"

extern "C" {

// Demag rows with the reference's FP32 semantics (LandauLifshitzEquationSolver.f90:706-754):
// out[a*R + r] = fl32(fl32(-sum_b K_ab(row r) m_b) * Ms_r) for R receiver rows whose
// tensor rows are Krows[c*R*n + r*n + j].  Used to time a bounded CPU sample.
int or_demag_rows(int n, int R, const float* Krows, const float* ms_rows, const double* m, double* out) {
  static const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  std::vector<float> m32(3 * (size_t)n);
  for (size_t k = 0; k < 3 * (size_t)n; ++k) m32[k] = (float)m[k];
  size_t plane = (size_t)R * n;
  for (int a = 0; a < 3; ++a) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; ++r) {
      float h = 0.0f;
      for (int b = 0; b < 3; ++b) {
        const float* Krow = Krows + comp[a][b] * plane + (size_t)r * n;
        const float* v = m32.data() + (size_t)b * n;
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += (double)Krow[j] * (double)v[j];
        h = (float)((double)h - s);
      }
      out[(size_t)a * R + r] = (double)(h * ms_rows[r]);
    }
  }
  return 0;
}

}  // extern "C"

// ============================================================================
// FP64 discrete adjoint of fixed-step RK4 (checker for the GPU adjoint).
// Operators are transposed explicitly (CSR scatter, dense block transpose), so
// the symmetry the GPU kernels rely on is checked rather than assumed.
// ============================================================================
namespace {

struct AdjCtx {
  Problem* P;
  double hext[3];
  std::vector<double> Hm64;
};

// negated total field H(m) and its pieces, FP64 (demag from the FP64 tensor when
// demag_fp64, else the reference's FP32 semantics)
void total_field(AdjCtx& A, const double* m, std::vector<double>& H, std::vector<double>& hm, std::vector<double>& hj,
                 std::vector<double>& hk) {
  Problem& P = *A.P;
  int n = P.n;
  P.hext_x.assign(2, A.hext[0]); P.hext_y.assign(2, A.hext[1]); P.hext_z.assign(2, A.hext[2]);
  P.hext_t = {0.0, 1.0};
  P.solver = 2;
  update_fields(P, m, 0.0);
  if (P.demag_fp64) demag_fp64(P, m, A.Hm64.data());
  H.resize(3 * (size_t)n); hm.resize(3 * (size_t)n); hj.resize(3 * (size_t)n); hk.resize(3 * (size_t)n);
  for (size_t k = 0; k < 3 * (size_t)n; ++k) {
    hm[k] = P.demag_fp64 ? A.Hm64[k] : (double)P.Hm[k];
    hj[k] = P.Hj[k];
    hk[k] = P.Hk[k];
    H[k] = -A.hext[k / n] + hj[k] + hm[k] + hk[k];
  }
}

void rhs_adj(AdjCtx& A, const double* m, double* out) {
  Problem& P = *A.P;
  int n = P.n;
  std::vector<double> H, hm, hj, hk;
  total_field(A, m, H, hm, hj, hk);
  for (int r = 0; r < n; ++r) {
    double mx = m[r], my = m[n + r], mz = m[2 * n + r];
    double Hx = H[r], Hy = H[n + r], Hz = H[2 * n + r];
    double cx = -(my * Hz - mz * Hy), cy = -(mz * Hx - mx * Hz), cz = -(mx * Hy - my * Hx);
    out[r] = -P.gamma * cx - P.alpha0 * (my * cz - mz * cy);
    out[n + r] = -P.gamma * cy - P.alpha0 * (mz * cx - mx * cz);
    out[2 * n + r] = -P.gamma * cz - P.alpha0 * (mx * cy - my * cx);
  }
}

// mu = (df/dm)^T u ; gh += dL/dhext ; gms += dL/dMs
void vjp_adj(AdjCtx& A, const double* m, const double* u, double* mu, double* gh, double* gms) {
  Problem& P = *A.P;
  int n = P.n;
  std::vector<double> H, hm, hj, hk;
  total_field(A, m, H, hm, hj, hk);
  std::vector<double> hb(3 * (size_t)n);
  for (int r = 0; r < n; ++r) {
    double mm[3] = {m[r], m[n + r], m[2 * n + r]}, uu[3] = {u[r], u[n + r], u[2 * n + r]};
    double HH[3] = {H[r], H[n + r], H[2 * n + r]};
    double c[3] = {-(mm[1] * HH[2] - mm[2] * HH[1]), -(mm[2] * HH[0] - mm[0] * HH[2]), -(mm[0] * HH[1] - mm[1] * HH[0])};
    double uxm[3] = {uu[1] * mm[2] - uu[2] * mm[1], uu[2] * mm[0] - uu[0] * mm[2], uu[0] * mm[1] - uu[1] * mm[0]};
    double cb[3], cxu[3], hxc[3];
    for (int k = 0; k < 3; ++k) cb[k] = -P.gamma * uu[k] - P.alpha0 * uxm[k];
    cxu[0] = c[1] * uu[2] - c[2] * uu[1]; cxu[1] = c[2] * uu[0] - c[0] * uu[2]; cxu[2] = c[0] * uu[1] - c[1] * uu[0];
    hxc[0] = HH[1] * cb[2] - HH[2] * cb[1]; hxc[1] = HH[2] * cb[0] - HH[0] * cb[2]; hxc[2] = HH[0] * cb[1] - HH[1] * cb[0];
    double hbr[3] = {-(cb[1] * mm[2] - cb[2] * mm[1]), -(cb[2] * mm[0] - cb[0] * mm[2]), -(cb[0] * mm[1] - cb[1] * mm[0])};
    double ms = P.demag_fp64 ? P.Ms64[r] : (double)P.Mfact[r];
    double g = 0.0;
    for (int k = 0; k < 3; ++k) {
      mu[(size_t)k * n + r] = -P.alpha0 * cxu[k] - hxc[k];
      hb[(size_t)k * n + r] = hbr[k];
      gh[k] += -hbr[k];
      g += hbr[k] * (hm[(size_t)k * n + r] - hj[(size_t)k * n + r] - hk[(size_t)k * n + r]);
    }
    gms[r] += g / ms;
  }
  // exchange transpose: H_j = -2 Jfact (A m)  ->  mu += A^T (-2 Jfact hb)
  for (int c3 = 0; c3 < 3; ++c3)
    for (int r = 0; r < n; ++r) {
      double w = -2.0 * P.Jfact[r] * hb[(size_t)c3 * n + r];
      for (int q = P.Aex.rowptr[r]; q < P.Aex.rowptr[r + 1]; ++q) mu[(size_t)c3 * n + P.Aex.col[q]] += P.Aex.val[q] * w;
    }
  // demag transpose: H_m,a(i) = -Ms_i sum_b K_ab(i,j) m_b(j)  ->  mu_b(j) += -sum_a,i K_ab(i,j) Ms_i hb_a(i)
  static const int comp[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  size_t nn = (size_t)n * n;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      int c6 = comp[a][b];
      for (int i = 0; i < n; ++i) {
        double w = -(P.demag_fp64 ? P.Ms64[i] : (double)P.Mfact[i]) * hb[(size_t)a * n + i];
        if (P.demag_fp64) {
          const double* Krow = P.Kd.data() + c6 * nn + (size_t)i * n;
          for (int j = 0; j < n; ++j) mu[(size_t)b * n + j] += Krow[j] * w;
        } else {
          const float* Krow = P.K.data() + c6 * nn + (size_t)i * n;
          for (int j = 0; j < n; ++j) mu[(size_t)b * n + j] += (double)Krow[j] * w;
        }
      }
    }
  // anisotropy transpose (symmetric 3x3 per cell): mu += -2 Kfact A_ani hb
  for (int r = 0; r < n; ++r) {
    double f = -2.0 * P.Kfact[r];
    double h0 = hb[r], h1 = hb[n + r], h2 = hb[2 * n + r];
    mu[r] += f * (P.Axx[r] * h0 + P.Axy[r] * h1 + P.Axz[r] * h2);
    mu[n + r] += f * (P.Axy[r] * h0 + P.Ayy[r] * h1 + P.Ayz[r] * h2);
    mu[2 * n + r] += f * (P.Axz[r] * h0 + P.Ayz[r] * h1 + P.Azz[r] * h2);
  }
}

}  // namespace

extern "C" {

// RK4 forward (mT) and, when dL_dmT is given, the discrete adjoint.  Uses
// q's grid, materials, alpha0, gamma and demag_fp64; the applied field is hext3.
int or_rk4_adjoint(const or_problem* q, double dt, long long nsteps, const double* hext3, const double* m0, double* mT,
                   const double* dL_dmT, double* grad_m0, double* grad_hext, double* grad_ms) {
  Problem P;
  int rc = build_problem(q, P);
  if (rc) return rc;
  int n = P.n;
  size_t n3 = 3 * (size_t)n;
  AdjCtx A;
  A.P = &P;
  A.hext[0] = hext3[0]; A.hext[1] = hext3[1]; A.hext[2] = hext3[2];
  A.Hm64.assign(n3, 0.0);
  std::vector<double> Y((size_t)(nsteps + 1) * n3);
  std::vector<double> k1(n3), k2(n3), k3(n3), k4(n3), tmp(n3);
  std::memcpy(Y.data(), m0, n3 * sizeof(double));
  for (long long s = 0; s < nsteps; ++s) {
    const double* y = Y.data() + (size_t)s * n3;
    double* yn = Y.data() + (size_t)(s + 1) * n3;
    rhs_adj(A, y, k1.data());
    for (size_t k = 0; k < n3; ++k) tmp[k] = y[k] + 0.5 * dt * k1[k];
    rhs_adj(A, tmp.data(), k2.data());
    for (size_t k = 0; k < n3; ++k) tmp[k] = y[k] + 0.5 * dt * k2[k];
    rhs_adj(A, tmp.data(), k3.data());
    for (size_t k = 0; k < n3; ++k) tmp[k] = y[k] + dt * k3[k];
    rhs_adj(A, tmp.data(), k4.data());
    for (size_t k = 0; k < n3; ++k) yn[k] = y[k] + dt / 6.0 * (k1[k] + 2.0 * k2[k] + 2.0 * k3[k] + k4[k]);
  }
  if (mT) std::memcpy(mT, Y.data() + (size_t)nsteps * n3, n3 * sizeof(double));
  if (!dL_dmT) return 0;
  std::vector<double> lam(dL_dmT, dL_dmT + n3), kb(n3), mu1(n3), mu2(n3), mu3(n3), mu4(n3);
  std::vector<double> Y2(n3), Y3(n3), Y4(n3);
  std::vector<double> gh(3, 0.0), gms(n, 0.0);
  for (long long s = nsteps - 1; s >= 0; --s) {
    const double* y = Y.data() + (size_t)s * n3;
    rhs_adj(A, y, k1.data());
    for (size_t k = 0; k < n3; ++k) Y2[k] = y[k] + 0.5 * dt * k1[k];
    rhs_adj(A, Y2.data(), k2.data());
    for (size_t k = 0; k < n3; ++k) Y3[k] = y[k] + 0.5 * dt * k2[k];
    rhs_adj(A, Y3.data(), k3.data());
    for (size_t k = 0; k < n3; ++k) Y4[k] = y[k] + dt * k3[k];
    for (size_t k = 0; k < n3; ++k) kb[k] = dt / 6.0 * lam[k];
    vjp_adj(A, Y4.data(), kb.data(), mu4.data(), gh.data(), gms.data());
    for (size_t k = 0; k < n3; ++k) kb[k] = dt / 3.0 * lam[k] + dt * mu4[k];
    vjp_adj(A, Y3.data(), kb.data(), mu3.data(), gh.data(), gms.data());
    for (size_t k = 0; k < n3; ++k) kb[k] = dt / 3.0 * lam[k] + 0.5 * dt * mu3[k];
    vjp_adj(A, Y2.data(), kb.data(), mu2.data(), gh.data(), gms.data());
    for (size_t k = 0; k < n3; ++k) kb[k] = dt / 6.0 * lam[k] + 0.5 * dt * mu2[k];
    vjp_adj(A, y, kb.data(), mu1.data(), gh.data(), gms.data());
    for (size_t k = 0; k < n3; ++k) lam[k] += mu1[k] + mu2[k] + mu3[k] + mu4[k];
  }
  if (grad_m0) std::memcpy(grad_m0, lam.data(), n3 * sizeof(double));
  if (grad_hext) std::memcpy(grad_hext, gh.data(), 3 * sizeof(double));
  if (grad_ms) std::memcpy(grad_ms, gms.data(), n * sizeof(double));
  return 0;
}

}  // extern "C"

// ----------------------------------------------------------------------------
// Magnetostatic prism tiles (getNFromTiles / getHFromTiles restricted to
// rectangular prisms, python/src/magtense/lib/FortranToPythonIO.f90:25-223).
// Arrays are Fortran-ordered as f2py passes them: a (n,3) array holds element
// (i,k) at i + n*k; the tensor N(n_tiles,n_pts,3,3) holds (i,p,a,b) at
// i + nt*(p + np*(a + 3b)).
// ----------------------------------------------------------------------------
namespace {

typedef double Mat3[3][3];

void matmul3(const Mat3 A, const Mat3 B, Mat3 C) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[r][c] = A[r][0] * B[0][c] + A[r][1] * B[1][c] + A[r][2] * B[2][c];
}

// DemagFieldGetSolution.f90:925-1011 (getRotationMatrices, getRotX/Y/Z)
void rotation_matrices(const double ang[3], Mat3 R, Mat3 Rinv) {
  double ax = -ang[0], ay = -ang[1], az = -ang[2];
  Mat3 X = {{1, 0, 0}, {0, std::cos(ax), -std::sin(ax)}, {0, std::sin(ax), std::cos(ax)}};
  Mat3 Y = {{std::cos(ay), 0, std::sin(ay)}, {0, 1, 0}, {-std::sin(ay), 0, std::cos(ay)}};
  Mat3 Z = {{std::cos(az), -std::sin(az), 0}, {std::sin(az), std::cos(az), 0}, {0, 0, 1}};
  Mat3 ZY, XtYt, Xt, Yt, Zt;
  matmul3(Z, Y, ZY);
  matmul3(ZY, X, R);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) { Xt[r][c] = X[c][r]; Yt[r][c] = Y[c][r]; Zt[r][c] = Z[c][r]; }
  matmul3(Xt, Yt, XtYt);
  matmul3(XtYt, Zt, Rinv);
}

// DemagFieldGetSolution.f90:833-896 (getSymmOpMatrices).  Every symm_M is a
// signed identity and every symm_H a diagonal mirror, so both are returned as
// the scalar sign c[j] and the diagonal s[j][3].
void symmetry_ops(const double ops[3], double c[8], double s[8][3]) {
  const double sx[3] = {-1, 1, 1}, sy[3] = {1, -1, 1}, sz[3] = {1, 1, -1};
  c[0] = 1; c[1] = ops[2]; c[2] = ops[1]; c[3] = ops[0];
  c[4] = c[2] * c[1]; c[5] = c[3] * c[1]; c[6] = c[3] * c[2]; c[7] = c[1] * (c[2] * c[3]);
  for (int k = 0; k < 3; ++k) {
    s[0][k] = 1; s[1][k] = sx[k]; s[2][k] = sy[k]; s[3][k] = sz[k];
    s[4][k] = sy[k] * sx[k]; s[5][k] = sz[k] * sx[k]; s[6][k] = sz[k] * sy[k];
    s[7][k] = sx[k] * (sy[k] * sz[k]);
  }
}

void k6_to_mat(const double k[6], Mat3 N) {
  N[0][0] = k[0]; N[0][1] = k[1]; N[0][2] = k[2];
  N[1][0] = k[1]; N[1][1] = k[3]; N[1][2] = k[4];
  N[2][0] = k[2]; N[2][1] = k[4]; N[2][2] = k[5];
}

struct PrismTile {
  double abc[3], off[3], M[3], ops[3];
  int sym;
  Mat3 R, Rinv;
  int type = 2;          // 2 prism, 5 tetrahedron
  double vert[4][3];     // tetrahedron vertices in the tile's frame
};

int tetra_n(const double vert[4][3], const double r[3], Mat3 N);

// tensor of the tile's own shape at the rotated relative position r
void shape_tensor(const PrismTile& T, const double r[3], Mat3 N) {
  if (T.type == 5) {
    if (tetra_n(T.vert, r, N))
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) N[a][b] = std::nan("");
    return;
  }
  double k[6];
  prism_k6(T.abc[0], T.abc[1], T.abc[2], r[0], r[1], r[2], k);
  k6_to_mat(k, N);
}

// getFieldFromTile (:685-730) and getFieldFromTile_symm (:744-822): tensor at p
void tile_tensor(const PrismTile& T, const double p[3], Mat3 N) {
  if (!T.sym) {
    double d[3] = {p[0] - T.off[0], p[1] - T.off[1], p[2] - T.off[2]}, r[3];
    for (int a = 0; a < 3; ++a) r[a] = T.R[a][0] * d[0] + T.R[a][1] * d[1] + T.R[a][2] * d[2];
    shape_tensor(T, r, N);
    return;
  }
  double c[8], s[8][3];
  symmetry_ops(T.ops, c, s);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) N[a][b] = 0.0;
  for (int j = 0; j < 8; ++j) {
    double d[3], r[3];
    Mat3 Nj;
    for (int a = 0; a < 3; ++a) d[a] = s[j][a] * p[a] - T.off[a];
    for (int a = 0; a < 3; ++a) r[a] = T.R[a][0] * d[0] + T.R[a][1] * d[1] + T.R[a][2] * d[2];
    shape_tensor(T, r, Nj);
    // N = symm_H * (N * symm_M): row a scaled by s[j][a], all scaled by c[j]
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) N[a][b] = s[j][a] * (Nj[a][b] * c[j]) + N[a][b];
  }
}

// ---- tetrahedron (tileType 5): sum over the four faces of the right-triangle
// face tensor (getN_tensor_tetrahedron, TileNComponents.f90:944-961;
// getN_Triangle and the Nxz / Nyz / Nzz integrals, TileTriangle.f90:26-286).
// FP64 throughout: the reference builds with default real = 8 bytes
// (-real-size 64 / -fdefault-real-8, python/src/magtense/lib/Makefile).
const double kTriPi = 3.14159265359;  // TileTriangle.f90:6, as written

double F_Nxz(const double r[3], double yp, double l, double h) {
  return h / std::sqrt(h * h + l * l) *
         std::atanh((l * l - l * r[0] + h * r[1] - h * yp * (1 + l * l / (h * h))) /
                    (std::sqrt(h * h + l * l) *
                     std::sqrt(l * l - 2 * l * r[0] + r[0] * r[0] + r[1] * r[1] - 2 * yp * (l * l - l * r[0] + h * r[1]) / h +
                               yp * yp * (1. + l * l / (h * h)) + r[2] * r[2])));
}
double G_Nxz(const double r[3], double yp) {
  double g = std::atanh((r[1] - yp) / (std::sqrt(r[0] * r[0] + (r[1] - yp) * (r[1] - yp) + r[2] * r[2])));
  return std::isfinite(g) ? g : 0.0;
}
double Nxz(const double r[3], double l, double h) {
  return -1. / (4. * kTriPi) * (F_Nxz(r, h, l, h) - F_Nxz(r, 0., l, h) - (G_Nxz(r, h) - G_Nxz(r, 0.)));
}
double K_Nyz(const double r[3], double xp, double l, double h) {
  return l / std::sqrt(h * h + l * l) *
         std::atanh((h * h - h * r[1] + l * r[0] - l * xp * (1 + h * h / (l * l))) /
                    (std::sqrt(h * h + l * l) *
                     std::sqrt(h * h - 2 * h * r[1] + r[0] * r[0] + r[1] * r[1] - 2 * xp * (h * h + l * r[0] - h * r[1]) / l +
                               xp * xp * (1 + h * h / (l * l)) + r[2] * r[2])));
}
double L_Nyz(const double r[3], double xp) {
  double g = std::atanh((r[0] - xp) / (std::sqrt((r[0] - xp) * (r[0] - xp) + r[1] * r[1] + r[2] * r[2])));
  return std::isfinite(g) ? g : 0.0;
}
double Nyz(const double r[3], double l, double h) {
  return -1. / (4. * kTriPi) * (K_Nyz(r, l, l, h) - K_Nyz(r, 0., l, h) - (L_Nyz(r, l) - L_Nyz(r, 0.)));
}
double P_Nzz(const double r[3], double xp, double l, double h) {
  return std::atan((r[0] * (h - r[1]) - xp * (h * (1 - r[0] / l) - r[1]) - h * (r[0] * r[0] + r[2] * r[2]) / l) /
                   (r[2] * std::sqrt(h * h + r[0] * r[0] + xp * xp * (1 + h * h / (l * l)) - 2 * h * r[1] + r[1] * r[1] -
                                     2 * xp * (h * h + l * r[0] - h * r[1]) / l + r[2] * r[2])));
}
double Q_Nzz(const double r[3], double xp) {
  return -std::atan((r[0] - xp) * r[1] / (r[2] * std::sqrt((r[0] - xp) * (r[0] - xp) + r[1] * r[1] + r[2] * r[2])));
}
double Nzz(const double r[3], double l, double h) {
  return -1. / (4. * kTriPi) * (P_Nzz(r, l, l, h) - P_Nzz(r, 0., l, h) - (Q_Nzz(r, l) - Q_Nzz(r, 0.)));
}

inline double norm3(const double a[3]) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
inline double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void sub3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2];
}
inline void cross3(const double a[3], const double b[3], double c[3]) {  // SpecialFunctions.f90:72-79
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// getN_Triangle (TileTriangle.f90:26-155): face v[0..2], v[3] behind the face;
// returns -1 for the reference's collinear / coplanar stops.
int triangle_n(const double vin[4][3], const double r[3], Mat3 N) {
  const double numErr = 1e-25, threshold = 1e-20;
  double v[4][3];
  std::memcpy(v, vin, sizeof(v));
  double t12[3], t23[3], t13[3];
  sub3(v[0], v[1], t12); sub3(v[1], v[2], t23); sub3(v[0], v[2], t13);
  if (std::fabs(norm3(t13) - (norm3(t12) + norm3(t23))) < numErr) return -1;
  double ang[3], a[3], b[3];
  sub3(v[0], v[1], a); sub3(v[0], v[2], b); ang[0] = std::acos(dot3(a, b) / (norm3(a) * norm3(b)));
  sub3(v[1], v[0], a); sub3(v[1], v[2], b); ang[1] = std::acos(dot3(a, b) / (norm3(a) * norm3(b)));
  sub3(v[2], v[1], a); sub3(v[2], v[0], b); ang[2] = std::acos(dot3(a, b) / (norm3(a) * norm3(b)));
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
  for (int i = 0; i < 3; ++i) std::memcpy(w[i], v[ind[i]], sizeof(w[i]));
  for (int i = 0; i < 3; ++i) std::memcpy(v[i], w[i], sizeof(w[i]));
  double tmp[3];
  std::memcpy(tmp, v[2], sizeof(tmp)); std::memcpy(v[2], v[1], sizeof(tmp)); std::memcpy(v[1], tmp, sizeof(tmp));
  ang[0] = srt[0]; ang[1] = srt[2]; ang[2] = srt[1];
  double e1[3], e2[3], e3[3], d32[3];
  auto frame = [&]() {
    sub3(v[0], v[2], e1);
    double n1 = norm3(e1);
    for (int k = 0; k < 3; ++k) e1[k] = e1[k] / n1;
    sub3(v[1], v[2], d32);
    cross3(e1, d32, e3);
    double n3 = norm3(e3);
    for (int k = 0; k < 3; ++k) e3[k] = e3[k] / n3;
  };
  frame();
  double d41[3];
  sub3(v[3], v[0], d41);
  if (std::fabs(dot3(d41, e3)) <= numErr) return -1;
  if (dot3(e3, d41) > 0.) {
    std::memcpy(tmp, v[0], sizeof(tmp)); std::memcpy(v[0], v[2], sizeof(tmp)); std::memcpy(v[2], tmp, sizeof(tmp));
    double t = ang[0]; ang[0] = ang[2]; ang[2] = t;
    frame();
  }
  cross3(e3, e1, e2);
  // P = [e1 e2 e3] (columns), Pinv = P^T
  Mat3 P = {{e1[0], e2[0], e3[0]}, {e1[1], e2[1], e3[1]}, {e1[2], e2[2], e3[2]}};
  sub3(v[1], v[2], d32);
  const double cd = std::cos(ang[2]) * norm3(d32);
  double D[3];
  for (int k = 0; k < 3; ++k) D[k] = cd * e1[k] + v[2][k];
  double vp[3][3], rt[3], q[3];
  for (int i = 0; i < 3; ++i) {
    sub3(v[i], D, q);
    for (int k = 0; k < 3; ++k) vp[i][k] = P[0][k] * q[0] + P[1][k] * q[1] + P[2][k] * q[2];
  }
  sub3(r, D, q);
  for (int k = 0; k < 3; ++k) rt[k] = P[0][k] * q[0] + P[1][k] * q[1] + P[2][k] * q[2];
  for (int k = 0; k < 3; ++k)
    if (std::fabs(rt[k]) < threshold) rt[k] = rt[k] < 0 ? -threshold : threshold;
  // local tensor: only column 3 (vp(1,1) = vp[0][0], vp(2,2) = vp[1][1], vp(1,3) = vp[2][0])
  const double n13 = Nxz(rt, vp[0][0], vp[1][1]) - Nxz(rt, vp[2][0], vp[1][1]);
  const double n23 = Nyz(rt, vp[0][0], vp[1][1]) - Nyz(rt, vp[2][0], vp[1][1]);
  const double n33 = Nzz(rt, vp[0][0], vp[1][1]) - Nzz(rt, vp[2][0], vp[1][1]);
  // N = (P N_loc) P^T, N_loc(:,3) = (n13, n23, n33)
  Mat3 PN = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int a2 = 0; a2 < 3; ++a2) PN[a2][2] = P[a2][0] * n13 + P[a2][1] * n23 + P[a2][2] * n33;
  for (int a2 = 0; a2 < 3; ++a2)
    for (int b2 = 0; b2 < 3; ++b2) N[a2][b2] = PN[a2][0] * P[b2][0] + PN[a2][1] * P[b2][1] + PN[a2][2] * P[b2][2];
  return 0;
}

// getN_tensor_tetrahedron: faces cshift(vert, i-1, 2), i = 1..4
int tetra_n(const double vert[4][3], const double r[3], Mat3 N) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) N[a][b] = 0.0;
  for (int i = 0; i < 4; ++i) {
    double v[4][3];
    for (int k = 0; k < 4; ++k) std::memcpy(v[k], vert[(k + i) % 4], sizeof(v[k]));
    Mat3 Nt;
    if (triangle_n(v, r, Nt)) return -1;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) N[a][b] = N[a][b] + Nt[a][b];
  }
  return 0;
}

int tetra_tensor_at(const double* vert12, const double r[3], double N9[9]) {
  double v[4][3];
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) v[k][c] = vert12[3 * k + c];
  Mat3 M;
  if (tetra_n(v, r, M)) return -1;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) N9[3 * a + b] = M[a][b];
  return 0;
}

// H contribution of one tile at p given its tensor (getDotProd + rotations)
void tile_field(const PrismTile& T, const Mat3 N, double h[3]) {
  double rm[3], dp[3];
  for (int a = 0; a < 3; ++a) rm[a] = T.R[a][0] * T.M[0] + T.R[a][1] * T.M[1] + T.R[a][2] * T.M[2];
  for (int a = 0; a < 3; ++a) dp[a] = N[a][0] * rm[0] + N[a][1] * rm[1] + N[a][2] * rm[2];
  for (int a = 0; a < 3; ++a) h[a] = T.Rinv[a][0] * dp[0] + T.Rinv[a][1] * dp[1] + T.Rinv[a][2] * dp[2];
}

int load_prism_tiles(int nt, const double* size, const double* offset, const double* rot, const double* M,
                     const int* tile_type, const int* use_sym, const double* sym_ops, std::vector<PrismTile>& T,
                     const double* vertices = nullptr) {
  T.resize(nt);
  for (int i = 0; i < nt; ++i) {
    const int type = tile_type ? tile_type[i] : 2;
    if (type != 2 && !(type == 5 && vertices)) return -2;
    PrismTile& t = T[i];
    t.type = type;
    if (type == 5)  // (nt,3,4) Fortran order
      for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 3; ++c) t.vert[k][c] = vertices[i + (size_t)nt * (c + 3 * (size_t)k)];
    double ang[3];
    for (int k = 0; k < 3; ++k) {
      t.abc[k] = size[i + (size_t)nt * k];
      t.off[k] = offset[i + (size_t)nt * k];
      t.M[k] = M ? M[i + (size_t)nt * k] : 0.0;
      t.ops[k] = sym_ops ? sym_ops[i + (size_t)nt * k] : 1.0;
      ang[k] = rot ? rot[i + (size_t)nt * k] : 0.0;
    }
    t.sym = use_sym ? (use_sym[i] == 1) : 0;
    rotation_matrices(ang, t.R, t.Rinv);
  }
  return 0;
}

}  // namespace

extern "C" {

// getRotationMatrices for one set of angles; R and Rinv row-major 3x3
void or_rotation_matrices(const double* ang, double* R, double* Rinv) {
  Mat3 a, b;
  rotation_matrices(ang, a, b);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) { R[3 * r + c] = a[r][c]; Rinv[3 * r + c] = b[r][c]; }
}

// getNFromTiles for prism tiles; returns -2 if a tile is not a prism
int or_tiles_tensor_v(int nt, const double* size, const double* offset, const double* rot, const int* tile_type,
                      const int* use_sym, const double* sym_ops, const double* vertices, int np, const double* pts,
                      double* N);
int or_tiles_tensor(int nt, const double* size, const double* offset, const double* rot, const int* tile_type,
                    const int* use_sym, const double* sym_ops, int np, const double* pts, double* N) {
  return or_tiles_tensor_v(nt, size, offset, rot, tile_type, use_sym, sym_ops, nullptr, np, pts, N);
}
// the same with tetrahedron tiles (tile_type 5, vertices (nt,3,4) Fortran order)
int or_tiles_tensor_v(int nt, const double* size, const double* offset, const double* rot, const int* tile_type,
                      const int* use_sym, const double* sym_ops, const double* vertices, int np, const double* pts,
                      double* N) {
  std::vector<PrismTile> T;
  int rc = load_prism_tiles(nt, size, offset, rot, nullptr, tile_type, use_sym, sym_ops, T, vertices);
  if (rc) return rc;
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < np; ++p) {
    double q[3] = {pts[p], pts[p + (size_t)np], pts[p + 2 * (size_t)np]};
    for (int i = 0; i < nt; ++i) {
      Mat3 Nm;
      tile_tensor(T[i], q, Nm);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) N[i + (size_t)nt * (p + (size_t)np * (a + 3 * b))] = Nm[a][b];
    }
  }
  return 0;
}

// getHFromTiles for prism tiles (H summed over tiles in tile order); N_in,
// when given, replaces the tensor evaluation (useStoredN)
int or_tiles_field_v(int nt, const double* size, const double* offset, const double* rot, const double* M,
                     const int* tile_type, const int* use_sym, const double* sym_ops, const double* vertices, int np,
                     const double* pts, const double* N_in, double* H);
int or_tiles_field(int nt, const double* size, const double* offset, const double* rot, const double* M,
                   const int* tile_type, const int* use_sym, const double* sym_ops, int np, const double* pts,
                   const double* N_in, double* H) {
  return or_tiles_field_v(nt, size, offset, rot, M, tile_type, use_sym, sym_ops, nullptr, np, pts, N_in, H);
}
int or_tiles_field_v(int nt, const double* size, const double* offset, const double* rot, const double* M,
                     const int* tile_type, const int* use_sym, const double* sym_ops, const double* vertices, int np,
                     const double* pts, const double* N_in, double* H) {
  std::vector<PrismTile> T;
  int rc = load_prism_tiles(nt, size, offset, rot, M, tile_type, use_sym, sym_ops, T, vertices);
  if (rc) return rc;
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < np; ++p) {
    double q[3] = {pts[p], pts[p + (size_t)np], pts[p + 2 * (size_t)np]};
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < nt; ++i) {
      Mat3 Nm;
      if (N_in) {
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) Nm[a][b] = N_in[i + (size_t)nt * (p + (size_t)np * (a + 3 * b))];
      } else {
        tile_tensor(T[i], q, Nm);
      }
      double h[3];
      tile_field(T[i], Nm, h);
      for (int a = 0; a < 3; ++a) acc[a] += h[a];
    }
    for (int a = 0; a < 3; ++a) H[p + (size_t)np * a] = acc[a];
  }
  return 0;
}

}  // extern "C"

extern "C" {

// FindThresholdFraction over the positive entries of a 6*n*n FP32 tensor
float or_threshold_fraction(const float* K, long long total, double fraction) {
  return threshold_fraction(K, (size_t)total, fraction);
}

}  // extern "C"

extern "C" {
// Receiver-averaged tensor of unstructured prisms (LandauLifshitzEquationSolver.f90:1044-1133),
// FP64, K6[c*n*n + i*n + j].
int or_tensor_unstructured(int n, const double* pts, const double* abc, const int* nave, double* K6) {
  if (n <= 0 || nave[0] < 1 || nave[1] < 1 || nave[2] < 1) return -1;
  std::vector<double> Kd;
  assemble_unstructured(n, pts, abc, nave, Kd);
  std::memcpy(K6, Kd.data(), sizeof(double) * Kd.size());
  return 0;
}
}  // extern "C"

extern "C" {
// Tetrahedron tensor at r (vertices vert[k*3 + c], k = 0..3), N row-major 3x3;
// -1 for a degenerate face (the reference stops).
int or_tetra_n(const double* vert, const double* r, double* N) {
  double v[4][3];
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) v[k][c] = vert[k * 3 + c];
  Mat3 M;
  if (tetra_n(v, r, M)) return -1;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) N[3 * a + b] = M[a][b];
  return 0;
}
}  // extern "C"

extern "C" {
// OpenMP threads used by the oracle's parallel loops (tensor assembly, FP32
// demag rows); returns the previous setting.
int or_set_threads(int n) {
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
}
}  // extern "C"

extern "C" {
// interp1_MagTense (util.f90:39-56) of one value on a table of n points;
// *bracket receives locate's index (util.f90:59-85, 1-based; 0 left of the
// table).  Exposed so the table lookup is checked on its own.
double or_interp1(int n, const double* x, const double* y, double xv, int* bracket) {
  std::vector<double> xs(x, x + n), ys(y, y + n);
  if (bracket) *bracket = locate(x, n, xv);
  return interp1c(xs, ys, xv);
}
}  // extern "C"
