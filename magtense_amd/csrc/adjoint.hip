// Differentiable LL integration: fixed-step RK4 forward pass and its
// hand-written discrete adjoint, with checkpointing and recomputation.
//
// There is no reference counterpart (MagTense integrates with RKSUITE and has
// no gradients); the right-hand side differentiated here is the reference's
// dmdt_fct (LandauLifshitzEquationSolver.f90:220-273) with its field terms
// (exchange :440-474, external :485-514, anisotropy :523-548, demag scaled by
// the receiver cell's Ms :749-754).
//
// With negated fields H (as the reference stores them),
//   c = -(m x H),  f = -gamma c - alpha (m x c),  H = H_h + H_j(m) + H_m(m) + H_k(m)
// the vector-Jacobian product for a cotangent u of f is
//   cbar = -gamma u - alpha (u x m)
//   mbar = -alpha (c x u) - (H x cbar) + L^T Hbar,   Hbar = -(cbar x m)
// where L^T Hbar = -2 Lap(Jfact Hbar) - K(Ms Hbar) - 2 A_ani(Kfact Hbar) (the
// Laplacian and the demag blocks are symmetric), and the parameter gradients
// are dL/dH_ext = -sum_i Hbar_i and dL/dMs_i = Hbar_i . (H_m - H_j - H_k)_i / Ms_i.
// For RK4, y' = y + h/6 (k1 + 2k2 + 2k3 + k4), the adjoint of one step is
//   kbar4 = h/6 lam,            mu4 = vjp(Y4, kbar4)
//   kbar3 = h/3 lam + h mu4,    mu3 = vjp(Y3, kbar3)
//   kbar2 = h/3 lam + h/2 mu3,  mu2 = vjp(Y2, kbar2)
//   kbar1 = h/6 lam + h/2 mu2,  mu1 = vjp(Y1, kbar1)
//   lam  <- lam + mu1 + mu2 + mu3 + mu4.

#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "mt_common.h"
#include "mt_kernels.h"
#include "../../include/magtense_hip.h"

namespace mt {

namespace {

constexpr double kMu0 = 4.0 * 3.141592653589793 * 1e-7;

struct VjpDesc {
  const double* m;      // stage input Y
  const void* S;        // K Y (positive demag sums), FP32 or FP64
  const double* u;      // cotangent of k = f(Y)
  double* mbar;         // out: direct part of the input cotangent
  double* hbar;         // out: field cotangent
  void* packed;         // out: (Ms * Hbar) for the transpose demag pass, FP32 or FP64
  double* grad_ms;      // accumulate: per-cell Ms gradient
  double* hext_part;    // accumulate: per-block partial sums of -Hbar (3 per block)
  double hext[3];
  double alpha, gamma;
};

__device__ inline double lap_at(const double* v, int i, int ix, int iy, int iz, const LlgGrid& g) {
  const int sy = g.nx, sz = g.nx * g.ny;
  double vi = v[i];
  double lx = ix == 0 ? (v[i + 1] - vi) : (ix == g.nx - 1 ? (v[i - 1] - vi) : (v[i - 1] - 2.0 * vi + v[i + 1]));
  double ly = iy == 0 ? (v[i + sy] - vi) : (iy == g.ny - 1 ? (v[i - sy] - vi) : (v[i - sy] - 2.0 * vi + v[i + sy]));
  double lap = lx * g.idx2 + ly * g.idy2;
  if (g.nz > 1) {
    double lz = iz == 0 ? (v[i + sz] - vi) : (iz == g.nz - 1 ? (v[i - sz] - vi) : (v[i - sz] - 2.0 * vi + v[i + sz]));
    lap += lz * g.idz2;
  }
  return lap;
}

// T: demag precision (float: the reference's FP32 product and FP32 Ms scaling;
// double: FP64 tensor, product and Ms)
template <typename T>
__global__ void __launch_bounds__(256) vjp_kernel(const LlgGrid g, const VjpDesc* __restrict__ descs) {
  const VjpDesc& D = descs[blockIdx.y];
  const int n = g.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double hx_part[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const int ix = i % g.nx, iy = (i / g.nx) % g.ny, iz = i / (g.nx * g.ny);
    double m[3], u[3], hj[3], hm[3], hk[3], H[3];
    for (int c = 0; c < 3; ++c) {
      m[c] = D.m[(size_t)c * n + i];
      u[c] = D.u[(size_t)c * n + i];
      hj[c] = (-2.0 * lap_at(D.m + (size_t)c * n, i, ix, iy, iz, g)) * g.jfact[i];
      const T sv = static_cast<const T*>(D.S)[(size_t)c * n + i];
      if constexpr (sizeof(T) == 8) {
        hm[c] = (-sv) * g.ms64[i];
      } else {
        float hf = (-sv) * g.ms32[i];
        hm[c] = (double)hf;
      }
    }
    if (g.aniso) {
      const double* A = g.aniso + (size_t)i * 6;
      double f = -2.0 * g.kfact[i];
      hk[0] = f * (A[0] * m[0] + A[1] * m[1] + A[2] * m[2]);
      hk[1] = f * (A[1] * m[0] + A[3] * m[1] + A[4] * m[2]);
      hk[2] = f * (A[2] * m[0] + A[4] * m[1] + A[5] * m[2]);
    } else {
      hk[0] = hk[1] = hk[2] = 0.0;
    }
    for (int c = 0; c < 3; ++c) H[c] = -D.hext[c] + hj[c] + hm[c] + hk[c];
    // c = -(m x H)
    double cv[3] = {-(m[1] * H[2] - m[2] * H[1]), -(m[2] * H[0] - m[0] * H[2]), -(m[0] * H[1] - m[1] * H[0])};
    // cbar = -gamma u - alpha (u x m)
    double uxm[3] = {u[1] * m[2] - u[2] * m[1], u[2] * m[0] - u[0] * m[2], u[0] * m[1] - u[1] * m[0]};
    double cb[3];
    for (int c = 0; c < 3; ++c) cb[c] = -D.gamma * u[c] - D.alpha * uxm[c];
    // mbar_direct = -alpha (c x u) - (H x cbar)
    double cxu[3] = {cv[1] * u[2] - cv[2] * u[1], cv[2] * u[0] - cv[0] * u[2], cv[0] * u[1] - cv[1] * u[0]};
    double hxc[3] = {H[1] * cb[2] - H[2] * cb[1], H[2] * cb[0] - H[0] * cb[2], H[0] * cb[1] - H[1] * cb[0]};
    // Hbar = -(cbar x m)
    double hb[3] = {-(cb[1] * m[2] - cb[2] * m[1]), -(cb[2] * m[0] - cb[0] * m[2]), -(cb[0] * m[1] - cb[1] * m[0])};
    double ms = sizeof(T) == 8 ? g.ms64[i] : (double)g.ms32[i];
    double gms = 0.0;
    for (int c = 0; c < 3; ++c) {
      D.mbar[(size_t)c * n + i] = -D.alpha * cxu[c] - hxc[c];
      D.hbar[(size_t)c * n + i] = hb[c];
      static_cast<T*>(D.packed)[(size_t)c * n + i] = (T)(ms * hb[c]);
      gms += hb[c] * (hm[c] - hj[c] - hk[c]);
      hx_part[c] = -hb[c];
    }
    D.grad_ms[i] += gms / ms;
  }
  // block partial of -sum Hbar for the applied-field gradient
  __shared__ double red[4][3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = 0; c < 3; ++c) {
    double x = hx_part[c];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) red[wave][c] = x;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double x = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
    D.hext_part[blockIdx.x * 3 + threadIdx.x] += x;
  }
}

struct FinishDesc {
  const double* mbar;   // direct part
  const double* hbar;   // field cotangent
  const void* T;        // K (Ms Hbar), FP32 or FP64
  double* mu;           // out: vjp w.r.t. the stage input
};

template <typename TT>
__global__ void __launch_bounds__(256) vjp_finish_kernel(const LlgGrid g, const FinishDesc* __restrict__ descs) {
  const FinishDesc& D = descs[blockIdx.y];
  const int n = g.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ix = i % g.nx, iy = (i / g.nx) % g.ny, iz = i / (g.nx * g.ny);
  double hb[3];
  for (int c = 0; c < 3; ++c) hb[c] = D.hbar[(size_t)c * n + i];
  double ak[3] = {0.0, 0.0, 0.0};
  if (g.aniso) {
    const double* A = g.aniso + (size_t)i * 6;
    double f = -2.0 * g.kfact[i];
    ak[0] = f * (A[0] * hb[0] + A[1] * hb[1] + A[2] * hb[2]);
    ak[1] = f * (A[1] * hb[0] + A[3] * hb[1] + A[4] * hb[2]);
    ak[2] = f * (A[2] * hb[0] + A[4] * hb[1] + A[5] * hb[2]);
  }
  // exchange transpose: -2 Lap(Jfact * Hbar); the weighted field is formed on the fly
  const double* jf = g.jfact;
  for (int c = 0; c < 3; ++c) {
    const double* h = D.hbar + (size_t)c * n;
    auto w = [&](int k) { return jf[k] * h[k]; };
    const int sy = g.nx, sz = g.nx * g.ny;
    double wi = w(i);
    double lx = ix == 0 ? (w(i + 1) - wi) : (ix == g.nx - 1 ? (w(i - 1) - wi) : (w(i - 1) - 2.0 * wi + w(i + 1)));
    double ly = iy == 0 ? (w(i + sy) - wi) : (iy == g.ny - 1 ? (w(i - sy) - wi) : (w(i - sy) - 2.0 * wi + w(i + sy)));
    double lap = lx * g.idx2 + ly * g.idy2;
    if (g.nz > 1) {
      double lz = iz == 0 ? (w(i + sz) - wi) : (iz == g.nz - 1 ? (w(i - sz) - wi) : (w(i - sz) - 2.0 * wi + w(i + sz)));
      lap += lz * g.idz2;
    }
    const double tv = (double)static_cast<const TT*>(D.T)[(size_t)c * n + i];
    D.mu[(size_t)c * n + i] = D.mbar[(size_t)c * n + i] - 2.0 * lap - tv + ak[c];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
struct Rk4 {
  int nx = 0, ny = 0, nz = 0, n = 0, B = 0;
  double dt = 0;
  long long nsteps = 0;
  int ck = 100;
  double gamma = 0, alpha = 0;
  std::vector<double> hext;  // B*3
  hipStream_t stream = nullptr;
  LlgGrid grid{};
  bool fp64 = false;       // FP64 tensor, product and Ms scaling
  size_t es = 4;           // bytes per demag element (4 or 8)
  float* K32 = nullptr;
  double* K64 = nullptr;
  double *jfact = nullptr, *kfact = nullptr, *aniso = nullptr;
  float* ms32 = nullptr;
  double* ms64 = nullptr;
  void* packd = nullptr;   // B x 3n demag input (FP32 or FP64)
  void* hbuf = nullptr;    // B x 3n demag sums (FP32 or FP64)
  double* ckpt = nullptr;  // (nck+1) x B x 3n
  double* seg = nullptr;   // ck+1 steps x B x (y + 4 stage inputs) x 3n
  uint8_t* segS = nullptr; // ck steps x 4 stages x B x 3n demag sums
  double* work = nullptr;  // B x 3n x 12 scratch vectors
  double* grad_ms = nullptr;   // B x n
  double* hext_part = nullptr; // B x nblocks x 3
  static constexpr size_t kSlotBytes = 65536;
  static constexpr int kSlots = 256;
  uint8_t* desc_d = nullptr;
  uint8_t* desc_h = nullptr;  // pinned
  int ring_next = 0;
  size_t desc_size = 0;
  // Descriptor cache: the fixed-step loops issue the same descriptor blocks
  // step after step (same vectors, same coefficients), so each distinct block
  // is copied to device memory once and reused; a copy is an SDMA transfer the
  // compute queue waits for, about 20 us each.
  static constexpr size_t kCacheBytes = (size_t)32 << 20;
  uint8_t* dcache_d = nullptr;
  size_t dcache_used = 0;
  std::unordered_map<std::string, size_t> dcache;
  int nck = 0;
  bool forward_done = false;

  ~Rk4() {
    auto f = [](void* p) {
      if (p) (void)hipFree(p);
    };
    f(K32); f(K64); f(jfact); f(kfact); f(aniso); f(ms32); f(ms64); f(packd); f(hbuf); f(ckpt); f(seg); f(segS);
    f(work);
    f(grad_ms); f(hext_part); f(desc_d); f(dcache_d);
    if (desc_h) (void)hipHostFree(desc_h);
    if (stream) (void)hipStreamDestroy(stream);
  }

  size_t n3() const { return (size_t)3 * n; }
  void* hcol(void* base, int b) const { return static_cast<uint8_t*>(base) + (size_t)b * n3() * es; }

  // out (default hbuf) = K * packd for all B states, columns n3 apart
  int demag_pass(void* out = nullptr) {
    if (!out) out = hbuf;
    if (fp64)
      return launch_matvec_f64(K64, n, static_cast<const double*>(packd), (int64_t)n3(), static_cast<double*>(out),
                               (int64_t)n3(), B, stream);
    return launch_matvec_f32(K32, n, static_cast<const float*>(packd), (int64_t)n3(), static_cast<float*>(out),
                             (int64_t)n3(), B, stream);
  }
  int nblocks() const { return (n + 255) / 256; }

  // Descriptors go through a ring of pinned host slots copied asynchronously;
  // the stream is synchronised only when the ring wraps, so a slot is never
  // rewritten while its copy may still be pending.
  template <typename T>
  int upload(const std::vector<T>& v, const T** out, size_t /*kind*/) {
    size_t bytes = v.size() * sizeof(T);
    if (bytes > kSlotBytes) {
      set_error("descriptor block too large");
      return kErrArg;
    }
    std::string key(reinterpret_cast<const char*>(v.data()), bytes);
    auto hit = dcache.find(key);
    if (hit != dcache.end()) {
      *out = reinterpret_cast<const T*>(dcache_d + hit->second);
      return kOk;
    }
    if (ring_next == kSlots) {
      MT_HIP(hipStreamSynchronize(stream));
      ring_next = 0;
    }
    size_t off = (size_t)ring_next++ * kSlotBytes;
    std::memcpy(desc_h + off, v.data(), bytes);
    const size_t cbytes = (bytes + 255) / 256 * 256;
    if (dcache_d && dcache_used + cbytes <= kCacheBytes) {
      // cached blocks are written once and only read afterwards
      MT_HIP(hipMemcpyAsync(dcache_d + dcache_used, desc_h + off, bytes, hipMemcpyHostToDevice, stream));
      *out = reinterpret_cast<const T*>(dcache_d + dcache_used);
      dcache.emplace(std::move(key), dcache_used);
      dcache_used += cbytes;
      return kOk;
    }
    MT_HIP(hipMemcpyAsync(desc_d + off, desc_h + off, bytes, hipMemcpyHostToDevice, stream));
    *out = reinterpret_cast<const T*>(desc_d + off);
    return kOk;
  }

  // k[b] = f(Y[b]) for every state; S (K Y) is written to S_out[b] when given.
  // packed: the FP32 demag input already holds Y (packed by the combination
  // that formed it).
  int rhs(double* const* Y, double* const* k, void* const* S_out, bool packed = false) {
    int rc = packed && !fp64 ? kOk : pack_input(Y);
    if (rc) return rc;
    // recording: when the states' S_out columns are laid out like the demag
    // buffer's, the product is written there directly instead of copied
    bool direct = S_out != nullptr;
    for (int b = 1; b < B && direct; ++b)
      direct = static_cast<uint8_t*>(S_out[b]) == static_cast<uint8_t*>(S_out[0]) + (size_t)b * n3() * es;
    void* out = direct ? S_out[0] : hbuf;
    rc = demag_pass(out);
    if (rc) return rc;
    return llg(Y, k, direct ? nullptr : S_out, out);
  }

  int pack_input(double* const* Y) {
    std::vector<VecDesc> vd(B);
    for (int b = 0; b < B; ++b) {
      VecDesc& d = vd[b];
      std::memset(&d, 0, sizeof(d));
      // FP32: identity combination used for the pack only; FP64: copy into the contiguous input
      d.dst = fp64 ? static_cast<double*>(hcol(packd, b)) : Y[b];
      d.base = Y[b];
      d.bcoef = 1.0;
      d.post_scale = 1.0;
      d.pack_col = fp64 ? -1 : b;
    }
    const VecDesc* dvd = nullptr;
    int rc = upload(vd, &dvd, 0);
    if (rc) return rc;
    PackSpec pk{fp64 ? kPackNone : kPackF32, fp64 ? nullptr : static_cast<float*>(packd), (int64_t)n3(), nullptr};
    return launch_vec(dvd, B, n, nullptr, nullptr, pk, false, stream);
  }

  int llg(double* const* Y, double* const* k, void* const* S_out, void* sums) {
    std::vector<LlgDesc> ld(B);
    for (int b = 0; b < B; ++b) {
      LlgDesc& L = ld[b];
      std::memset(&L, 0, sizeof(L));
      L.m = Y[b];
      L.out = k[b];
      if (fp64) L.hpart64 = static_cast<const double*>(hcol(sums, b));
      else L.hpart = static_cast<const float*>(hcol(sums, b));
      L.nsplit = 1;
      for (int c = 0; c < 3; ++c) L.hext[c] = hext[b * 3 + c];
      L.alpha = alpha;
      L.gamma = gamma;
    }
    const LlgDesc* dl = nullptr;
    int rc = upload(ld, &dl, 1);
    if (rc) return rc;
    rc = launch_llg(grid, dl, B, stream);
    if (rc) return rc;
    if (S_out)
      for (int b = 0; b < B; ++b)
        MT_HIP(hipMemcpyAsync(S_out[b], hcol(sums, b), n3() * es, hipMemcpyDeviceToDevice, stream));
    return kOk;
  }

  // dst[b] = base[b] + sum_k c_k src_k[b]; pack: also write dst into the FP32
  // demag input (not in the FP64 mode, whose input is a plain copy)
  int combine(double* const* dst, double* const* base, std::vector<std::pair<double* const*, double>> terms,
              double bcoef = 1.0, bool pack = false) {
    pack = pack && !fp64;
    std::vector<VecDesc> vd(B);
    for (int b = 0; b < B; ++b) {
      VecDesc& d = vd[b];
      std::memset(&d, 0, sizeof(d));
      d.dst = dst[b];
      d.base = base ? base[b] : nullptr;
      d.bcoef = bcoef;
      d.post_scale = 1.0;
      d.pack_col = pack ? b : -1;
      for (auto& t : terms) {
        d.src[d.nterm] = t.first[b];
        d.coef[d.nterm] = t.second;
        d.nterm += 1;
      }
    }
    const VecDesc* dvd = nullptr;
    int rc = upload(vd, &dvd, 2);
    if (rc) return rc;
    PackSpec pk{pack ? kPackF32 : kPackNone, pack ? static_cast<float*>(packd) : nullptr, pack ? (int64_t)n3() : 0,
                nullptr};
    return launch_vec(dvd, B, n, nullptr, nullptr, pk, false, stream);
  }

  // pointer tables
  std::vector<double*> vecs(double* base, size_t stride) {
    std::vector<double*> v(B);
    for (int b = 0; b < B; ++b) v[b] = base + (size_t)b * stride;
    return v;
  }

  double* seg_y(long long j) { return seg + (size_t)j * B * 5 * n3(); }                        // B x 3n
  double* seg_stage(long long j, int s) { return seg_y(j) + (size_t)(1 + s) * B * n3(); }    // stage input s
  uint8_t* seg_S(long long j, int s) { return segS + ((size_t)j * 4 + s) * B * n3() * es; }

  // One RK4 step from y (B vectors) to ynext; records stage inputs / demag sums when asked.
  int step(double* const* y, double* const* ynext, double* const* Yst /*4*B*/, void* const* Sst /*4*B*/) {
    const size_t N3 = n3();
    auto W = [&](int k) { return vecs(work + (size_t)k * B * N3, N3); };
    auto k1 = W(0), k2 = W(1), k3 = W(2), k4 = W(3);
    std::vector<double*> Y1(y, y + B), Y2 = Yst ? std::vector<double*>(Yst + B, Yst + 2 * B) : W(5),
                                       Y3 = Yst ? std::vector<double*>(Yst + 2 * B, Yst + 3 * B) : W(6),
                                       Y4 = Yst ? std::vector<double*>(Yst + 3 * B, Yst + 4 * B) : W(7);
    int rc;
    if (Yst) {
      rc = combine(Yst, Y1.data(), {});
      if (rc) return rc;
    }
    rc = rhs(Y1.data(), k1.data(), Sst ? Sst : nullptr);
    if (rc) return rc;
    // each stage input is packed for its demag pass by the combination that forms it
    rc = combine(Y2.data(), Y1.data(), {{k1.data(), 0.5 * dt}}, 1.0, true);
    if (rc) return rc;
    rc = rhs(Y2.data(), k2.data(), Sst ? Sst + B : nullptr, true);
    if (rc) return rc;
    rc = combine(Y3.data(), Y1.data(), {{k2.data(), 0.5 * dt}}, 1.0, true);
    if (rc) return rc;
    rc = rhs(Y3.data(), k3.data(), Sst ? Sst + 2 * B : nullptr, true);
    if (rc) return rc;
    rc = combine(Y4.data(), Y1.data(), {{k3.data(), dt}}, 1.0, true);
    if (rc) return rc;
    rc = rhs(Y4.data(), k4.data(), Sst ? Sst + 3 * B : nullptr, true);
    if (rc) return rc;
    // y' = y + h/6 (k1 + 2 k2 + 2 k3 + k4); the kernel reads every operand
    // element before writing it, so ynext may be y itself
    return combine(ynext, Y1.data(),
                   {{k1.data(), dt / 6.0}, {k2.data(), dt / 3.0}, {k3.data(), dt / 3.0}, {k4.data(), dt / 6.0}});
  }

  int forward(const double* m0_host, double* mT_host) {
    const size_t N3 = n3();
    MT_HIP(hipMemcpyAsync(ckpt, m0_host, (size_t)B * N3 * sizeof(double), hipMemcpyHostToDevice, stream));
    auto cur = vecs(work + (size_t)8 * B * N3, N3);
    auto nxt = vecs(work + (size_t)9 * B * N3, N3);
    auto ck0 = vecs(ckpt, N3);
    int rc = combine(cur.data(), ck0.data(), {});
    if (rc) return rc;
    for (long long s = 0; s < nsteps; ++s) {
      rc = step(cur.data(), nxt.data(), nullptr, nullptr);
      if (rc) return rc;
      std::swap(cur, nxt);
      if ((s + 1) % ck == 0 && (s + 1) < nsteps) {
        auto cp = vecs(ckpt + (size_t)((s + 1) / ck) * B * N3, N3);
        rc = combine(cp.data(), cur.data(), {});
        if (rc) return rc;
      }
    }
    if (mT_host)
      for (int b = 0; b < B; ++b)
        MT_HIP(hipMemcpyAsync(mT_host + (size_t)b * N3, cur[b], N3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    MT_HIP(hipStreamSynchronize(stream));
    forward_done = true;
    return kOk;
  }

  // mu[b] = vjp of f at stage input Y with cotangent u; accumulates parameter gradients
  int vjp(double* const* Y, void* const* S, double* const* u, double* const* mu) {
    const size_t N3 = n3();
    auto mbar = vecs(work + (size_t)10 * B * N3, N3);
    auto hbar = vecs(work + (size_t)11 * B * N3, N3);
    std::vector<VjpDesc> vd(B);
    for (int b = 0; b < B; ++b) {
      VjpDesc& d = vd[b];
      d.m = Y[b];
      d.S = S[b];
      d.u = u[b];
      d.mbar = mbar[b];
      d.hbar = hbar[b];
      d.packed = hcol(packd, b);
      d.grad_ms = grad_ms + (size_t)b * n;
      d.hext_part = hext_part + (size_t)b * nblocks() * 3;
      for (int c = 0; c < 3; ++c) d.hext[c] = hext[b * 3 + c];
      d.alpha = alpha;
      d.gamma = gamma;
    }
    const VjpDesc* dvd = nullptr;
    int rc = upload(vd, &dvd, 3);
    if (rc) return rc;
    if (fp64) hipLaunchKernelGGL(vjp_kernel<double>, dim3(nblocks(), B), dim3(256), 0, stream, grid, dvd);
    else hipLaunchKernelGGL(vjp_kernel<float>, dim3(nblocks(), B), dim3(256), 0, stream, grid, dvd);
    MT_HIP(hipGetLastError());
    rc = demag_pass();
    if (rc) return rc;
    std::vector<FinishDesc> fd(B);
    for (int b = 0; b < B; ++b) {
      fd[b].mbar = mbar[b];
      fd[b].hbar = hbar[b];
      fd[b].T = hcol(hbuf, b);
      fd[b].mu = mu[b];
    }
    const FinishDesc* dfd = nullptr;
    rc = upload(fd, &dfd, 4);
    if (rc) return rc;
    if (fp64) hipLaunchKernelGGL(vjp_finish_kernel<double>, dim3(nblocks(), B), dim3(256), 0, stream, grid, dfd);
    else hipLaunchKernelGGL(vjp_finish_kernel<float>, dim3(nblocks(), B), dim3(256), 0, stream, grid, dfd);
    MT_HIP(hipGetLastError());
    return kOk;
  }

  int backward(const double* gT_host, double* grad_m0, double* grad_hext_out, double* grad_ms_out) {
    if (!forward_done) {
      set_error("mt_rk4_backward before mt_rk4_forward");
      return kErrState;
    }
    const size_t N3 = n3();
    MT_HIP(hipMemsetAsync(grad_ms, 0, (size_t)B * n * sizeof(double), stream));
    MT_HIP(hipMemsetAsync(hext_part, 0, (size_t)B * nblocks() * 3 * sizeof(double), stream));
    auto lam = vecs(work + (size_t)12 * B * N3, N3);
    auto W = [&](int k) { return vecs(work + (size_t)k * B * N3, N3); };
    auto kb = W(0), mu4 = W(1), mu3 = W(2), mu2 = W(3), mu1 = W(4);
    for (int b = 0; b < B; ++b)
      MT_HIP(hipMemcpyAsync(lam[b], gT_host + (size_t)b * N3, N3 * sizeof(double), hipMemcpyHostToDevice, stream));
    int rc;
    for (int seg_i = nck - 1; seg_i >= 0; --seg_i) {
      long long s0 = (long long)seg_i * ck;
      long long s1 = std::min(nsteps, s0 + ck);
      long long len = s1 - s0;
      // recompute the segment, recording stage inputs and their demag sums
      auto y0 = vecs(seg_y(0), N3);
      auto cp = vecs(ckpt + (size_t)seg_i * B * N3, N3);
      rc = combine(y0.data(), cp.data(), {});
      if (rc) return rc;
      for (long long j = 0; j < len; ++j) {
        std::vector<double*> Yst(4 * B);
        std::vector<void*> Sst(4 * B);
        for (int s = 0; s < 4; ++s)
          for (int b = 0; b < B; ++b) {
            Yst[s * B + b] = seg_stage(j, s) + (size_t)b * N3;
            Sst[s * B + b] = seg_S(j, s) + (size_t)b * N3 * es;
          }
        auto yj = vecs(seg_y(j), N3);
        auto yn = vecs(seg_y(j + 1), N3);
        rc = step(yj.data(), yn.data(), Yst.data(), Sst.data());
        if (rc) return rc;
      }
      // backward through the segment
      for (long long j = len - 1; j >= 0; --j) {
        auto st = [&](int s) {
          std::vector<double*> v(B);
          for (int b = 0; b < B; ++b) v[b] = seg_stage(j, s) + (size_t)b * N3;
          return v;
        };
        auto sS = [&](int s) {
          std::vector<void*> v(B);
          for (int b = 0; b < B; ++b) v[b] = seg_S(j, s) + (size_t)b * N3 * es;
          return v;
        };
        auto Y1 = st(0), Y2 = st(1), Y3 = st(2), Y4 = st(3);
        auto S1 = sS(0), S2 = sS(1), S3 = sS(2), S4 = sS(3);
        rc = combine(kb.data(), nullptr, {{lam.data(), dt / 6.0}});
        if (rc) return rc;
        rc = vjp(Y4.data(), S4.data(), kb.data(), mu4.data());
        if (rc) return rc;
        rc = combine(kb.data(), nullptr, {{lam.data(), dt / 3.0}, {mu4.data(), dt}});
        if (rc) return rc;
        rc = vjp(Y3.data(), S3.data(), kb.data(), mu3.data());
        if (rc) return rc;
        rc = combine(kb.data(), nullptr, {{lam.data(), dt / 3.0}, {mu3.data(), 0.5 * dt}});
        if (rc) return rc;
        rc = vjp(Y2.data(), S2.data(), kb.data(), mu2.data());
        if (rc) return rc;
        rc = combine(kb.data(), nullptr, {{lam.data(), dt / 6.0}, {mu2.data(), 0.5 * dt}});
        if (rc) return rc;
        rc = vjp(Y1.data(), S1.data(), kb.data(), mu1.data());
        if (rc) return rc;
        rc = combine(lam.data(), lam.data(), {{mu1.data(), 1.0}, {mu2.data(), 1.0}, {mu3.data(), 1.0}, {mu4.data(), 1.0}});
        if (rc) return rc;
      }
    }
    if (grad_m0)
      for (int b = 0; b < B; ++b)
        MT_HIP(hipMemcpyAsync(grad_m0 + (size_t)b * N3, lam[b], N3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (grad_ms_out)
      MT_HIP(hipMemcpyAsync(grad_ms_out, grad_ms, (size_t)B * n * sizeof(double), hipMemcpyDeviceToHost, stream));
    std::vector<double> hp((size_t)B * nblocks() * 3);
    MT_HIP(hipMemcpyAsync(hp.data(), hext_part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    MT_HIP(hipStreamSynchronize(stream));
    if (grad_hext_out)
      for (int b = 0; b < B; ++b)
        for (int c = 0; c < 3; ++c) {
          double s = 0.0;
          for (int k = 0; k < nblocks(); ++k) s += hp[((size_t)b * nblocks() + k) * 3 + c];
          grad_hext_out[b * 3 + c] = s;
        }
    return kOk;
  }
};

}  // namespace mt

using mt::Rk4;

extern "C" {

int mt_rk4_create(const mt_micromag_problem* q, double dt, long long nsteps, int checkpoint_every, mt_rk4** out) {
  using namespace mt;
  if (!q || !out) {
    set_error("null argument");
    return kErrArg;
  }
  *out = nullptr;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) {
    set_error("no HIP device visible: the MI355X backend needs a GPU");
    return kErrHip;
  }
  MT_CHECK(q->grid_type == 0 || q->grid_type == 1, "the RK4 adjoint runs on uniform grids (grid_type 1)");
  MT_CHECK(q->nx >= 2 && q->ny >= 2 && q->nz >= 1, "the uniform-grid exchange operator needs nx >= 2 and ny >= 2");
  MT_CHECK(q->n_states >= 1 && q->n_states <= 1024, "n_states must be in [1, 1024]");
  MT_CHECK(dt > 0 && nsteps >= 1 && checkpoint_every >= 1, "need dt > 0, nsteps >= 1, checkpoint_every >= 1");
  MT_CHECK(q->Ms && q->hext && q->nt_hext >= 1, "Ms and an applied field are required");
  MT_CHECK(q->precision == MT_PREC_FP32 || q->precision == MT_PREC_FP64,
           "the RK4 adjoint runs the demag product in FP32 (MT_PREC_FP32) or FP64 (MT_PREC_FP64)");
  MT_CHECK(q->precision == MT_PREC_FP32 || !q->K_in, "K_in is an FP32 tensor: use MT_PREC_FP32 with it");
  std::unique_ptr<Rk4> R(new Rk4());
  Rk4& P = *R;
  P.nx = q->nx; P.ny = q->ny; P.nz = q->nz; P.n = q->nx * q->ny * q->nz; P.B = q->n_states;
  P.dt = dt; P.nsteps = nsteps; P.ck = checkpoint_every;
  P.gamma = q->gamma; P.alpha = q->alpha0;
  P.fp64 = q->precision == MT_PREC_FP64;
  P.es = P.fp64 ? sizeof(double) : sizeof(float);
  P.hext.resize((size_t)P.B * 3);
  for (int b = 0; b < P.B; ++b) {
    const double* row = q->hext + (size_t)(q->hext_per_state ? b : 0) * q->nt_hext * 4;
    for (int c = 0; c < 3; ++c) P.hext[b * 3 + c] = row[1 + c];  // constant field: first table row
  }
  const int n = P.n;
  const size_t N3 = P.n3();
  double dx = q->nx > 1 ? q->Lx / q->nx : q->Lx, dy = q->ny > 1 ? q->Ly / q->ny : q->Ly,
         dz = q->nz > 1 ? q->Lz / q->nz : q->Lz;
  MT_HIP(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
  // tensor (FP32, receiver-major) from the assembly kernel or the caller
  const size_t nn = (size_t)n * n;
  if (P.fp64) MT_HIP(hipMalloc(&P.K64, 6 * nn * sizeof(double)));
  else MT_HIP(hipMalloc(&P.K32, 6 * nn * sizeof(float)));
  if (q->K_in) {
    MT_HIP(hipMemcpyAsync(P.K32, q->K_in, 6 * nn * sizeof(float), hipMemcpyHostToDevice, P.stream));
  } else {
    std::vector<double> pts((size_t)n * 3);
    auto axis = [](int m, double L, double d, int i) { return m > 1 ? -L / 2 + i * d + d / 2 : 0.0; };
    for (int k = 0; k < P.nz; ++k)
      for (int j = 0; j < P.ny; ++j)
        for (int i = 0; i < P.nx; ++i) {
          size_t ind = (size_t)i + (size_t)j * P.nx + (size_t)k * P.nx * P.ny;
          pts[ind * 3] = axis(P.nx, q->Lx, dx, i);
          pts[ind * 3 + 1] = axis(P.ny, q->Ly, dy, j);
          pts[ind * 3 + 2] = axis(P.nz, q->Lz, dz, k);
        }
    double *dp = nullptr, *da = nullptr;
    double abc[3] = {dx, dy, dz};
    MT_HIP(hipMalloc(&dp, pts.size() * sizeof(double)));
    MT_HIP(hipMalloc(&da, sizeof(abc)));
    MT_HIP(hipMemcpyAsync(dp, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, P.stream));
    MT_HIP(hipMemcpyAsync(da, abc, sizeof(abc), hipMemcpyHostToDevice, P.stream));
    int rc = launch_assemble(dp, da, 0, nullptr, dp, n, n, P.K32, P.K64, nullptr, P.stream);
    if (rc) return rc;
    MT_HIP(hipStreamSynchronize(P.stream));
    (void)hipFree(dp);
    (void)hipFree(da);
  }
  std::vector<double> jf(n), kf(n), an;
  std::vector<float> ms(n);
  bool any_k = false;
  for (int i = 0; i < n; ++i) {
    double Ms = q->Ms[i], K0 = q->K0 ? q->K0[i] : 0.0;
    jf[i] = q->A0 / (kMu0 * Ms);
    kf[i] = K0 / (kMu0 * Ms);
    ms[i] = (float)Ms;
    any_k |= K0 != 0.0;
  }
  MT_HIP(hipMalloc(&P.jfact, n * sizeof(double)));
  MT_HIP(hipMalloc(&P.kfact, n * sizeof(double)));
  MT_HIP(hipMalloc(&P.ms32, n * sizeof(float)));
  MT_HIP(hipMalloc(&P.ms64, n * sizeof(double)));
  MT_HIP(hipMemcpyAsync(P.ms64, q->Ms, n * sizeof(double), hipMemcpyHostToDevice, P.stream));
  MT_HIP(hipMemcpyAsync(P.jfact, jf.data(), n * sizeof(double), hipMemcpyHostToDevice, P.stream));
  MT_HIP(hipMemcpyAsync(P.kfact, kf.data(), n * sizeof(double), hipMemcpyHostToDevice, P.stream));
  MT_HIP(hipMemcpyAsync(P.ms32, ms.data(), n * sizeof(float), hipMemcpyHostToDevice, P.stream));
  if (any_k && q->u_ea) {
    an.resize((size_t)n * 6);
    for (int i = 0; i < n; ++i) {
      double ux = q->u_ea[i * 3], uy = q->u_ea[i * 3 + 1], uz = q->u_ea[i * 3 + 2];
      double* A = &an[(size_t)i * 6];
      A[0] = ux * ux; A[1] = ux * uy; A[2] = ux * uz; A[3] = uy * uy; A[4] = uy * uz; A[5] = uz * uz;
    }
    MT_HIP(hipMalloc(&P.aniso, an.size() * sizeof(double)));
    MT_HIP(hipMemcpyAsync(P.aniso, an.data(), an.size() * sizeof(double), hipMemcpyHostToDevice, P.stream));
  }
  P.grid.n = n; P.grid.nx = P.nx; P.grid.ny = P.ny; P.grid.nz = P.nz;
  P.grid.idx2 = 1.0 / (dx * dx); P.grid.idy2 = 1.0 / (dy * dy); P.grid.idz2 = 1.0 / (dz * dz);
  P.grid.jfact = P.jfact; P.grid.kfact = P.kfact; P.grid.ms32 = P.ms32; P.grid.ms64 = P.ms64;
  P.grid.aniso = P.aniso;
  P.nck = (int)((nsteps + P.ck - 1) / P.ck);
  MT_HIP(hipMalloc(&P.packd, (size_t)P.B * N3 * P.es));
  MT_HIP(hipMalloc(&P.hbuf, (size_t)P.B * N3 * P.es));
  MT_HIP(hipMalloc(&P.ckpt, (size_t)P.nck * P.B * N3 * sizeof(double)));
  MT_HIP(hipMalloc(&P.seg, (size_t)(P.ck + 1) * P.B * 5 * N3 * sizeof(double)));
  MT_HIP(hipMalloc(&P.segS, (size_t)P.ck * 4 * P.B * N3 * P.es));
  MT_HIP(hipMalloc(&P.work, (size_t)13 * P.B * N3 * sizeof(double)));
  MT_HIP(hipMalloc(&P.grad_ms, (size_t)P.B * n * sizeof(double)));
  MT_HIP(hipMalloc(&P.hext_part, (size_t)P.B * P.nblocks() * 3 * sizeof(double)));
  P.desc_size = Rk4::kSlotBytes * Rk4::kSlots;
  MT_HIP(hipMalloc(&P.desc_d, P.desc_size));
  MT_HIP(hipHostMalloc(&P.desc_h, P.desc_size, hipHostMallocDefault));
  MT_HIP(hipMalloc(&P.dcache_d, Rk4::kCacheBytes));
  MT_CHECK(P.B * sizeof(VecDesc) <= Rk4::kSlotBytes, "too many states for one descriptor block");
  MT_HIP(hipStreamSynchronize(P.stream));
  *out = reinterpret_cast<mt_rk4*>(R.release());
  return kOk;
}

int mt_rk4_forward(mt_rk4* h, const double* m0, double* mT) {
  if (!h || !m0) {
    mt::set_error("null argument");
    return mt::kErrArg;
  }
  return reinterpret_cast<Rk4*>(h)->forward(m0, mT);
}

int mt_rk4_backward(mt_rk4* h, const double* dL_dmT, double* grad_m0, double* grad_hext, double* grad_ms) {
  if (!h || !dL_dmT) {
    mt::set_error("null argument");
    return mt::kErrArg;
  }
  return reinterpret_cast<Rk4*>(h)->backward(dL_dmT, grad_m0, grad_hext, grad_ms);
}

void mt_rk4_destroy(mt_rk4* h) { delete reinterpret_cast<Rk4*>(h); }

}  // extern "C"
