// SHT on iso-latitude rings (see ring.h).
//
// Replaces the reference's GL-grid SHTns calls (synthesis_cupy / adjoint_synthesis_cupy /
// synthesis_der1_cupy, cunusht/sht/GPU_sht_transformer.py:56-96) and, for spin 1, the pointing
// kernel launched after synth_grad (cunusht/deflection/GPU_nufft_transformer.py:83-101).
#include "ring.h"

#include "stages.h"
#include "timing.h"

#include <algorithm>
#include <cmath>

namespace nusht {
namespace {

// 768 threads: a radix-13 stage over 8190 points (Rader for nphi = 8191) needs one butterfly per thread,
// and 768 rather than 640 threads halves its register spill.
constexpr int kRingNT = 768;
constexpr int kRingNMAX = 8320;
constexpr int kRingE = (kRingNMAX + kRingNT - 1) / kRingNT;  // ring values per thread held in registers

struct RingArgs {
  int lmax, spin, gstride, mode;
  RingLen one;            // the ring length of a launch whose rings all share one length
  const int* rings;       // rings of this launch (one DFT method)
  const int* lenidx;      // per ring: index into lens
  const RingLen* lens;
  const double* theta;
  const double* phi0;
  const long* ofs;
  double2* G;
  double* o0;
  double* o1;
  double* o2;
  const double* i0;
  const double* i1;
  const double* wt;  // per-ring weights applied to the adjoint's input (null: none)
};

// Forward DFT of buf[0..n) by Bluestein's algorithm on the padded length L.mlen:
// X_k = c_k sum_j (x_j c_j) conj(c)_{k-j},  c_j = e^{-i pi j^2 / n}.
__device__ __forceinline__ void bluestein_lds(double2* buf, int n, const RingLen& L) {
  constexpr int NT = kRingNT;
  const int M = L.mlen;
  for (int j = threadIdx.x; j < M; j += NT) buf[j] = j < n ? cmul(buf[j], L.chirp[j]) : double2{0.0, 0.0};
  __syncthreads();
  fft_lds<double, NT, kRingNMAX, true>(buf, L.fd);
  for (int j = threadIdx.x; j < M; j += NT) buf[j] = cconj(cmul(buf[j], L.kern[j]));
  __syncthreads();
  fft_lds<double, NT, kRingNMAX, true>(buf, L.fd);
  const double inv = 1.0 / M;
  for (int k = threadIdx.x; k < n; k += NT) buf[k] = cscale(cmul(L.chirp[k], cconj(buf[k])), inv);
  __syncthreads();
}

// In-place forward DFT of buf[0..n) (n = L.n) in LDS.
template <int METHOD>
__device__ __forceinline__ void ring_fwd_dft(double2* buf, const RingLen& L) {
  const int n = L.n;
  constexpr int NT = kRingNT;
  constexpr int E = kRingE;
  if constexpr (METHOD == (int)RingMethod::Direct) {
    fft_lds<double, NT, kRingNMAX, true>(buf, L.fd);
    return;
  }
  if constexpr (METHOD == (int)RingMethod::Bluestein) {
    bluestein_lds(buf, n, L);
    return;
  }
  if constexpr (METHOD == (int)RingMethod::Split) {
    // n = r b, j = j1 + r j2:  X_k = sum_{j1} w_n^{j1 k} Y_{j1}[k mod b],  Y_{j1} = DFT_b(x_{j1 + r j2})
    const int r = L.r, b = L.b;
    double2 x[E], acc[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int q = threadIdx.x + i * NT;
      x[i] = q < n ? buf[q] : double2{0.0, 0.0};
      acc[i] = double2{0.0, 0.0};
    }
    __syncthreads();
    for (int j1 = 0; j1 < r; ++j1) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int q = threadIdx.x + i * NT;
        if (q < n && q % r == j1) buf[q / r] = x[i];
      }
      __syncthreads();
      if (L.sub_bluestein) {
        bluestein_lds(buf, b, L);
      } else {
        fft_lds<double, NT, kRingNMAX, true>(buf, L.fd);
      }
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < n) {
          const long e = (long)j1 * k % n;
          double sn, cs;
          sincospi(-2.0 * (double)e / (double)n, &sn, &cs);
          acc[i] = cadd(acc[i], cmul(double2{cs, sn}, buf[k % b]));
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int q = threadIdx.x + i * NT;
      if (q < n) buf[q] = acc[i];
    }
    __syncthreads();
    return;
  }
  if constexpr (METHOD != (int)RingMethod::Rader) return;
  // Rader (n prime): X[g^-p] = x0 + (a (*) b)_p with a_q = x[g^q], b_q = w^{g^-q}
  const int n1 = n - 1;
  const double2 x0 = buf[0];
  double2 v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int q = threadIdx.x + i * NT;
    if (q < n1) v[i] = buf[L.perm_in[q]];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int q = threadIdx.x + i * NT;
    if (q < n1) buf[q] = v[i];
  }
  __syncthreads();
  fft_lds<double, NT, kRingNMAX, true>(buf, L.fd);
  const double2 a0 = buf[0];
  __syncthreads();
  for (int q = threadIdx.x; q < n1; q += NT) buf[q] = cconj(cmul(buf[q], L.kern[q]));
  __syncthreads();
  fft_lds<double, NT, kRingNMAX, true>(buf, L.fd);
  const double inv = 1.0 / n1;
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int p = threadIdx.x + i * NT;
    if (p < n1) v[i] = cadd(x0, cscale(cconj(buf[p]), inv));
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int p = threadIdx.x + i * NT;
    if (p < n1) buf[L.perm_out[p]] = v[i];
  }
  if (threadIdx.x == 0) buf[0] = cadd(x0, a0);
  __syncthreads();
}

// G_m e^{i m phi0} for row m of this ring
__device__ __forceinline__ double2 ring_coef(const double2* g, int m, int L, int spin, long gstride, double ph0) {
  double2 v = g[(long)(spin ? m + L : m) * gstride];
  if (ph0 != 0.0) {
    double s, c;
    sincos(m * ph0, &s, &c);
    v = cmul(v, double2{c, s});
  }
  return v;
}

// UNIFORM: every ring of the launch has the length a.one (read from the kernel arguments); otherwise each
// ring reads its description from a.lens.
template <int METHOD, bool UNIFORM>
__global__ __launch_bounds__(kRingNT) void k_ring(RingArgs a) {
  extern __shared__ unsigned char smem_raw[];
  double2* buf = reinterpret_cast<double2*>(smem_raw);
  // neighbouring rings run on the same XCD at the same time, so the m-major G columns they read
  // share that XCD's L2 lines
  const int ring = a.rings[xcd_remap(blockIdx.x, gridDim.x)];
  const RingLen& Ld = UNIFORM ? a.one : a.lens[a.lenidx[ring]];
  const int n = Ld.n;
  const int L = a.lmax;
  const double ph0 = a.phi0[ring];
  const long ofs = a.ofs[ring];
  const double2* g = a.G + ring;  // G[row][ring], row stride gstride
  if (a.mode < 3) {
    // conjugated spectrum for the inverse DFT: X[m mod n] = sum conj(G_m e^{i m phi0}) over the m that alias there
    const int mlo = a.spin ? -L : 0;
    if (n >= 2 * L + 1) {
      for (int k = threadIdx.x; k < n; k += kRingNT) buf[k] = double2{0.0, 0.0};
      __syncthreads();
      for (int m = mlo + (int)threadIdx.x; m <= L; m += kRingNT) {
        const double2 v = ring_coef(g, m, L, a.spin, a.gstride, ph0);
        int k = m >= 0 ? m : m + n;
        buf[k] = cconj(v);
        if (!a.spin && m > 0) buf[n - m] = v;
      }
    } else {
      for (int k = threadIdx.x; k < n; k += kRingNT) {
        double2 acc{0.0, 0.0};
        for (int m = mlo + ((k - mlo) % n + n) % n; m <= L; m += n)
          acc = cadd(acc, cconj(ring_coef(g, m, L, a.spin, a.gstride, ph0)));
        if (!a.spin)  // the m < 0 half of a real field: conj(G_{-m}) = G_m at k = -m mod n
          for (int m = (n - k) % n == 0 ? n : (n - k) % n; m <= L; m += n)
            acc = cadd(acc, ring_coef(g, m, L, 0, a.gstride, ph0));
        buf[k] = acc;
      }
    }
    __syncthreads();
    ring_fwd_dft<METHOD>(buf, Ld);
    const double th = a.theta[ring];
    for (int k = threadIdx.x; k < n; k += kRingNT) {
      double2 f = cconj(buf[k]);
      if (a.mode == 0) {
        a.o0[ofs + k] = f.x;
      } else if (a.mode == 1) {
        a.o0[ofs + k] = f.x;
        a.o1[ofs + k] = f.y;
      } else {
        double tp, pp;
        deflect_point(th, ph0 + k * (2.0 * M_PI / n), f.x, f.y, tp, pp, a.o2 ? &a.o2[ofs + k] : nullptr);
        a.o0[ofs + k] = tp;
        a.o1[ofs + k] = pp;
      }
    }
  } else {
    const double w = a.wt ? a.wt[ring] : 1.0;
    for (int k = threadIdx.x; k < n; k += kRingNT)
      buf[k] = double2{w * a.i0[ofs + k], (a.spin && a.i1) ? w * a.i1[ofs + k] : 0.0};
    __syncthreads();
    ring_fwd_dft<METHOD>(buf, Ld);
    const int mlo = a.spin ? -L : 0;
    for (int m = mlo + (int)threadIdx.x; m <= L; m += kRingNT) {
      const int k = (m % n + n) % n;
      double2 v = buf[k];
      if (ph0 != 0.0) {
        double s, c;
        sincos(-m * ph0, &s, &c);
        v = cmul(v, double2{c, s});
      }
      a.G[ring + (long)(a.spin ? m + L : m) * a.gstride] = v;
    }
  }
}

// One-off table transform: out = FFT(in) of length fd.n.
__global__ __launch_bounds__(kRingNT) void k_fft_table(FftDesc<double> fd, const double2* in, double2* out) {
  extern __shared__ unsigned char smem_raw[];
  double2* buf = reinterpret_cast<double2*>(smem_raw);
  for (int k = threadIdx.x; k < fd.n; k += kRingNT) buf[k] = in[k];
  __syncthreads();
  fft_lds<double, kRingNT, kRingNMAX, true>(buf, fd);
  for (int k = threadIdx.x; k < fd.n; k += kRingNT) out[k] = buf[k];
}

bool is_prime(long n) {
  if (n < 2) return false;
  for (long d = 2; d * d <= n; ++d)
    if (n % d == 0) return false;
  return true;
}

long powmod(long b, long e, long m) {
  long r = 1;
  b %= m;
  while (e > 0) {
    if (e & 1) r = r * b % m;
    b = b * b % m;
    e >>= 1;
  }
  return r;
}

long primitive_root(long n) {
  std::vector<long> fac;
  long m = n - 1;
  for (long d = 2; d * d <= m; ++d)
    if (m % d == 0) { fac.push_back(d); while (m % d == 0) m /= d; }
  if (m > 1) fac.push_back(m);
  for (long g = 2; g < n; ++g) {
    bool ok = true;
    for (long q : fac)
      if (powmod(g, (n - 1) / q, n) == 1) { ok = false; break; }
    if (ok) return g;
  }
  return -1;
}

__global__ void k_sqrt_ll(int lmax, double* out) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l <= lmax) out[l] = sqrt((double)l * (l + 1));
}

}  // namespace

const FftLength<double>& RingPlan::fft_of(int len) {
  auto it = ffts_.find(len);
  if (it == ffts_.end()) it = ffts_.emplace(len, std::make_unique<FftLength<double>>(len)).first;
  return *it->second;
}

// Bluestein tables for a transform of length n padded to M: chirp c_j = e^{-i pi j^2 / n} and the FFT of
// the conjugate chirp wrapped onto the cyclic length M.
static void bluestein_tables(int n, int M, const FftDesc<double>& fd, DevBuf& chirp, DevBuf& kern, hipStream_t st) {
  std::vector<double2> c(n), b(M, double2{0.0, 0.0});
  for (long j = 0; j < n; ++j) {
    long j2 = (j * j) % (2L * n);
    double ang = -M_PI * (double)j2 / n;
    c[j] = double2{std::cos(ang), std::sin(ang)};
  }
  for (long j = 0; j < n; ++j) {
    b[j] = double2{c[j].x, -c[j].y};
    if (j > 0) b[M - j] = b[j];
  }
  upload(chirp, c);
  DevBuf bt;
  upload(bt, b);
  kern.alloc(sizeof(double2) * M);
  hipLaunchKernelGGL(k_fft_table, dim3(1), dim3(kRingNT), sizeof(double2) * M, st, fd, bt.as<double2>(),
                     kern.as<double2>());
  NUSHT_LAUNCH_CHECK();
  NUSHT_HIP(hipStreamSynchronize(st));
}

int RingPlan::add_length(int n, hipStream_t st) {
  auto it = len_index_.find(n);
  if (it != len_index_.end()) return it->second;
  NUSHT_REQUIRE(n >= 1 && n <= kRingNMAX, kErrUnsupported,
                "ringplan: ring length " + std::to_string(n) + " outside 1.." + std::to_string(kRingNMAX));
  auto t = std::make_unique<LenTables>();
  RingLen& d = t->d;
  d.n = n;
  if (fft_supported(n)) {
    d.method = (int)RingMethod::Direct;
    d.mlen = n;
  } else if (is_prime(n) && fft_supported(n - 1)) {
    d.method = (int)RingMethod::Rader;
    d.mlen = n - 1;
  } else if (fft_good_size(2 * n - 1) <= kRingNMAX) {
    d.method = (int)RingMethod::Bluestein;
    d.mlen = fft_good_size(2 * n - 1);
  } else {
    // n = r b with the fewest sub-transforms whose length b runs in LDS
    d.method = (int)RingMethod::Split;
    for (int r = 2; r <= n && d.r == 0; ++r) {
      if (n % r) continue;
      const int b = n / r;
      if (fft_supported(b)) {
        d.r = r; d.b = b; d.mlen = b; d.sub_bluestein = 0;
      } else if (fft_good_size(2 * b - 1) <= kRingNMAX) {
        d.r = r; d.b = b; d.mlen = fft_good_size(2 * b - 1); d.sub_bluestein = 1;
      }
    }
    NUSHT_REQUIRE(d.r > 0 && d.r < n, kErrUnsupported,
                  "ringplan: ring length " + std::to_string(n) + " not supported (prime with a non-smooth n - 1)");
    NUSHT_REQUIRE(d.r <= 64, kErrUnsupported, "ringplan: ring length " + std::to_string(n) + " needs too many sub-transforms");
  }
  d.fd = fft_of(d.mlen).desc();
  if (d.method == (int)RingMethod::Bluestein) {
    bluestein_tables(n, d.mlen, d.fd, t->chirp, t->kern, st);
  } else if (d.method == (int)RingMethod::Split && d.sub_bluestein) {
    bluestein_tables(d.b, d.mlen, d.fd, t->chirp, t->kern, st);
  } else if (d.method == (int)RingMethod::Rader) {
    long gr = primitive_root(n);
    long ginv = powmod(gr, n - 2, n);
    std::vector<int> pin(n - 1), pout(n - 1);
    std::vector<double2> b(n - 1);
    long gi = 1, go = 1;
    for (int q = 0; q < n - 1; ++q) {
      pin[q] = (int)gi;
      pout[q] = (int)go;
      double ang = -2.0 * M_PI * (double)go / n;
      b[q] = double2{std::cos(ang), std::sin(ang)};
      gi = gi * gr % n;
      go = go * ginv % n;
    }
    upload(t->perm_in, pin);
    upload(t->perm_out, pout);
    DevBuf bt;
    upload(bt, b);
    t->kern.alloc(sizeof(double2) * (n - 1));
    hipLaunchKernelGGL(k_fft_table, dim3(1), dim3(kRingNT), sizeof(double2) * (n - 1), st, d.fd, bt.as<double2>(),
                       t->kern.as<double2>());
    NUSHT_LAUNCH_CHECK();
    NUSHT_HIP(hipStreamSynchronize(st));
  }
  d.chirp = t->chirp.as<double2>();
  d.kern = t->kern.as<double2>();
  d.perm_in = t->perm_in.as<int>();
  d.perm_out = t->perm_out.as<int>();
  const int idx = (int)lens_h_.size();
  lens_h_.push_back(std::move(t));
  len_index_[n] = idx;
  return idx;
}

RingPlan::RingPlan(int lmax, int spin, const std::vector<double>& theta, const std::vector<double>& phi0,
                   const std::vector<long>& nph, const std::vector<long>& ofs, const std::vector<double>& weight,
                   hipStream_t st)
    : lmax_(lmax), spin_(spin), nring_((int)theta.size()) {
  NUSHT_REQUIRE(nring_ >= 1, kErrArg, "ringplan: no rings");
  NUSHT_REQUIRE(phi0.size() == theta.size() && nph.size() == theta.size() && ofs.size() == theta.size(), kErrArg,
                "ringplan: ring arrays differ in length");
  NUSHT_REQUIRE(weight.empty() || weight.size() == theta.size(), kErrArg, "ringplan: one weight per ring");
  for (int i = 0; i < nring_; ++i) {
    NUSHT_REQUIRE(theta[i] >= 0.0 && theta[i] <= M_PI, kErrArg, "ringplan: theta outside [0, pi]");
    NUSHT_REQUIRE(nph[i] >= 1 && ofs[i] >= 0, kErrArg, "ringplan: bad nph or ofs");
  }
  npix_ = 0;
  for (int i = 0; i < nring_; ++i) npix_ = std::max(npix_, ofs[i] + nph[i]);
  const size_t table_lds = sizeof(double2) * kRingNMAX;
  NUSHT_HIP(hipFuncSetAttribute((const void*)k_fft_table, hipFuncAttributeMaxDynamicSharedMemorySize, (int)table_lds));
  // one DFT description per distinct ring length; rings grouped by DFT method
  std::vector<int> lidx(nring_);
  std::vector<std::vector<int>> by_method(kRingMethods);
  int longest = 0;
  for (int i = 0; i < nring_; ++i) {
    lidx[i] = add_length((int)nph[i], st);
    const RingLen& d = lens_h_[lidx[i]]->d;
    by_method[d.method].push_back(i);
    lds_[d.method] = std::max(lds_[d.method], sizeof(double2) * (size_t)std::max(d.n, d.mlen));
    if (nph[i] > nph[longest]) longest = i;
  }
  method_ = (RingMethod)lens_h_[lidx[longest]]->d.method;
  std::vector<RingLen> lv;
  for (auto& t : lens_h_) lv.push_back(t->d);
  upload(lens_, lv);
  upload(lenidx_, lidx);
  for (int m = 0; m < kRingMethods; ++m) {
    nrings_m_[m] = (int)by_method[m].size();
    if (nrings_m_[m]) upload(rings_[m], by_method[m]);
    uniform_[m] = -1;
    for (int i : by_method[m]) {
      if (uniform_[m] == -1) uniform_[m] = lidx[i];
      else if (uniform_[m] != lidx[i]) { uniform_[m] = -2; break; }
    }
  }
  for (const void* f : {(const void*)k_ring<0, true>, (const void*)k_ring<1, true>, (const void*)k_ring<2, true>,
                        (const void*)k_ring<3, true>, (const void*)k_ring<0, false>, (const void*)k_ring<1, false>,
                        (const void*)k_ring<2, false>, (const void*)k_ring<3, false>})
    NUSHT_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)table_lds));
  upload(theta_, theta);
  upload(phi0_, phi0);
  upload(ofs_, ofs);
  if (!weight.empty()) upload(weight_, weight);
  tab_.tag = "ring";
  tab_.build(lmax, spin, theta, st);
  nrows_ = spin ? 2 * lmax + 1 : lmax + 1;
  gstride_ = nring_;
  G_.alloc(sizeof(double2) * (size_t)gstride_ * nrows_);
  coef_.alloc(legendre_coef_bytes(tab_, 1));
  sqll_.alloc(sizeof(double) * (lmax + 1));
  hipLaunchKernelGGL(k_sqrt_ll, dim3(ceil_div(lmax + 1, 256)), dim3(256), 0, st, lmax, sqll_.as<double>());
  NUSHT_LAUNCH_CHECK();
  legendre_autotune(tab_, spin ? 2 : 1, coef_.as<double2>(), G_.as<double2>(), gstride_, 1, st);
}

size_t RingPlan::bytes() const {
  size_t b = tab_.bytes() + theta_.bytes() + phi0_.bytes() + ofs_.bytes() + weight_.bytes() + sqll_.bytes() +
             G_.bytes() + coef_.bytes() + lens_.bytes() + lenidx_.bytes();
  for (const auto& t : lens_h_) b += t->chirp.bytes() + t->kern.bytes() + t->perm_in.bytes() + t->perm_out.bytes();
  for (int m = 0; m < kRingMethods; ++m) b += rings_[m].bytes();
  return b;
}

void RingPlan::ring_pass(int mode, double* o0, double* o1, double* o2, hipStream_t st, const double* wt) {
  RingArgs a;
  a.lmax = lmax_; a.spin = spin_; a.gstride = gstride_; a.mode = mode;
  a.lenidx = lenidx_.as<int>();
  a.lens = lens_.as<RingLen>();
  a.theta = theta_.as<double>(); a.phi0 = phi0_.as<double>(); a.ofs = ofs_.as<long>();
  a.G = G_.as<double2>();
  a.o0 = o0; a.o1 = o1; a.o2 = o2;
  a.i0 = o0; a.i1 = o1;
  a.wt = wt;
  static const char* names[4] = {"ring_dft_synth_s0", "ring_dft_synth_spin", "ring_dft_pointing", "ring_dft_adjoint"};
  StageScope sc(names[mode], st);
  for (int m = 0; m < kRingMethods; ++m) {
    if (!nrings_m_[m]) continue;
    a.rings = rings_[m].as<int>();
    const bool uni = uniform_[m] >= 0;
    if (uni) a.one = lens_h_[uniform_[m]]->d;
    const dim3 grid(nrings_m_[m]), block(kRingNT);
    switch ((RingMethod)m) {
      case RingMethod::Direct:
        if (uni) hipLaunchKernelGGL((k_ring<0, true>), grid, block, lds_[m], st, a);
        else hipLaunchKernelGGL((k_ring<0, false>), grid, block, lds_[m], st, a);
        break;
      case RingMethod::Rader:
        if (uni) hipLaunchKernelGGL((k_ring<1, true>), grid, block, lds_[m], st, a);
        else hipLaunchKernelGGL((k_ring<1, false>), grid, block, lds_[m], st, a);
        break;
      case RingMethod::Bluestein:
        if (uni) hipLaunchKernelGGL((k_ring<2, true>), grid, block, lds_[m], st, a);
        else hipLaunchKernelGGL((k_ring<2, false>), grid, block, lds_[m], st, a);
        break;
      case RingMethod::Split:
        if (uni) hipLaunchKernelGGL((k_ring<3, true>), grid, block, lds_[m], st, a);
        else hipLaunchKernelGGL((k_ring<3, false>), grid, block, lds_[m], st, a);
        break;
    }
    NUSHT_LAUNCH_CHECK();
  }
}

void RingPlan::synthesis(const double2* alm, int ncomp, double* map0, double* map1, hipStream_t st) {
  legendre_synth(tab_, 1, ncomp, alm, 0, nullptr, coef_.as<double2>(), G_.as<double2>(), gstride_, 1, 0, st);
  ring_pass(spin_ ? 1 : 0, map0, map1, nullptr, st);
}

void RingPlan::adjoint(const double* map0, const double* map1, double2* alm, int ncomp, hipStream_t st) {
  ring_pass(3, const_cast<double*>(map0), const_cast<double*>(map1), nullptr, st);
  legendre_adjoint(tab_, ncomp, G_.as<double2>(), gstride_, 1, alm, st);
}

void RingPlan::analysis(const double* map0, const double* map1, double2* alm, int ncomp, hipStream_t st) {
  NUSHT_REQUIRE(has_weights(), kErrState, "ring plan created without ring weights: no analysis");
  ring_pass(3, const_cast<double*>(map0), const_cast<double*>(map1), nullptr, st, weight_.as<double>());
  legendre_adjoint(tab_, ncomp, G_.as<double2>(), gstride_, 1, alm, st);
}

void RingPlan::synthesis_grad(const double2* alm, double* dtheta, double* dphi, hipStream_t st) {
  NUSHT_REQUIRE(spin_ == 1, kErrState, "gradient synthesis needs a spin-1 ring plan");
  legendre_synth(tab_, 1, 1, alm, 0, sqll_.as<double>(), coef_.as<double2>(), G_.as<double2>(), gstride_, 1, 0, st);
  ring_pass(1, dtheta, dphi, nullptr, st);
}

void RingPlan::dlm2pointing(const double2* dlm_scaled, double* ptheta, double* pphi, double* gamma, hipStream_t st) {
  NUSHT_REQUIRE(spin_ == 1, kErrState, "dlm2pointing needs a spin-1 ring plan");
  legendre_synth(tab_, 1, 1, dlm_scaled, 0, sqll_.as<double>(), coef_.as<double2>(), G_.as<double2>(), gstride_, 1, 0,
                 st);
  ring_pass(2, ptheta, pphi, gamma, st);
}

}  // namespace nusht
