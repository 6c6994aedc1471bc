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

struct RingArgs {
  int lmax, spin, nphi, nrows, gstride, method, mlen, mode;
  const double* theta;
  const double* phi0;
  const long* ofs;
  double2* G;
  FftDesc<double> fd;
  const double2* chirp;
  const double2* kern;
  const int* perm_in;
  const int* perm_out;
  double* o0;
  double* o1;
  double* o2;
  const double* i0;
  const double* i1;
  const double* wt;  // per-ring weights applied to the adjoint's input (null: none)
};

// In-place forward DFT of buf[0..n) (n = nphi) in LDS.
template <int METHOD>
__device__ __forceinline__ void ring_fwd_dft(double2* buf, const RingArgs& a) {
  const int n = a.nphi;
  constexpr int NT = kRingNT;
  if constexpr (METHOD == (int)RingMethod::Direct) {
    fft_lds<double, NT, kRingNMAX, true>(buf, a.fd);
    return;
  }
  if constexpr (METHOD == (int)RingMethod::Bluestein) {
    const int M = a.mlen;
    for (int j = threadIdx.x; j < M; j += NT) buf[j] = j < n ? cmul(buf[j], a.chirp[j]) : double2{0.0, 0.0};
    __syncthreads();
    fft_lds<double, NT, kRingNMAX, true>(buf, a.fd);
    for (int j = threadIdx.x; j < M; j += NT) buf[j] = cconj(cmul(buf[j], a.kern[j]));
    __syncthreads();
    fft_lds<double, NT, kRingNMAX, true>(buf, a.fd);
    const double inv = 1.0 / M;
    for (int k = threadIdx.x; k < n; k += NT) buf[k] = cscale(cmul(a.chirp[k], cconj(buf[k])), inv);
    __syncthreads();
    return;
  }
  if constexpr (METHOD != (int)RingMethod::Rader) return;
  // Rader (n prime): X[g^-p] = x0 + (a (*) b)_p with a_q = x[g^q], b_q = w^{g^-q}
  constexpr int E = (kRingNMAX + NT - 1) / NT;
  const int n1 = n - 1;
  const double2 x0 = buf[0];
  double2 v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int q = threadIdx.x + i * NT;
    if (q < n1) v[i] = buf[a.perm_in[q]];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int q = threadIdx.x + i * NT;
    if (q < n1) buf[q] = v[i];
  }
  __syncthreads();
  fft_lds<double, NT, kRingNMAX, true>(buf, a.fd);
  const double2 a0 = buf[0];
  __syncthreads();
  for (int q = threadIdx.x; q < n1; q += NT) buf[q] = cconj(cmul(buf[q], a.kern[q]));
  __syncthreads();
  fft_lds<double, NT, kRingNMAX, true>(buf, a.fd);
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
    if (p < n1) buf[a.perm_out[p]] = v[i];
  }
  if (threadIdx.x == 0) buf[0] = cadd(x0, a0);
  __syncthreads();
}

template <int METHOD>
__global__ __launch_bounds__(kRingNT) void k_ring(RingArgs a) {
  extern __shared__ unsigned char smem_raw[];
  double2* buf = reinterpret_cast<double2*>(smem_raw);
  // neighbouring rings run on the same XCD at the same time, so the m-major G columns they read
  // share that XCD's L2 lines
  const int ring = xcd_remap(blockIdx.x, gridDim.x);
  const int n = a.nphi;
  const int L = a.lmax;
  const double ph0 = a.phi0[ring];
  const long ofs = a.ofs[ring];
  double2* g = a.G + ring;  // G[row][ring], row stride gstride
  if (a.mode < 3) {
    for (int k = threadIdx.x; k < n; k += kRingNT) buf[k] = double2{0.0, 0.0};
    __syncthreads();
    // conjugated spectrum for the inverse DFT: X[m mod n] = conj(G_m e^{i m phi0})
    const int mlo = a.spin ? -L : 0;
    for (int m = mlo + (int)threadIdx.x; m <= L; m += kRingNT) {
      double2 v = g[(long)(a.spin ? m + L : m) * a.gstride];
      if (ph0 != 0.0) {
        double s, c;
        sincos(m * ph0, &s, &c);
        v = cmul(v, double2{c, s});
      }
      int k = m >= 0 ? m : m + n;
      buf[k] = cconj(v);
      if (!a.spin && m > 0) buf[n - m] = v;
    }
    __syncthreads();
    ring_fwd_dft<METHOD>(buf, a);
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
    ring_fwd_dft<METHOD>(buf, a);
    const int mlo = a.spin ? -L : 0;
    for (int m = mlo + (int)threadIdx.x; m <= L; m += kRingNT) {
      int k = m >= 0 ? m : m + n;
      double2 v = buf[k];
      if (ph0 != 0.0) {
        double s, c;
        sincos(-m * ph0, &s, &c);
        v = cmul(v, double2{c, s});
      }
      g[(long)(a.spin ? m + L : m) * a.gstride] = v;
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

RingPlan::RingPlan(int lmax, int spin, const std::vector<double>& theta, const std::vector<double>& phi0,
                   const std::vector<long>& nph, const std::vector<long>& ofs, const std::vector<double>& weight,
                   hipStream_t st)
    : lmax_(lmax), spin_(spin), nring_((int)theta.size()) {
  NUSHT_REQUIRE(nring_ >= 1, kErrArg, "ringplan: no rings");
  NUSHT_REQUIRE(phi0.size() == theta.size() && nph.size() == theta.size() && ofs.size() == theta.size(), kErrArg,
                "ringplan: ring arrays differ in length");
  NUSHT_REQUIRE(weight.empty() || weight.size() == theta.size(), kErrArg, "ringplan: one weight per ring");
  nphi_ = (int)nph[0];
  for (long v : nph)
    NUSHT_REQUIRE(v == nphi_, kErrUnsupported, "ringplan: rings must share one nphi");
  NUSHT_REQUIRE(nphi_ >= 2 * lmax + 1, kErrUnsupported, "ringplan: nphi must be >= 2 lmax + 1");
  npix_ = 0;
  for (int i = 0; i < nring_; ++i) npix_ = std::max(npix_, ofs[i] + nph[i]);
  const int maxlen = kRingNMAX;
  const int n = nphi_;
  if (fft_supported(n) && n <= maxlen) {
    method_ = RingMethod::Direct;
    mlen_ = n;
  } else if (is_prime(n) && fft_supported(n - 1) && n <= maxlen) {
    method_ = RingMethod::Rader;
    mlen_ = n - 1;
  } else {
    int M = fft_good_size(2 * n - 1);
    NUSHT_REQUIRE(M <= maxlen, kErrUnsupported, "ringplan: nphi " + std::to_string(n) + " not supported");
    method_ = RingMethod::Bluestein;
    mlen_ = M;
  }
  fft_.init(mlen_);
  const size_t lds = sizeof(double2) * (size_t)std::max(n, mlen_);
  NUSHT_HIP(hipFuncSetAttribute((const void*)k_fft_table, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  NUSHT_HIP(hipFuncSetAttribute((const void*)k_ring<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  NUSHT_HIP(hipFuncSetAttribute((const void*)k_ring<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  NUSHT_HIP(hipFuncSetAttribute((const void*)k_ring<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (method_ == RingMethod::Bluestein) {
    std::vector<double2> c(n), b(mlen_, double2{0.0, 0.0});
    for (long j = 0; j < n; ++j) {
      long j2 = (j * j) % (2L * n);
      double ang = -M_PI * (double)j2 / n;
      c[j] = double2{std::cos(ang), std::sin(ang)};
    }
    for (long j = 0; j < n; ++j) {
      b[j] = double2{c[j].x, -c[j].y};
      if (j > 0) b[mlen_ - j] = b[j];
    }
    upload(chirp_, c);
    DevBuf bt;
    upload(bt, b);
    kern_.alloc(sizeof(double2) * mlen_);
    hipLaunchKernelGGL(k_fft_table, dim3(1), dim3(kRingNT), lds, st, fft_.desc(), bt.as<double2>(), kern_.as<double2>());
    NUSHT_LAUNCH_CHECK();
    NUSHT_HIP(hipStreamSynchronize(st));
  } else if (method_ == RingMethod::Rader) {
    long g = primitive_root(n);
    long ginv = powmod(g, n - 2, n);
    std::vector<int> pin(n - 1), pout(n - 1);
    std::vector<double2> b(n - 1);
    long gi = 1, go = 1;
    for (int q = 0; q < n - 1; ++q) {
      pin[q] = (int)gi;
      pout[q] = (int)go;
      double ang = -2.0 * M_PI * (double)go / n;
      b[q] = double2{std::cos(ang), std::sin(ang)};
      gi = gi * g % n;
      go = go * ginv % n;
    }
    upload(perm_in_, pin);
    upload(perm_out_, pout);
    DevBuf bt;
    upload(bt, b);
    kern_.alloc(sizeof(double2) * (n - 1));
    hipLaunchKernelGGL(k_fft_table, dim3(1), dim3(kRingNT), lds, st, fft_.desc(), bt.as<double2>(), kern_.as<double2>());
    NUSHT_LAUNCH_CHECK();
    NUSHT_HIP(hipStreamSynchronize(st));
  }
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
  return tab_.bytes() + theta_.bytes() + phi0_.bytes() + ofs_.bytes() + weight_.bytes() + chirp_.bytes() + kern_.bytes() +
         perm_in_.bytes() + perm_out_.bytes() + sqll_.bytes() + G_.bytes() + coef_.bytes();
}

void RingPlan::ring_pass(int mode, double* o0, double* o1, double* o2, hipStream_t st, const double* wt) {
  RingArgs a;
  a.lmax = lmax_; a.spin = spin_; a.nphi = nphi_; a.nrows = nrows_; a.gstride = gstride_; a.method = (int)method_;
  a.mlen = mlen_;
  a.mode = mode;
  a.theta = theta_.as<double>(); a.phi0 = phi0_.as<double>(); a.ofs = ofs_.as<long>();
  a.G = G_.as<double2>();
  a.fd = fft_.desc();
  a.chirp = chirp_.as<double2>(); a.kern = kern_.as<double2>();
  a.perm_in = perm_in_.as<int>(); a.perm_out = perm_out_.as<int>();
  a.o0 = o0; a.o1 = o1; a.o2 = o2;
  a.i0 = o0; a.i1 = o1;
  a.wt = wt;
  const size_t lds = sizeof(double2) * (size_t)std::max(nphi_, mlen_);
  static const char* names[4] = {"ring_dft_synth_s0", "ring_dft_synth_spin", "ring_dft_pointing", "ring_dft_adjoint"};
  StageScope sc(names[mode], st);
  switch (method_) {
    case RingMethod::Direct: hipLaunchKernelGGL(k_ring<0>, dim3(nring_), dim3(kRingNT), lds, st, a); break;
    case RingMethod::Rader: hipLaunchKernelGGL(k_ring<1>, dim3(nring_), dim3(kRingNT), lds, st, a); break;
    case RingMethod::Bluestein: hipLaunchKernelGGL(k_ring<2>, dim3(nring_), dim3(kRingNT), lds, st, a); break;
  }
  NUSHT_LAUNCH_CHECK();
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

void RingPlan::dlm2pointing(const double2* dlm_scaled, double* ptheta, double* pphi, double* gamma, hipStream_t st) {
  NUSHT_REQUIRE(spin_ == 1, kErrState, "dlm2pointing needs a spin-1 ring plan");
  legendre_synth(tab_, 1, 1, dlm_scaled, 0, sqll_.as<double>(), coef_.as<double2>(), G_.as<double2>(), gstride_, 1, 0,
                 st);
  ring_pass(2, ptheta, pphi, gamma, st);
}

}  // namespace nusht
