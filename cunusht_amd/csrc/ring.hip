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

constexpr int kRingNT = kDftNT;

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
      // The thread's rows (at most kDftE: L - mlo < n <= kDftNMAX) are loaded unconditionally, from rows
      // clamped to L, so that every load is in flight before the first LDS write.
      double2 gv[kDftE];
#pragma unroll
      for (int i = 0; i < kDftE; ++i) {
        const int m = min(mlo + (int)threadIdx.x + i * kRingNT, L);
        gv[i] = g[(long)(a.spin ? m + L : m) * a.gstride];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kDftE; ++i) {
        const int m = mlo + (int)threadIdx.x + i * kRingNT;
        if (m <= L) {
          double2 v = gv[i];
          if (ph0 != 0.0) {
            double s, c;
            sincos(m * ph0, &s, &c);
            v = cmul(v, double2{c, s});
          }
          const int k = m >= 0 ? m : m + n;
          buf[k] = cconj(v);
          if (!a.spin && m > 0) buf[n - m] = v;
        }
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
    dft_lds<METHOD, double>(buf, Ld);
    const double th = a.theta[ring];
    const double sth = sin(th), cth = cos(th);  // constant along the ring
    for (int k = threadIdx.x; k < n; k += kRingNT) {
      double2 f = cconj(buf[k]);
      if (a.mode == 0) {
        a.o0[ofs + k] = f.x;
      } else if (a.mode == 1) {
        a.o0[ofs + k] = f.x;
        a.o1[ofs + k] = f.y;
      } else {
        double tp, pp;
        deflect_point_sc(sth, cth, ph0 + k * (2.0 * M_PI / n), f.x, f.y, tp, pp, a.o2 ? &a.o2[ofs + k] : nullptr);
        a.o0[ofs + k] = tp;
        a.o1[ofs + k] = pp;
      }
    }
  } else {
    const double w = a.wt ? a.wt[ring] : 1.0;
    for (int k = threadIdx.x; k < n; k += kRingNT)
      buf[k] = double2{w * a.i0[ofs + k], (a.spin && a.i1) ? w * a.i1[ofs + k] : 0.0};
    __syncthreads();
    dft_lds<METHOD, double>(buf, Ld);
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
  for (int i = 0; i < nring_; ++i) {
    NUSHT_REQUIRE(theta[i] >= 0.0 && theta[i] <= M_PI, kErrArg, "ringplan: theta outside [0, pi]");
    NUSHT_REQUIRE(nph[i] >= 1 && ofs[i] >= 0, kErrArg, "ringplan: bad nph or ofs");
  }
  npix_ = 0;
  for (int i = 0; i < nring_; ++i) npix_ = std::max(npix_, ofs[i] + nph[i]);
  const size_t table_lds = sizeof(double2) * kDftNMAX;
  // one DFT description per distinct ring length; rings grouped by DFT method
  std::vector<int> lidx(nring_);
  std::vector<std::vector<int>> by_method(kDftMethods);
  int longest = 0;
  for (int i = 0; i < nring_; ++i) {
    lidx[i] = dft_.add((int)nph[i], st);
    const RingLen& d = dft_.at(lidx[i]);
    by_method[d.method].push_back(i);
    lds_[d.method] = std::max(lds_[d.method], sizeof(double2) * (size_t)d.lds_values());
    if (nph[i] > nph[longest]) longest = i;
  }
  method_ = (RingMethod)dft_.at(lidx[longest]).method;
  std::vector<RingLen> lv;
  for (int i = 0; i < dft_.size(); ++i) lv.push_back(dft_.at(i));
  upload(lens_, lv);
  upload(lenidx_, lidx);
  for (int m = 0; m < kDftMethods; ++m) {
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
  if (spin) legendre_autotune(tab_, 1, coef_.as<double2>(), G_.as<double2>(), gstride_, 1, st);  // E only
}

size_t RingPlan::bytes() const {
  size_t b = tab_.bytes() + theta_.bytes() + phi0_.bytes() + ofs_.bytes() + weight_.bytes() + sqll_.bytes() +
             G_.bytes() + coef_.bytes() + lens_.bytes() + lenidx_.bytes();
  b += dft_.bytes();
  for (int m = 0; m < kDftMethods; ++m) b += rings_[m].bytes();
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
  for (int m = 0; m < kDftMethods; ++m) {
    if (!nrings_m_[m]) continue;
    a.rings = rings_[m].as<int>();
    const bool uni = uniform_[m] >= 0;
    if (uni) a.one = dft_.at(uniform_[m]);
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
