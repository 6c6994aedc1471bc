// Exact synthesis at arbitrary points (few points; no nuFFT approximation).
//
// The Legendre tables are built on the points' colatitudes (each point is a "ring" of one
// sample), the usual synthesis gives G_m(theta_p) for every m, and one workgroup per point sums
// f(theta_p, phi_p) = sum_m G_m(theta_p) e^{i m phi_p}, optionally rotated by e^{i s gamma_p}.
// Serves gclm2lenpixs (cunusht/deflection/CPU_nufft_transformer.py:531-558).
#pragma once

#include "common.h"
#include "legendre.h"

namespace nusht {

class PointPlan {
 public:
  PointPlan(int lmax, int spin, const std::vector<double>& theta, hipStream_t st);
  int lmax() const { return lmax_; }
  int spin() const { return spin_; }
  long npts() const { return npts_; }
  // alm (ncomp x nalm) -> out0 (spin 0 value, or Q), out1 (U); phi, gamma (may be null) on the device
  void synthesis(const double2* alm, int ncomp, const double* phi, const double* gamma, double* out0, double* out1,
                 hipStream_t st);

 private:
  int lmax_, spin_;
  long npts_;
  int nrows_ = 0;
  LegendreTables tab_;
  DevBuf G_, coef_;
};

// Exact synthesis at points without any device allocation (no plan): the recursion seeds and coefficients are
// computed on the device per call and the ring coefficients G_m(theta_p) go into a file-scope device array of
// kNoAllocG complex values, one per GPU, shared by every call on that GPU (calls that overlap on one GPU race).
// theta, phi, gamma (may be null), alm, out0 / out1 are device arrays; requires (rows = lmax + 1 for spin 0,
// 2 lmax + 1 otherwise) x npts <= kNoAllocG.
constexpr long kNoAllocG = 1L << 21;  // 32 MiB of double2
void point_synthesis_noalloc(int lmax, int spin, long npts, const double* theta, const double2* alm, int ncomp,
                             const double* phi, const double* gamma, double* out0, double* out1, hipStream_t st);

}  // namespace nusht
