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

}  // namespace nusht
