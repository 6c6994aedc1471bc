// SHT on iso-latitude rings with uniform phi sampling per ring (the lensing/pointing grid), FP64.
//
// Synthesis = Legendre ring coefficients G_m(theta_r) followed by one length-nph DFT per ring;
// the adjoint runs the transposes.  Rings may differ in length (HEALPix, thinned Gauss-Legendre,
// any ragged geometry); a ring shorter than 2 lmax + 1 receives the aliased sum of its G_m
// (m folded modulo nph), as the exact sampling of the band-limited ring requires.  Every distinct
// length gets its own in-LDS DFT: a direct mixed-radix FFT when nph factors into 2, 3, 5, 7, 13,
// Rader's algorithm when nph is prime with a smooth nph - 1, Bluestein's chirp-z algorithm when
// its padded length fits in LDS, and otherwise a split nph = r * b into r sub-transforms of
// length b (direct or Bluestein) recombined with twiddles.  For spin 1 the deflected pointing
// (theta', phi', gamma) is computed straight from the ring values in LDS, so the deflection field
// itself is never written to HBM.
#pragma once

#include "common.h"
#include "dft_any.h"
#include "legendre.h"

namespace nusht {

using RingMethod = DftMethod;
using RingLen = DftLen<double>;

class RingPlan {
 public:
  // weight: per-ring quadrature weights (may be empty; analysis() then fails)
  RingPlan(int lmax, int spin, const std::vector<double>& theta, const std::vector<double>& phi0,
           const std::vector<long>& nph, const std::vector<long>& ofs, const std::vector<double>& weight,
           hipStream_t st);

  int lmax() const { return lmax_; }
  int spin() const { return spin_; }
  long npix() const { return npix_; }
  int nring() const { return nring_; }
  int nlengths() const { return dft_.size(); }
  RingMethod method() const { return method_; }  // method of the longest ring
  const LegendreTables& tables() const { return tab_; }

  // alm (ncomp x nalm complex) -> map0 (spin 0: real map; spin s: Q) and map1 (U)
  void synthesis(const double2* alm, int ncomp, double* map0, double* map1, hipStream_t st);
  // exact adjoint
  void adjoint(const double* map0, const double* map1, double2* alm, int ncomp, hipStream_t st);
  // adjoint of the map scaled by the ring weights (quadrature analysis: the inverse of synthesis
  // for band-limited maps on a Gauss-Legendre grid)
  void analysis(const double* map0, const double* map1, double2* alm, int ncomp, hipStream_t st);
  bool has_weights() const { return weight_.bytes() != 0; }
  // spin-1 plans: gradient (d/dtheta, d/dphi / sin theta) of the scalar field with coefficients alm
  void synthesis_grad(const double2* alm, double* dtheta, double* dphi, hipStream_t st);
  // spin-1 GRAD synthesis of dlm = sqrt(l(l+1)) dlm_scaled, then deflected angles (and gamma)
  void dlm2pointing(const double2* dlm_scaled, double* ptheta, double* pphi, double* gamma, hipStream_t st);

  size_t bytes() const;

 private:
  void ring_pass(int mode, double* o0, double* o1, double* o2, hipStream_t st, const double* wt = nullptr);

  int lmax_, spin_, nring_;
  long npix_;
  RingMethod method_ = RingMethod::Direct;
  LegendreTables tab_;
  DftSet<double> dft_;                                       // one DFT description per distinct ring length
  DevBuf lens_;                                              // RingLen[nlengths]
  DevBuf lenidx_;                                            // int[nring]
  DevBuf rings_[kDftMethods];                                // ring ids per DFT method
  int nrings_m_[kDftMethods] = {0, 0, 0, 0};
  int uniform_[kDftMethods] = {-1, -1, -1, -1};              // the one length index of a method's rings, or -2
  size_t lds_[kDftMethods] = {0, 0, 0, 0};
  DevBuf theta_, phi0_, ofs_, weight_;
  DevBuf sqll_;  // sqrt(l(l+1))
  DevBuf G_;     // [nrows][nring] complex (m-major)
  DevBuf coef_;
  int nrows_ = 0;
  int gstride_ = 0;
};

}  // namespace nusht
