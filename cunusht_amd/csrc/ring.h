// SHT on iso-latitude rings with uniform phi sampling (the lensing/pointing grid), FP64.
//
// Synthesis = Legendre ring coefficients G_m(theta_r) followed by one length-nph DFT per ring;
// the adjoint runs the transposes.  Ring lengths fixed by the caller (nph = 2 lmax + 1 is
// typical: 8191 is prime, 511 = 7 * 73, ...) are handled in LDS by a direct mixed-radix FFT
// when nph factors into 2, 3, 5, 7, 11, 13, by Rader's algorithm when nph is prime with a
// smooth nph - 1, and by Bluestein's chirp-z algorithm otherwise.  For spin 1 the deflected
// pointing (theta', phi', gamma) is computed straight from the ring values in LDS, so the
// deflection field itself is never written to HBM.
#pragma once

#include "common.h"
#include "fft_engine.h"
#include "legendre.h"

namespace nusht {

enum class RingMethod : int { Direct = 0, Rader = 1, Bluestein = 2 };

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
  RingMethod method() const { return method_; }
  const LegendreTables& tables() const { return tab_; }

  // alm (ncomp x nalm complex) -> map0 (spin 0: real map; spin s: Q) and map1 (U)
  void synthesis(const double2* alm, int ncomp, double* map0, double* map1, hipStream_t st);
  // exact adjoint
  void adjoint(const double* map0, const double* map1, double2* alm, int ncomp, hipStream_t st);
  // adjoint of the map scaled by the ring weights (quadrature analysis: the inverse of synthesis
  // for band-limited maps on a Gauss-Legendre grid)
  void analysis(const double* map0, const double* map1, double2* alm, int ncomp, hipStream_t st);
  bool has_weights() const { return weight_.bytes() != 0; }
  // spin-1 GRAD synthesis of dlm = sqrt(l(l+1)) dlm_scaled, then deflected angles (and gamma)
  void dlm2pointing(const double2* dlm_scaled, double* ptheta, double* pphi, double* gamma, hipStream_t st);

  size_t bytes() const;

 private:
  void ring_pass(int mode, double* o0, double* o1, double* o2, hipStream_t st, const double* wt = nullptr);
  int lmax_, spin_, nring_, nphi_;
  long npix_;
  RingMethod method_ = RingMethod::Direct;
  int mlen_ = 0;  // FFT length used inside the ring transform
  LegendreTables tab_;
  FftLength<double> fft_;
  DevBuf theta_, phi0_, ofs_, weight_;
  DevBuf chirp_, kern_, perm_in_, perm_out_;  // Bluestein / Rader tables
  DevBuf sqll_;                               // sqrt(l(l+1))
  DevBuf G_;                                  // [nrows][nring] complex (m-major)
  DevBuf coef_;
  int nrows_ = 0;
  int gstride_ = 0;
};

}  // namespace nusht
