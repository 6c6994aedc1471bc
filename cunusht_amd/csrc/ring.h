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
#include "fft_engine.h"
#include "legendre.h"

#include <map>
#include <memory>

namespace nusht {

enum class RingMethod : int { Direct = 0, Rader = 1, Bluestein = 2, Split = 3 };
constexpr int kRingMethods = 4;

// One distinct ring length and how its DFT runs in LDS (kernel-argument view).
struct RingLen {
  int n;              // ring length
  int method;         // RingMethod
  int mlen;           // FFT length run in LDS: n (direct), n - 1 (Rader), padded (Bluestein, split sub)
  int r, b;           // split: n = r * b, r sub-transforms of length b
  int sub_bluestein;  // split: sub-transforms by Bluestein (else a direct FFT of length b)
  FftDesc<double> fd; // of length mlen
  const double2* chirp;  // Bluestein: e^{-i pi j^2 / len}, len = n (or b for split)
  const double2* kern;   // Bluestein: FFT of the wrapped conjugate chirp (mlen); Rader: FFT of the permuted roots
  const int* perm_in;    // Rader: g^q mod n
  const int* perm_out;   // Rader: g^-q mod n
};

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
  int nlengths() const { return (int)lens_h_.size(); }
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
  struct LenTables {
    RingLen d{};
    DevBuf chirp, kern, perm_in, perm_out;
  };
  int add_length(int n, hipStream_t st);
  const FftLength<double>& fft_of(int len);
  void ring_pass(int mode, double* o0, double* o1, double* o2, hipStream_t st, const double* wt = nullptr);

  int lmax_, spin_, nring_;
  long npix_;
  RingMethod method_ = RingMethod::Direct;
  LegendreTables tab_;
  std::map<int, std::unique_ptr<FftLength<double>>> ffts_;  // by FFT length
  std::vector<std::unique_ptr<LenTables>> lens_h_;
  std::map<int, int> len_index_;                             // ring length -> index into lens_h_
  DevBuf lens_;                                              // RingLen[nlengths]
  DevBuf lenidx_;                                            // int[nring]
  DevBuf rings_[kRingMethods];                               // ring ids per DFT method
  int nrings_m_[kRingMethods] = {0, 0, 0, 0};
  int uniform_[kRingMethods] = {-1, -1, -1, -1};             // the one length index of a method's rings, or -2
  size_t lds_[kRingMethods] = {0, 0, 0, 0};
  DevBuf theta_, phi0_, ofs_, weight_;
  DevBuf sqll_;  // sqrt(l(l+1))
  DevBuf G_;     // [nrows][nring] complex (m-major)
  DevBuf coef_;
  int nrows_ = 0;
  int gstride_ = 0;
};

}  // namespace nusht
