// Stand-alone stage kernels (doubling, adjoint doubling, pointing) and the shared pointing math.
#pragma once

#include "common.h"

namespace nusht {

// Deflected angles for one pixel (cunusht/c/pointing.cu:205-223 formulas) and the
// polarisation rotation angle (cunusht/deflection/CPU_nufft_transformer.py:158-160):
//   a = |d|;  n1 = sin(t) cos(a) + cos(t) d_t sin(a)/a;  n2 = -d_p sin(a)/a;
//   n3 = cos(t) cos(a) - sin(t) d_t sin(a)/a;  t' = atan2(sqrt(n1^2+n2^2), n3);
//   p' = p - atan2(n2, n1), reduced by 2 pi only when >= 2 pi;
//   gamma = atan2(d_p, d_t) - atan2(d_p, a sin(a) cot(t) + d_t cos(a)).
// The same with the colatitude's sine and cosine given (constant along a ring).
__host__ __device__ inline void deflect_point_sc(double st, double ct, double p, double dt, double dp, double& tp,
                                                 double& pp, double* gamma);

__host__ __device__ inline void deflect_point(double t, double p, double dt, double dp, double& tp, double& pp,
                                              double* gamma) {
  deflect_point_sc(sin(t), cos(t), p, dt, dp, tp, pp, gamma);
}

__host__ __device__ inline void deflect_point_sc(double st, double ct, double p, double dt, double dp, double& tp,
                                                 double& pp, double* gamma) {
  const double a2 = dt * dt + dp * dp;
  const double a = sqrt(a2);
  double sa, ca;
  if (a < 1e-3) {
    sa = 1.0 - a2 / 6.0 * (1.0 - a2 / 20.0 * (1.0 - a2 / 42.0));
    ca = 1.0 - a2 / 2.0 * (1.0 - a2 / 12.0 * (1.0 - a2 / 30.0 * (1.0 - a2 / 56.0)));
  } else {
    sa = sin(a) / a;
    ca = cos(a);
  }
  const double n1 = st * ca + ct * dt * sa;
  const double n2 = -dp * sa;
  const double n3 = ct * ca - st * dt * sa;
  tp = atan2(sqrt(n1 * n1 + n2 * n2), n3);
  double q = p - atan2(n2, n1);
  pp = (q >= 2.0 * M_PI) ? q - 2.0 * M_PI : q;
  if (gamma) {
    // a sin(a) = a^2 * sin(a)/a.  Both angles have the ordinate dp, so they lie in one half-plane and their
    // difference is the principal value atan2(dp (x - dt), dt x + dp^2) (signed zeros included); at the poles
    // (cot t infinite) the two terms are kept
    const double x = a2 * sa * ct / st + dt * ca;
    *gamma = st != 0.0 ? atan2(dp * (x - dt), dt * x + dp * dp) : atan2(dp, dt) - atan2(dp, x);
  }
}

template <typename T>
void doubling(const T* map, int ntheta, int nphi, int spin, T* dmap, hipStream_t st);
template <typename T>
void adjoint_doubling(const T* dmap, int ntheta, int nphi, int spin, T* map, hipStream_t st);

// Lensing magnification determinant from convergence, shear, deflection and curl fields (w may be null).
void magnification(const double* k, const double* g1, const double* g2, const double* d1, const double* d2,
                   const double* w, long n, bool exact, double* out, hipStream_t st);
// out = fl[l] * alm (ncomp blocks of nalm complex; out may alias alm).
void almxfl(const double2* alm, int lmax, int ncomp, const double* fl, double2* out, hipStream_t st);

void pointing(const double* theta, const double* phi0, const long* nph, const long* ofs, int nring, long max_nph,
              long npix,
              const double* dt, const double* dp, double* pt, double* pp, double* gamma, hipStream_t st);

}  // namespace nusht
