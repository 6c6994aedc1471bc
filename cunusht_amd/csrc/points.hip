// Exact synthesis at arbitrary points (see points.h).
#include "points.h"

#include "timing.h"

namespace nusht {
namespace {

// One workgroup per point: threads stride over the rows of G, block reduction of the sum.
__global__ __launch_bounds__(256) void k_point_sum(const double2* __restrict__ G, long gstride, int lmax, int spin,
                                                   long npts, const double* __restrict__ phi,
                                                   const double* __restrict__ gamma, double* __restrict__ o0,
                                                   double* __restrict__ o1) {
  __shared__ double red[2][256];
  const long p = blockIdx.x;
  const double ph = phi[p];
  const int nrows = spin ? 2 * lmax + 1 : lmax + 1;
  double sr = 0.0, si = 0.0;
  for (int row = threadIdx.x; row < nrows; row += blockDim.x) {
    const int m = spin ? row - lmax : row;
    const double2 g = G[(long)row * gstride + p];
    double sn, cs;
    sincos((double)m * ph, &sn, &cs);
    const double vr = g.x * cs - g.y * sn, vi = g.x * sn + g.y * cs;
    if (spin) {
      sr += vr;
      si += vi;
    } else {
      sr += (m == 0 ? 1.0 : 2.0) * vr;  // real field: m and -m
    }
  }
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double q = red[0][0], u = red[1][0];
  if (!spin) {
    o0[p] = q;
    return;
  }
  if (gamma) {
    double sn, cs;
    sincos(spin * gamma[p], &sn, &cs);
    const double x = q * cs - u * sn, y = q * sn + u * cs;
    q = x;
    u = y;
  }
  o0[p] = q;
  o1[p] = u;
}

}  // namespace

PointPlan::PointPlan(int lmax, int spin, const std::vector<double>& theta, hipStream_t st)
    : lmax_(lmax), spin_(spin), npts_((long)theta.size()) {
  NUSHT_REQUIRE(npts_ >= 1 && npts_ < (1L << 31), kErrArg, "pointplan: need 1 .. 2^31 points");
  for (double t : theta) NUSHT_REQUIRE(t >= 0.0 && t <= M_PI, kErrArg, "pointplan: theta outside [0, pi]");
  tab_.tag = "points";
  tab_.build(lmax, spin, theta, st);
  nrows_ = spin ? 2 * lmax + 1 : lmax + 1;
  G_.alloc(sizeof(double2) * (size_t)nrows_ * npts_);
  coef_.alloc(legendre_coef_bytes(tab_, 1));
  tab_.variant = 0;
}

void PointPlan::synthesis(const double2* alm, int ncomp, const double* phi, const double* gamma, double* out0,
                          double* out1, hipStream_t st) {
  NUSHT_REQUIRE(alm && phi && out0, kErrArg, "pointplan: null array");
  NUSHT_REQUIRE(spin_ == 0 || out1, kErrArg, "pointplan: spin > 0 needs two outputs");
  NUSHT_REQUIRE(ncomp == 1 || (spin_ > 0 && ncomp == 2), kErrArg, "pointplan: ncomp must be 1, or 2 for spin > 0");
  legendre_synth(tab_, 1, ncomp, alm, 0, nullptr, coef_.as<double2>(), G_.as<double2>(), npts_, 1, 0, st);
  StageScope sc("point_sum", st);
  hipLaunchKernelGGL(k_point_sum, dim3((unsigned)npts_), dim3(256), 0, st, G_.as<double2>(), npts_, lmax_, spin_,
                     npts_, phi, gamma, out0, out1);
  NUSHT_LAUNCH_CHECK();
}

}  // namespace nusht
