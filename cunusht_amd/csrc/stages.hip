// Stand-alone stage kernels: theta doubling and its adjoint, deflected pointing, error state.
//
// Doubling replaces cunusht/c/doubling.cu:180-220 (theta-contiguous (nphi, ntheta) ->
// (nphi, 2 ntheta - 2)) and the adjoint replaces doubling.cu:53-92 (ring-major
// (2 ntheta - 2, nphi) -> (ntheta, nphi)); both carry the (-1)^s mirror sign of the spin-s
// CPU backend (cunusht/deflection/CPU_nufft_transformer.py:284-318).  Unlike the reference's
// one-thread-per-column loops, every thread here writes one element, so reads and writes are
// coalesced along the contiguous axis.
//
// Pointing replaces cunusht/c/pointing.cu:197-255 (one thread per ring, serial pixels): one
// thread per pixel, exact sin(a)/a with a series below a = 1e-3, and optional gamma
// (CPU_nufft_transformer.py:158-160).
#include "stages.h"

#include <algorithm>

#include <mutex>

namespace nusht {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

namespace {

template <typename T>
__device__ __forceinline__ T neg_if(T v, bool neg);
template <>
__device__ __forceinline__ float neg_if(float v, bool neg) { return neg ? -v : v; }
template <>
__device__ __forceinline__ double neg_if(double v, bool neg) { return neg ? -v : v; }
template <>
__device__ __forceinline__ float2 neg_if(float2 v, bool neg) { return neg ? float2{-v.x, -v.y} : v; }
template <>
__device__ __forceinline__ double2 neg_if(double2 v, bool neg) { return neg ? double2{-v.x, -v.y} : v; }
template <typename T>
__device__ __forceinline__ T add_(T a, T b) { return a + b; }
template <>
__device__ __forceinline__ float2 add_(float2 a, float2 b) { return float2{a.x + b.x, a.y + b.y}; }
template <>
__device__ __forceinline__ double2 add_(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }

template <typename T>
__global__ void k_doubling(const T* __restrict__ map, int ntheta, int nphi, bool neg, T* __restrict__ dmap) {
  const int ntd = 2 * ntheta - 2;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (r >= ntd) return;
  T v;
  if (r < ntheta) {
    v = map[(long)p * ntheta + r];
  } else {
    int q = p + nphi / 2;
    if (q >= nphi) q -= nphi;
    v = neg_if(map[(long)q * ntheta + (ntd - r)], neg);
  }
  dmap[(long)p * ntd + r] = v;
}

template <typename T>
__global__ void k_adjoint_doubling(const T* __restrict__ dmap, int ntheta, int nphi, bool neg, T* __restrict__ map) {
  const int ntd = 2 * ntheta - 2;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (p >= nphi) return;
  T v = dmap[(long)r * nphi + p];
  if (r > 0 && r < ntheta - 1) {
    int q = p + nphi / 2;
    if (q >= nphi) q -= nphi;
    v = add_(v, neg_if(dmap[(long)(ntd - r) * nphi + q], neg));
  }
  map[(long)r * nphi + p] = v;
}

// Pixels k of a ring with ofs + k outside [0, npix), or k >= max_nph, are not touched: a ring table that
// does not fit the arrays (wrong integer width, wrong offsets) leaves them unwritten instead of faulting
// (the reference guards its ring index only, cunusht/c/pointing.cu:202).
__global__ void k_pointing(const double* __restrict__ theta, const double* __restrict__ phi0,
                           const long* __restrict__ nph, const long* __restrict__ ofs, long max_nph, long npix,
                           const double* __restrict__ dt, const double* __restrict__ dp, double* __restrict__ pt,
                           double* __restrict__ pp, double* __restrict__ gamma) {
  const int ring = blockIdx.y;
  const long n = nph[ring], o = ofs[ring];
  if (n <= 0 || o < 0 || o >= npix) return;
  const long kend = min(min(n, max_nph), npix - o);
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < kend; k += (long)gridDim.x * blockDim.x) {
    const long i = o + k;
    double th = theta[ring];
    double ph = phi0[ring] + k * (2.0 * M_PI / n);
    deflect_point(th, ph, dt[i], dp[i], pt[i], pp[i], gamma ? &gamma[i] : nullptr);
  }
}

}  // namespace

template <typename T>
void doubling(const T* map, int ntheta, int nphi, int spin, T* dmap, hipStream_t st) {
  NUSHT_REQUIRE(ntheta >= 2 && nphi >= 2 && nphi % 2 == 0, kErrArg, "doubling: need ntheta >= 2 and even nphi");
  dim3 grid(ceil_div(2 * ntheta - 2, 256), nphi);
  hipLaunchKernelGGL(k_doubling<T>, grid, dim3(256), 0, st, map, ntheta, nphi, (spin & 1) != 0, dmap);
  NUSHT_LAUNCH_CHECK();
}

template <typename T>
void adjoint_doubling(const T* dmap, int ntheta, int nphi, int spin, T* map, hipStream_t st) {
  NUSHT_REQUIRE(ntheta >= 2 && nphi >= 2 && nphi % 2 == 0, kErrArg, "adjoint_doubling: need ntheta >= 2 and even nphi");
  dim3 grid(ceil_div(nphi, 256), ntheta);
  hipLaunchKernelGGL(k_adjoint_doubling<T>, grid, dim3(256), 0, st, dmap, ntheta, nphi, (spin & 1) != 0, map);
  NUSHT_LAUNCH_CHECK();
}

template void doubling<float>(const float*, int, int, int, float*, hipStream_t);
template void doubling<double>(const double*, int, int, int, double*, hipStream_t);
template void doubling<float2>(const float2*, int, int, int, float2*, hipStream_t);
template void doubling<double2>(const double2*, int, int, int, double2*, hipStream_t);
template void adjoint_doubling<float>(const float*, int, int, int, float*, hipStream_t);
template void adjoint_doubling<double>(const double*, int, int, int, double*, hipStream_t);
template void adjoint_doubling<float2>(const float2*, int, int, int, float2*, hipStream_t);
template void adjoint_doubling<double2>(const double2*, int, int, int, double2*, hipStream_t);

// ------------------------------------------------------------------ lensing magnification
// Determinant of the magnification matrix of the deflection (cunusht/helper_GPU.py:113-161):
//   A = f0 ((1-k)^2 - g1^2 - g2^2 + w^2) + f1 (1-k) + sg (f1/|d|^2) ((d1^2 - d2^2) g1 + 2 d1 d2 g2),
//   f0 = sin|d|/|d|, f1 = cos|d| - f0, series for |d| < 1e-4 (f0 = 1 - d^2/6, f1/d^2 = -1/3 + d^2/30).
// sg = +1 gives the solid-angle ratio dOmega'/dOmega of the map n -> cos|d| n + sin|d| d_hat
// exactly; sg = -1 is the reference's sign, correct through second order in |d|.
__global__ void k_magnification(const double* __restrict__ k, const double* __restrict__ g1,
                                const double* __restrict__ g2, const double* __restrict__ d1,
                                const double* __restrict__ d2, const double* __restrict__ w, long n, double sg,
                                double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double a = d1[i], b = d2[i], kk = k[i], s1 = g1[i], s2 = g2[i];
    const double ww = w ? w[i] : 0.0;
    const double dd = a * a + b * b;
    const double d = sqrt(dd);
    double f0, f1, f1od2;
    if (d < 1e-4) {
      f0 = 1.0 - dd / 6.0;
      f1od2 = -1.0 / 3.0 + dd / 30.0;
      f1 = f1od2 * dd;
    } else {
      f0 = sin(d) / d;
      f1 = cos(d) - f0;
      f1od2 = f1 / dd;
    }
    const double omk = 1.0 - kk;
    out[i] = f0 * (omk * omk - s1 * s1 - s2 * s2 + ww * ww) + f1 * omk +
             sg * f1od2 * ((a * a - b * b) * s1 + 2.0 * a * b * s2);
  }
}

void magnification(const double* k, const double* g1, const double* g2, const double* d1, const double* d2,
                   const double* w, long n, bool exact, double* out, hipStream_t st) {
  NUSHT_REQUIRE(n >= 0, kErrArg, "magnification: negative size");
  if (n == 0) return;
  hipLaunchKernelGGL(k_magnification, dim3((unsigned)std::min<long>(ceil_div(n, 256), 65536)), dim3(256), 0, st, k,
                     g1, g2, d1, d2, w, n, exact ? 1.0 : -1.0, out);
  NUSHT_LAUNCH_CHECK();
}

// out_lm = fl[l] alm_lm over the healpy m-major layout (out may alias alm)
__global__ void k_almxfl(const double2* __restrict__ alm, int lmax, int ncomp, const double* __restrict__ fl,
                         double2* __restrict__ out) {
  const int m = blockIdx.y;
  const long nalm = (long)(lmax + 1) * (lmax + 2) / 2;
  const long base = (long)m * (2 * lmax + 1 - m) / 2;
  for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l <= lmax; l += gridDim.x * blockDim.x) {
    const double f = fl[l];
    for (int c = 0; c < ncomp; ++c) {
      const double2 v = alm[c * nalm + base + l];
      out[c * nalm + base + l] = double2{v.x * f, v.y * f};
    }
  }
}

void almxfl(const double2* alm, int lmax, int ncomp, const double* fl, double2* out, hipStream_t st) {
  NUSHT_REQUIRE(lmax >= 0 && ncomp >= 1, kErrArg, "almxfl: bad lmax or ncomp");
  hipLaunchKernelGGL(k_almxfl, dim3(ceil_div(lmax + 1, 256), lmax + 1), dim3(256), 0, st, alm, lmax, ncomp, fl, out);
  NUSHT_LAUNCH_CHECK();
}

void pointing(const double* theta, const double* phi0, const long* nph, const long* ofs, int nring, long max_nph,
              long npix, const double* dt, const double* dp, double* pt, double* pp, double* gamma, hipStream_t st) {
  NUSHT_REQUIRE(nring >= 1 && nring <= 65535, kErrArg, "pointing: nring must be in [1, 65535]");
  NUSHT_REQUIRE(max_nph >= 1 && npix >= 1, kErrArg, "pointing: max_nph and npix must be >= 1");
  dim3 grid(ceil_div(max_nph, 256), nring);
  hipLaunchKernelGGL(k_pointing, grid, dim3(256), 0, st, theta, phi0, nph, ofs, max_nph, npix, dt, dp, pt, pp,
                     gamma);
  NUSHT_LAUNCH_CHECK();
}

}  // namespace nusht
