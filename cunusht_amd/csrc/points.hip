// Exact synthesis at arbitrary points (see points.h).
#include "points.h"

#include <map>
#include <mutex>

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

// ------------------------------------------------------------------ allocation-free variant
//
// sY_lm(theta, 0) = (-1)^s N_l d^l_{m,-s}(theta), N_l = sqrt((2l+1)/4pi), and for m < 0
// sY_{l,m} = (-1)^m N_l d^l_{|m|,+s}: every row needs lambda_l = N_l d^l_{a,b}(theta) with a = |m| >= 0 and
// b = -s (m >= 0) or +s (m < 0), for l from l0 = max(a, s) to lmax.
// Seed at l0: for a >= s the product over a' = s+1..a of -sqrt(2a'(2a'+1)/(a'^2 - s^2)) sin(t/2) cos(t/2)
// from N_s sin^{2s}(t/2) (b = -s) or N_s cos^{2s}(t/2) (b = +s); for a < s (l0 = s) the closed forms
// N_s sqrt(C(2s, s+a)) cos^{s-a} sin^{s+a} (-1)^{a+s} (b = -s) and N_s sqrt(C(2s, s+a)) cos^{s+a} sin^{s-a}.
// Three-term recursion in l (textbook Wigner-d relation, normalised):
//   lambda_{l+1} = (alpha_l x - beta_l) lambda_l - gamma_l lambda_{l-1},  x = cos theta,
//   alpha_l = f (l+1) / d1,  beta_l = f a b / (l d1),  gamma_l = sqrt((2l+3)/(2l-1)) (l+1) d0 / (l d1),
//   f = sqrt((2l+3)(2l+1)), d1 = sqrt(((l+1)^2-a^2)((l+1)^2-b^2)), d0 = sqrt((l^2-a^2)(l^2-b^2)).
// Values below 2^-300 are carried as v 2^(-600 k) and add nothing until k returns to 0.
// G rows: spin 0 -> G_m = sum_l a_lm lambda_l (m = 0..lmax); spin s -> G_m = -sgn sum_l (e + i b) lambda_l with
// (e, b) = (E, B) for m >= 0 and (-1)^m (conj E, conj B) for m < 0, sgn = (-1)^s (m >= 0) or (-1)^m (m < 0).

namespace {

static __device__ __align__(256) double2 scratch[kNoAllocG];

constexpr int kNaChunk = 64;  // l values whose coefficients one wave stages in LDS at a time

__global__ __launch_bounds__(64) void k_point_rows(int lmax, int spin, int ncomp, long npts,
                                                   const double* __restrict__ theta, const double2* __restrict__ alm,
                                                   double2* __restrict__ G) {
  __shared__ double ca[kNaChunk], cb[kNaChunk], cg[kNaChunk];
  __shared__ double2 cw[kNaChunk];
  const int row = blockIdx.y;
  const int m = spin ? row - lmax : row;
  const int a = m < 0 ? -m : m;
  const int b = m >= 0 ? -spin : spin;
  const int s = spin;
  const int l0 = a > s ? a : s;
  const long p = (long)blockIdx.x * kNaChunk + threadIdx.x;
  const bool live = p < npts;
  const double th = live ? theta[p] : 0.5 * M_PI;
  const double x = cos(th), sh = sin(0.5 * th), ch = cos(0.5 * th);
  const double ns = sqrt((2.0 * s + 1.0) / (4.0 * M_PI));
  // seed lambda_{l0} = v * 2^(-600 k)
  double v;
  int k = 0;
  if (a >= s) {
    v = ns;
    for (int j = 0; j < 2 * s; ++j) v *= (b < 0 ? sh : ch);
    for (int j = s + 1; j <= a; ++j) {
      v *= -sqrt(2.0 * j * (2.0 * j + 1.0) / ((double)j * j - (double)s * s)) * sh * ch;
      if (v != 0.0 && fabs(v) < 0x1p-300) {
        v *= 0x1p600;
        ++k;
      }
    }
  } else {
    double bin = 1.0;  // C(2s, s+a)
    for (int j = 1; j <= s - a; ++j) bin = bin * (s + a + j) / j;
    v = ns * sqrt(bin);
    if (b < 0) {
      for (int j = 0; j < s - a; ++j) v *= ch;
      for (int j = 0; j < s + a; ++j) v *= sh;
      if ((a + s) & 1) v = -v;
    } else {
      for (int j = 0; j < s + a; ++j) v *= ch;
      for (int j = 0; j < s - a; ++j) v *= sh;
    }
  }
  const long arow = (long)a * (2 * lmax + 1 - a) / 2;  // healpy index of (l = 0, m = a)
  const long nalm = (long)(lmax + 1) * (lmax + 2) / 2;
  const double sg = (a & 1) ? -1.0 : 1.0;               // (-1)^m for m < 0
  double vp = 0.0, accr = 0.0, acci = 0.0;
  for (int lb = l0; lb <= lmax; lb += kNaChunk) {
    // stage the chunk's recursion coefficients (step l -> l + 1) and alm combinations
    {
      const int l = lb + (int)threadIdx.x;
      if (l <= lmax) {
        double al, be, ga;
        if (l == 0) {
          al = sqrt(3.0);
          be = ga = 0.0;
        } else {
          const double d1 = sqrt(((l + 1.0) * (l + 1.0) - (double)a * a) * ((l + 1.0) * (l + 1.0) - (double)b * b));
          const double d0 = sqrt(((double)l * l - (double)a * a) * ((double)l * l - (double)b * b));
          const double f = sqrt((2.0 * l + 3.0) * (2.0 * l + 1.0));
          al = d1 > 0.0 ? f * (l + 1.0) / d1 : 0.0;
          be = d1 > 0.0 ? f * (double)a * b / (l * d1) : 0.0;
          ga = d1 > 0.0 ? sqrt((2.0 * l + 3.0) / (2.0 * l - 1.0)) * (l + 1.0) * d0 / (l * d1) : 0.0;
        }
        ca[threadIdx.x] = al;
        cb[threadIdx.x] = be;
        cg[threadIdx.x] = ga;
        const double2 e = alm[arow + l];
        double2 w;
        if (spin == 0) {
          w = e;
        } else {
          const double2 bb = ncomp > 1 ? alm[nalm + arow + l] : double2{0.0, 0.0};
          w = m >= 0 ? double2{e.x - bb.y, e.y + bb.x} : double2{sg * (e.x + bb.y), sg * (bb.x - e.y)};
        }
        cw[threadIdx.x] = w;
      }
    }
    __syncthreads();
    const int n = min(kNaChunk, lmax - lb + 1);
    for (int j = 0; j < n; ++j) {
      if (k == 0) {
        const double2 w = cw[j];
        accr = fma(w.x, v, accr);
        acci = fma(w.y, v, acci);
      }
      const double nv = fma(fma(ca[j], x, -cb[j]), v, -cg[j] * vp);
      vp = v;
      v = nv;
      if (k > 0 && fabs(v) > 0x1p300) {
        v *= 0x1p-600;
        vp *= 0x1p-600;
        --k;
      }
    }
    __syncthreads();
  }
  if (!live) return;
  double pref = 1.0;
  if (spin) pref = -(m >= 0 ? ((s & 1) ? -1.0 : 1.0) : sg);
  G[(long)row * npts + p] = double2{pref * accr, pref * acci};
}

std::mutex g_scratch_mu;
std::map<int, double2*> g_scratch_ptr;  // device ordinal -> address of that device's `scratch`

// Device address of this GPU's scratch array, looked up once per device (with that device current).
double2* scratch_ptr() {
  int dev = 0;
  NUSHT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* ptr = nullptr;
  NUSHT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  NUSHT_REQUIRE(ptr != nullptr, kErrHip, "no device address for the point-synthesis scratch array");
  g_scratch_ptr[dev] = static_cast<double2*>(ptr);
  return static_cast<double2*>(ptr);
}

}  // namespace

void point_synthesis_noalloc(int lmax, int spin, long npts, const double* theta, const double2* alm, int ncomp,
                             const double* phi, const double* gamma, double* out0, double* out1, hipStream_t st) {
  NUSHT_REQUIRE(theta && alm && phi && out0, kErrArg, "point synthesis: null array");
  NUSHT_REQUIRE(lmax >= 0 && spin >= 0 && spin <= 3 && lmax >= spin, kErrArg,
                "point synthesis: need 0 <= spin <= 3 and lmax >= spin");
  NUSHT_REQUIRE(spin == 0 || out1, kErrArg, "point synthesis: spin > 0 needs two outputs");
  NUSHT_REQUIRE(ncomp == 1 || (spin > 0 && ncomp == 2), kErrArg, "point synthesis: ncomp must be 1, or 2 for spin > 0");
  NUSHT_REQUIRE(npts >= 1, kErrArg, "point synthesis: need at least one point");
  const long nrows = spin ? 2L * lmax + 1 : lmax + 1L;
  NUSHT_REQUIRE(nrows * npts <= kNoAllocG && npts <= 65535L * kNaChunk, kErrUnsupported,
                "point synthesis without allocation: (rows = " + std::to_string(nrows) + ") x (points = " +
                    std::to_string(npts) + ") exceeds the " + std::to_string(kNoAllocG) +
                    " ring coefficients of the scratch array");
  NUSHT_REQUIRE(nrows <= 65535, kErrUnsupported, "point synthesis without allocation: lmax too large");
  double2* G = scratch_ptr();
  {
    StageScope sc("point_rows_noalloc", st);
    hipLaunchKernelGGL(k_point_rows, dim3((unsigned)ceil_div(npts, kNaChunk), (unsigned)nrows), dim3(kNaChunk), 0, st,
                       lmax, spin, ncomp, npts, theta, alm, G);
    NUSHT_LAUNCH_CHECK();
  }
  StageScope sc("point_sum", st);
  hipLaunchKernelGGL(k_point_sum, dim3((unsigned)npts), dim3(256), 0, st, G, npts, lmax, spin, npts, phi, gamma, out0,
                     out1);
  NUSHT_LAUNCH_CHECK();
}

}  // namespace nusht
