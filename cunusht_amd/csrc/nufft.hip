// Double-Fourier-sphere passes and ES nuFFT (see nufft.h).
//
// Kernel parameters follow the published ES-kernel method (Barnett, Magland & af Klinteberg,
// SIAM J. Sci. Comput. 41 (2019)): for upsampling sigma, width w = ceil(ln(1/eps) / (pi sqrt(1 - 1/sigma))),
// beta = 0.97 pi w (1 - 1/(2 sigma)); deconvolution by the kernel's Fourier transform, here
// integrated with Gauss-Legendre quadrature on the host.  Point coordinates stay FP64 in every
// precision: the integer cell and the fractional offset are formed in FP64 before the kernel
// is evaluated in the working precision.
#include "nufft.h"

#include "timing.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace nusht {
namespace {

constexpr int kMaxW = 24;
constexpr int kLdsBytes = 160 * 1024;

template <typename T>
__device__ __forceinline__ T es_kernel(T z, T beta) {
  T t = T(1) - z * z;
  return t > T(0) ? exp(beta * (sqrt(t) - T(1))) : T(0);
}

template <typename T>
struct WorkT;  // working precision helpers
template <>
struct WorkT<float> {
  __device__ static float ker(float z, float beta) {
    float t = 1.0f - z * z;
    return t > 0.0f ? __expf(beta * (__fsqrt_rn(t) - 1.0f)) : 0.0f;
  }
};
template <>
struct WorkT<double> {
  __device__ static double ker(double z, double beta) {
    double t = 1.0 - z * z;
    return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
  }
};

// ------------------------------------------------------------------ theta pass, type 2
template <typename T>
struct ThetaArgs {
  int lmax, spin, ntheta, ntd, nf, nm, row_lo, nrow, nmrows;
  const double2* G;
  long gsm;
  const double* deconv;
  FftDesc<T> ftd, fnf;
  c2_t<T>* Tout;  // [nrow][nm]
};

// Two columns per workgroup.  The doubled sequence of column m is even or odd (x_{-n} = +-x_n, sign (-1)^(s+m)),
// and rows 2p, 2p + 1 have opposite signs; so one transform pair serves both: with X = XA + XB and
// XA_{-k} = sA XA_k, XB_{-k} = -sA XB_k, the deconvolved spectrum d_k (fA XA_k + fB XB_k) is
// d_k (cp X_k + cm X_{-k}), cp = (fA + fB) / 2, cm = sA (fA - fB) / 2, and the inverse transform T = TA + TB
// separates by TA_t = (T_t + sA T_{-t}) / 2, TB_t = (T_t - sA T_{-t}) / 2.
template <typename T, int NT, int NMAX>
__global__ __launch_bounds__(NT) void k_theta2(ThetaArgs<T> a) {
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  T2* buf = reinterpret_cast<T2*>(smem_raw);
  const int rowA = 2 * xcd_remap(blockIdx.x, gridDim.x), rowB = rowA + 1;
  const bool hasB = rowB < a.nmrows;
  const int mA = a.spin ? rowA - a.lmax : rowA, mB = mA + 1;
  const int amA = mA < 0 ? -mA : mA, amB = mB < 0 ? -mB : mB;
  const double sA = ((a.spin + amA) & 1) ? -1.0 : 1.0;
  const double2* gA = a.G + (long)rowA * a.gsm;
  const double2* gB = a.G + (long)rowB * a.gsm;
  // ring r and its mirror ntd - r (ntheta - 1 = ntd / 2 rings apart from the poles)
  for (int r = threadIdx.x; r < a.ntheta; r += NT) {
    double2 u = gA[r], w{0.0, 0.0};
    if (hasB) w = gB[r];
    buf[r] = T2{(T)(u.x + w.x), (T)(u.y + w.y)};
    if (r > 0 && r < a.ntheta - 1) buf[a.ntd - r] = T2{(T)(sA * (u.x - w.x)), (T)(sA * (u.y - w.y))};
  }
  __syncthreads();
  fft_lds<T, NT, NMAX, false, true, std::is_same<T, float>::value>(buf, a.ftd);
  // split the parities, deconvolve, pad to nf (conjugated for the inverse transform)
  constexpr int E = (NMAX + NT - 1) / NT;
  T2 v[E];
  const double fA = a.deconv[amA] / a.ntd, fB = hasB ? a.deconv[amB] / a.ntd : 0.0;
  const double cp = 0.5 * (fA + fB), cm = 0.5 * sA * (fA - fB);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = threadIdx.x + i * NT;
    v[i] = T2{T(0), T(0)};
    if (k < a.nf) {
      int kk = -1, r = 0, rm = 0;  // |frequency|, its buffer index and the mirrored one
      if (k <= a.lmax) {
        kk = k;
        r = k;
        rm = k == 0 ? 0 : a.ntd - k;
      } else if (k >= a.nf - a.lmax) {
        kk = a.nf - k;
        r = a.ntd - kk;
        rm = kk;
      }
      if (kk >= 0) {
        const T2 s = buf[r], u = buf[rm];
        const double d = a.deconv[kk];
        const T fp = (T)(cp * d), fn = (T)(cm * d);
        v[i] = T2{fp * s.x + fn * u.x, -(fp * s.y + fn * u.y)};
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = threadIdx.x + i * NT;
    if (k < a.nf) buf[k] = v[i];
  }
  __syncthreads();
  fft_lds<T, NT, NMAX, false, true, std::is_same<T, float>::value>(buf, a.fnf);
  // rows of the kept band, both columns; spin 0 also fills the conjugate columns.  A lane pair shares a row: the
  // even lane writes column A, the odd lane column B, so the row's two values leave as one contiguous segment
  // (half the rows per store instruction of one value per lane).
  const int half = threadIdx.x & 1;
  const T hs = (T)(half ? -0.5 * sA : 0.5 * sA);
  const int m = half ? mB : mA;
  for (int q = threadIdx.x >> 1; q < a.nrow; q += NT / 2) {
    int t = a.row_lo + q;
    t %= a.nf;
    if (t < 0) t += a.nf;
    const T2 s = buf[t], u = buf[t == 0 ? 0 : a.nf - t];
    const T2 val{T(0.5) * s.x + hs * u.x, -(T(0.5) * s.y + hs * u.y)};
    T2* orow = a.Tout + (long)q * a.nm;
    if (!half || hasB) {
      orow[m + a.lmax] = val;
      if (!a.spin && m > 0) orow[a.lmax - m] = T2{val.x, -val.y};
    }
  }
}

// ------------------------------------------------------------------ phi pass, type 2
template <typename T>
struct PhiArgs {
  int lmax, nf, nm, nrow;
  FftDesc<T> fnf;
  c2_t<T>* Tbuf;  // [nrow][nm]
  c2_t<T>* grid;  // [nrow][nf]
};

template <typename T, int NT, int NMAX>
__global__ __launch_bounds__(NT) void k_phi2(PhiArgs<T> a) {
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  T2* buf = reinterpret_cast<T2*>(smem_raw);
  const int q = blockIdx.x;
  const T2* trow = a.Tbuf + (long)q * a.nm;
  // Every load of the thread is issued before the first LDS write (nf <= NMAX): the loads are unconditional,
  // from clamped in-range addresses, and positions outside the band are zeroed afterwards.
  constexpr int E = (NMAX + NT - 1) / NT;
  T2 v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = threadIdx.x + i * NT;
    const bool in = k <= a.lmax || (k < a.nf && k >= a.nf - a.lmax);
    const int m = k <= a.lmax ? k : k - a.nf;
    v[i] = trow[in ? m + a.lmax : 0];
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = threadIdx.x + i * NT;
    const bool in = k <= a.lmax || (k < a.nf && k >= a.nf - a.lmax);
    v[i] = in ? T2{v[i].x, -v[i].y} : T2{T(0), T(0)};
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int k = threadIdx.x + i * NT;
    if (k < a.nf) buf[k] = v[i];
  }
  __syncthreads();
  fft_lds<T, NT, NMAX, false, true, std::is_same<T, float>::value>(buf, a.fnf);
  T2* grow = a.grid + (long)q * a.nf;
  for (int k = threadIdx.x; k < a.nf; k += NT) {
    T2 s = buf[k];
    grow[k] = T2{s.x, -s.y};
  }
}

// ------------------------------------------------------------------ phi pass, type 1 (adjoint)
template <typename T, int NT, int NMAX>
__global__ __launch_bounds__(NT) void k_phi1(PhiArgs<T> a) {
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  T2* buf = reinterpret_cast<T2*>(smem_raw);
  const int q = blockIdx.x;
  const T2* grow = a.grid + (long)q * a.nf;
  for (int k = threadIdx.x; k < a.nf; k += NT) buf[k] = grow[k];
  __syncthreads();
  fft_lds<T, NT, NMAX, false, true, std::is_same<T, float>::value>(buf, a.fnf);
  T2* trow = a.Tbuf + (long)q * a.nm;
  for (int i = threadIdx.x; i < a.nm; i += NT) {
    int m = i - a.lmax;
    int k = m >= 0 ? m : a.nf + m;
    trow[i] = buf[k];
  }
}

// ------------------------------------------------------------------ theta pass, type 1 (adjoint)
// Exact transpose of k_theta2, two columns per workgroup: the kept rows of both columns enter the nf-point
// buffer through the transposed parity split (u_t = (uA_t + sA uA_{-t}) / 2 + (uB_t - sA uB_{-t}) / 2), the
// spectrum through the transposed deconvolution d_k (cp Y_k + cm Y_{-k}), and the two columns leave through
// their own foldings (signs sA and -sA).
template <typename T, int NT, int NMAX>
__global__ __launch_bounds__(NT) void k_theta1(ThetaArgs<T> a) {
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  T2* buf = reinterpret_cast<T2*>(smem_raw);
  const int rowA = 2 * xcd_remap(blockIdx.x, gridDim.x), rowB = rowA + 1;
  const bool hasB = rowB < a.nmrows;
  const int mA = a.spin ? rowA - a.lmax : rowA, mB = mA + 1;
  const int amA = mA < 0 ? -mA : mA, amB = mB < 0 ? -mB : mB;
  const double sA = ((a.spin + amA) & 1) ? -1.0 : 1.0;
  const T hs = (T)(0.5 * sA);
  const T2* colA = a.Tout + mA + a.lmax;
  const T2* colB = a.Tout + mB + a.lmax;
  // positions k and -k (the same for k = 0 and nf / 2) from the kept rows there, if any.  Every load of the
  // thread is issued before the first LDS write (nf / 2 + 1 <= NMAX / 2 + 1 positions): the loads are
  // unconditional, from clamped rows (row 0 and column A stand in), and zeroed afterwards.
  {
    constexpr int EK = (NMAX / 2 + NT) / NT;
    const T2* colX = hasB ? colB : colA;
    T2 uA[EK], uAm[EK], uB[EK], uBm[EK];
#pragma unroll
    for (int i = 0; i < EK; ++i) {
      const int k = threadIdx.x + i * NT;
      const int km = k == 0 ? 0 : a.nf - k;
      int q = k - a.row_lo;
      if (q >= a.nf) q -= a.nf;
      int qm = km - a.row_lo;
      if (qm >= a.nf) qm -= a.nf;
      const bool in = k <= a.nf / 2;
      const long oq = in && q < a.nrow ? (long)q * a.nm : 0, oqm = in && qm < a.nrow ? (long)qm * a.nm : 0;
      uA[i] = colA[oq];
      uAm[i] = colA[oqm];
      uB[i] = colX[oq];
      uBm[i] = colX[oqm];
    }
#pragma unroll
    for (int i = 0; i < EK; ++i) {
      const int k = threadIdx.x + i * NT;
      const int km = k == 0 ? 0 : a.nf - k;
      int q = k - a.row_lo;
      if (q >= a.nf) q -= a.nf;
      int qm = km - a.row_lo;
      if (qm >= a.nf) qm -= a.nf;
      const bool in = k <= a.nf / 2;
      const T2 zero{T(0), T(0)};
      if (!(in && q < a.nrow)) uA[i] = uB[i] = zero;
      if (!(in && qm < a.nrow)) uAm[i] = uBm[i] = zero;
      if (!hasB) uB[i] = uBm[i] = zero;
    }
#pragma unroll
    for (int i = 0; i < EK; ++i) {
      const int k = threadIdx.x + i * NT;
      if (k <= a.nf / 2) {
        const int km = k == 0 ? 0 : a.nf - k;
        buf[k] = T2{T(0.5) * (uA[i].x + uB[i].x) + hs * (uAm[i].x - uBm[i].x),
                    T(0.5) * (uA[i].y + uB[i].y) + hs * (uAm[i].y - uBm[i].y)};
        if (km != k)
          buf[km] = T2{T(0.5) * (uAm[i].x + uBm[i].x) + hs * (uA[i].x - uB[i].x),
                       T(0.5) * (uAm[i].y + uBm[i].y) + hs * (uA[i].y - uB[i].y)};
      }
    }
  }
  __syncthreads();
  fft_lds<T, NT, NMAX, false, true, std::is_same<T, float>::value>(buf, a.fnf);
  constexpr int E = (NMAX + NT - 1) / NT;
  T2 v[E];
  const double fA = a.deconv[amA] / a.ntd, fB = hasB ? a.deconv[amB] / a.ntd : 0.0;
  const double cp = 0.5 * (fA + fB), cm = 0.5 * sA * (fA - fB);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int r = threadIdx.x + i * NT;
    v[i] = T2{T(0), T(0)};
    if (r < a.ntd) {
      int kk = -1, pos = 0, posm = 0;  // |frequency|, its nf-buffer position and the mirrored one
      if (r <= a.lmax) {
        kk = r;
        pos = r;
        posm = r == 0 ? 0 : a.nf - r;
      } else if (r >= a.ntd - a.lmax) {
        kk = a.ntd - r;
        pos = a.nf - kk;
        posm = kk;
      }
      if (kk >= 0) {
        const T2 s = buf[pos], u = buf[posm];
        const double d = a.deconv[kk];
        const T fp = (T)(cp * d), fn = (T)(cm * d);
        v[i] = T2{fp * s.x + fn * u.x, -(fp * s.y + fn * u.y)};
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int r = threadIdx.x + i * NT;
    if (r < a.ntd) buf[r] = v[i];
  }
  __syncthreads();
  fft_lds<T, NT, NMAX, false, true, std::is_same<T, float>::value>(buf, a.ftd);
  // fold back onto the ntheta rings (conjugate back from the inverse trick), both columns
  double2* gA = const_cast<double2*>(a.G) + (long)rowA * a.gsm;
  double2* gB = const_cast<double2*>(a.G) + (long)rowB * a.gsm;
  for (int j = threadIdx.x; j < a.ntheta; j += NT) {
    const T2 s = buf[j];
    const double re = (double)s.x, im = -(double)s.y;
    double ur = 0.0, ui = 0.0;
    if (j > 0 && j < a.ntheta - 1) {
      const T2 u = buf[a.ntd - j];
      ur = sA * (double)u.x;
      ui = -sA * (double)u.y;
    }
    gA[j] = double2{re + ur, im + ui};
    if (hasB) gB[j] = double2{re - ur, im - ui};
  }
}

// ------------------------------------------------------------------ interpolation (type 2)
template <typename T>
struct InterpArgs {
  int nf, w, row_lo, nrow, spin, tr, tc, ntc, ntiles;
  T beta;
  const T* kpoly;  // [(deg + 1)][w] kernel polynomials, one per footprint cell
  int deg;
  long npts;
  const double2* pts;
  const int* idx;
  const int* tile_start;
  const c2_t<T>* grid;
  const double* gamma;
  double* out_re;
  double* out_im;
};

// Kernel weights of one point along both axes for a compile-time width W:
// k[i] = psi((f + i) 2/W) with f = (first cell) - (position) in [-W/2, -W/2 + 1), from the plan's
// Horner polynomials in u = f + W/2 - 1/2 (one per cell, coefficients wave-uniform).  The two outer
// cells hold the square-root branch point of psi at |z| = 1, where a polynomial converges only
// algebraically; there psi is below e^{-beta} * beta (~1e-8 at eps 1e-7) and the plan's fit still meets
// max(eps/100, 1e-15) with the same degree as the interior cells (checked at plan time on 513 samples
// per cell), so no exp / sqrt is evaluated per point.
template <typename T, int W>
__device__ __forceinline__ void ker_weights(double ft, double fp, T beta, const T* __restrict__ cp, int deg,
                                            T (&kt)[W], T (&kp)[W]) {
  (void)beta;
  const T ut = (T)(ft + 0.5 * W - 0.5), up = (T)(fp + 0.5 * W - 0.5);
  const T* c = cp + (long)deg * W;
#pragma unroll
  for (int i = 0; i < W; ++i) { kt[i] = c[i]; kp[i] = c[i]; }
  for (int k = deg - 1; k >= 0; --k) {
    c -= W;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      kt[i] = fma(kt[i], ut, c[i]);
      kp[i] = fma(kp[i], up, c[i]);
    }
  }
}

// The workgroup after the last tile: points binned past the tiles (footprint outside the formed rows, or
// not finite; only set_pts_async lets them through) get NaN.
template <typename T>
__device__ __forceinline__ void interp_mark_bad(const InterpArgs<T>& a) {
  const double nan = __builtin_nan("");
  for (long p = a.tile_start[a.ntiles] + threadIdx.x; p < a.npts; p += blockDim.x) {
    const int o = a.idx[p];
    a.out_re[o] = nan;
    if (a.spin) a.out_im[o] = nan;
  }
}

// Points per tile that the interpolation orders by bank class in LDS (tiles with more keep the binned order).
constexpr int kInterpSortCap = 1024;

// One workgroup per fine-grid tile: the tile plus its (W-1)-cell halo is staged in LDS once and
// every point binned to the tile is interpolated from it (fully unrolled W x W footprint).
// A 16-byte LDS read serves 16 lanes per pass, and lane k's reads fall in bank group (base_k + offset) mod 16,
// base_k = lr * tw + lc the footprint's first cell, offset the same for all lanes; consecutive points of a ring
// are about 1.25 cells apart, so 16 of them span about 20 columns and collide.  The workgroup therefore counts
// its points per class base mod 16 (LDS atomics), lists them class by class, and gives lane position k the
// list entry (k mod 16) S + k / 16, S = ceil(n / 16): the 16 lanes of a pass take entries S apart, from
// (nearly always) different classes.
template <typename T, int W>
__global__ __launch_bounds__(256) void k_interp_w(InterpArgs<T> a) {
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  T2* tile = reinterpret_cast<T2*>(smem_raw);
  const int tb = blockIdx.x;
  if (tb == a.ntiles) {
    interp_mark_bad(a);
    return;
  }
  const int s0 = a.tile_start[tb], s1 = a.tile_start[tb + 1];
  if (s0 == s1) return;
  const int tw = a.tc + W, th = a.tr + W;
  int* ccnt = reinterpret_cast<int*>(smem_raw + sizeof(T2) * (size_t)tw * th);  // [16] class counts, then starts
  int* order = ccnt + 16;                                                          // [kInterpSortCap]
  const int n = s1 - s0;
  const bool sorted = n <= kInterpSortCap;
  if (sorted && threadIdx.x < 16) ccnt[threadIdx.x] = 0;
  const int tile_r = tb / a.ntc, tile_c = tb - (tb / a.ntc) * a.ntc;
  const int r0 = tile_r * a.tr, c0 = tile_c * a.tc;
  for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
    const int rr = i / tw, cc = i - rr * tw;
    const int grow = r0 + rr;
    int gc = c0 + cc;
    while (gc >= a.nf) gc -= a.nf;
    tile[i] = grow < a.nrow ? a.grid[(long)grow * a.nf + gc] : T2{T(0), T(0)};
  }
  __syncthreads();
  const double scale = a.nf / (2.0 * M_PI);
  const double hw = 0.5 * W;
  // first cell of a point's footprint relative to the tile
  auto first_cell = [&](double2 tp, double& ut, double& up, int& i0, int& j0u, int& lr, int& lc) {
    ut = tp.x * scale;
    up = tp.y * scale;
    up -= a.nf * floor(up / a.nf);
    i0 = (int)ceil(ut - hw);
    j0u = (int)ceil(up - hw);
    const int j0 = j0u < 0 ? j0u + a.nf : j0u;
    lr = i0 - a.row_lo - r0;
    lc = j0 - c0;
  };
  if (sorted) {
    constexpr int EI = kInterpSortCap / 256;
    int cls[EI], rank[EI];
#pragma unroll
    for (int e = 0; e < EI; ++e) {
      const int i = threadIdx.x + e * 256;
      cls[e] = rank[e] = 0;
      if (i < n) {
        double ut, up;
        int i0, j0u, lr, lc;
        first_cell(a.pts[s0 + i], ut, up, i0, j0u, lr, lc);
        cls[e] = (lr * tw + lc) & 15;
        rank[e] = atomicAdd(&ccnt[cls[e]], 1);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int c = 0; c < 16; ++c) {
        const int v = ccnt[c];
        ccnt[c] = acc;
        acc += v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EI; ++e) {
      const int i = threadIdx.x + e * 256;
      if (i < n) order[ccnt[cls[e]] + rank[e]] = i;
    }
    __syncthreads();
  }
  const int S = (n + 15) >> 4;
  const int nk = sorted ? 16 * S : n;
  for (int k = threadIdx.x; k < nk; k += blockDim.x) {
    int pi = k;
    if (sorted) {
      const int q = (k & 15) * S + (k >> 4);
      if (q >= n) continue;
      pi = order[q];
    }
    const int p = s0 + pi;
    double ut, up;
    int i0, j0u, lr, lc;
    first_cell(a.pts[p], ut, up, i0, j0u, lr, lc);
    const int o = a.idx[p];
    const bool ok = lr >= 0 && lr < a.tr && lc >= 0 && lc < a.tc;  // always true for binned points
    T kt[W], kp[W];
    ker_weights<T, W>(i0 - ut, j0u - up, a.beta, a.kpoly, a.deg, kt, kp);
    T sr = 0, si = 0;
    if (ok) {
      const T2* base = tile + lr * tw + lc;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const T2* row = base + i * tw;
        T rr = 0, ri = 0;
#pragma unroll
        for (int j = 0; j < W; ++j) {
          const T2 g = row[j];
          rr = fma(g.x, kp[j], rr);
          ri = fma(g.y, kp[j], ri);
        }
        sr = fma(rr, kt[i], sr);
        si = fma(ri, kt[i], si);
      }
    }
    double vr = ok ? (double)sr : __builtin_nan("");
    double vi = ok ? (double)si : __builtin_nan("");
    if (a.spin == 0) {
      a.out_re[o] = vr;
    } else {
      if (a.gamma) {
        double sn, cs;
        sincos(a.spin * a.gamma[o], &sn, &cs);
        const double xr = vr * cs - vi * sn, xi = vr * sn + vi * cs;
        vr = xr;
        vi = xi;
      }
      a.out_re[o] = vr;
      a.out_im[o] = vi;
    }
  }
}

// Same for any width (runtime w, exact kernel evaluation); used above the templated widths.
template <typename T>
__global__ __launch_bounds__(256) void k_interp(InterpArgs<T> a) {
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  T2* tile = reinterpret_cast<T2*>(smem_raw);
  const int tb = blockIdx.x;
  if (tb == a.ntiles) {
    interp_mark_bad(a);
    return;
  }
  const int s0 = a.tile_start[tb], s1 = a.tile_start[tb + 1];
  if (s0 == s1) return;
  const int tw = a.tc + a.w, th = a.tr + a.w;
  const int tile_r = tb / a.ntc, tile_c = tb - (tb / a.ntc) * a.ntc;
  const int r0 = tile_r * a.tr, c0 = tile_c * a.tc;
  for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
    const int rr = i / tw, cc = i - rr * tw;
    const int grow = r0 + rr;
    int gc = c0 + cc;
    while (gc >= a.nf) gc -= a.nf;
    tile[i] = grow < a.nrow ? a.grid[(long)grow * a.nf + gc] : T2{T(0), T(0)};
  }
  __syncthreads();
  const double scale = a.nf / (2.0 * M_PI);
  const double hw = 0.5 * a.w;
  const T inv = (T)(2.0 / a.w);
  for (int p = s0 + threadIdx.x; p < s1; p += blockDim.x) {
    const double2 tp = a.pts[p];
    const double ut = tp.x * scale;
    double up = tp.y * scale;
    up -= a.nf * floor(up / a.nf);
    const int i0 = (int)ceil(ut - hw);
    const int j0u = (int)ceil(up - hw);
    const int j0 = j0u < 0 ? j0u + a.nf : j0u;
    const int lr = i0 - a.row_lo - r0, lc = j0 - c0;
    const int o = a.idx[p];
    const bool ok = lr >= 0 && lr < a.tr && lc >= 0 && lc < a.tc;  // always true for binned points
    T kt[kMaxW], kp[kMaxW];
    const double ft = i0 - ut, fp = j0u - up;
#pragma unroll
    for (int i = 0; i < kMaxW; ++i) {
      if (i < a.w) {
        kt[i] = WorkT<T>::ker((T)(ft + i) * inv, a.beta);
        kp[i] = WorkT<T>::ker((T)(fp + i) * inv, a.beta);
      }
    }
    T sr = 0, si = 0;
    if (ok) {
      for (int i = 0; i < a.w; ++i) {
        const T2* row = tile + (lr + i) * tw + lc;
        T rr = 0, ri = 0;
#pragma unroll 4
        for (int j = 0; j < a.w; ++j) {
          const T2 g = row[j];
          rr = fma(g.x, kp[j], rr);
          ri = fma(g.y, kp[j], ri);
        }
        sr = fma(rr, kt[i], sr);
        si = fma(ri, kt[i], si);
      }
    }
    double vr = ok ? (double)sr : __builtin_nan("");
    double vi = ok ? (double)si : __builtin_nan("");
    if (a.spin == 0) {
      a.out_re[o] = vr;
    } else {
      if (a.gamma) {
        double s, c;
        sincos(a.spin * a.gamma[o], &s, &c);
        const double xr = vr * c - vi * s, xi = vr * s + vi * c;
        vr = xr;
        vi = xi;
      }
      a.out_re[o] = vr;
      a.out_im[o] = vi;
    }
  }
}

// ------------------------------------------------------------------ spreading (type 1)
template <typename T>
struct SpreadArgs {
  int nf, w, row_lo, nrow, spin, tr, tc, ntc;
  T beta;
  const T* kpoly;
  int deg;
  const double2* pts;
  const int* idx;
  const int* tile_start;
  const double* in_re;
  const double* in_im;
  const double* gamma;
  c2_t<T>* grid;
};

// Broadcast of lane I of each 16-lane row to the whole row (DPP row_newbcast).
template <int I>
__device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + I, 0xF, 0xF, false));
}
template <int I>
__device__ __forceinline__ double row_bcast(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + I, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + I, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Exact transpose of k_interp_w: points spread into an LDS copy of the tile (+ halo) with LDS
// atomics, then the tile is added to the fine grid with global atomics.  Each point is spread by one
// 16-lane row of a wave, lane j owning footprint column j (W <= 16): the lane evaluates the kernel
// polynomial of cell j along both axes, takes the row weights of the other lanes by DPP broadcasts and
// adds W row cells of its column.  The per-point set-up (cell, rotation by gamma, polynomial arguments)
// runs one point per lane and reaches the spreading rows by lane shuffles.  The W lanes of a point hit
// W adjacent cells of one tile row per instruction, so they never share an address; with one lane per
// point (64 neighbouring points, 1.25 cells apart, per instruction) bank and address conflicts took
// about half of the LDS cycles (at lmax 4095, w = 12: 2.4e9 of 4.7e9; 1.3e9 of 3.6e9 now).  The LDS
// tile accumulates in FP64 for both grid precisions: ds_add_f64 sustains several times the rate of
// ds_add_f32 on these patterns, and the sums lose nothing before the final rounding to the grid's type.
template <typename T, int W>
__global__ __launch_bounds__(256) void k_spread_w(SpreadArgs<T> a) {
  static_assert(W <= 16, "one 16-lane row per point");
  extern __shared__ unsigned char smem_raw[];
  double* tile = reinterpret_cast<double*>(smem_raw);  // interleaved re, im
  const int tb = blockIdx.x;
  const int s0 = a.tile_start[tb], s1 = a.tile_start[tb + 1];
  if (s0 == s1) return;
  const int tw = a.tc + W, th = a.tr + W;
  const int ncell = tw * th;
  for (int i = threadIdx.x; i < 2 * ncell; i += blockDim.x) tile[i] = 0.0;
  __syncthreads();
  const int tile_r = tb / a.ntc, tile_c = tb % a.ntc;
  const int r0 = tile_r * a.tr, c0 = tile_c * a.tc;
  const double scale = a.nf / (2.0 * M_PI);
  const double hw = 0.5 * W;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row16 = lane >> 4;  // this lane's 16-lane row of the wave
  const int j = lane & 15;
  const bool own = j < W;
  const int jc = own ? j : W - 1;  // lanes past the footprint evaluate the last cell and add nothing
  // Each wave takes 64 points at a time: lane-per-point set-up (position, rotation), then 16 rounds in
  // which row r spreads point 4 t + r, its set-up fetched from that lane.
  for (int pb = s0 + 64 * wave; pb < s1; pb += 64 * (blockDim.x >> 6)) {
    const int p = pb + lane;
    int cell = -1;  // lr * tw + lc, or -1: no point
    double vr = 0.0, vi = 0.0;
    T ut_ = T(0), up_ = T(0);
    if (p < s1) {
      const double2 tp = a.pts[p];
      const double ut = tp.x * scale;
      double up = tp.y * scale;
      up -= a.nf * floor(up / a.nf);
      const int i0 = (int)ceil(ut - hw);
      const int j0u = (int)ceil(up - hw);
      const int j0 = j0u < 0 ? j0u + a.nf : j0u;
      const int lr = i0 - a.row_lo - r0, lc = j0 - c0;
      if (lr >= 0 && lr < a.tr && lc >= 0 && lc < a.tc) {  // always, for points binned by set_pts
        cell = lr * tw + lc;
        const int o = a.idx[p];
        vr = a.in_re[o];
        vi = (a.spin && a.in_im) ? a.in_im[o] : 0.0;
        if (a.spin && a.gamma) {
          double sn, cs;
          sincos(-a.spin * a.gamma[o], &sn, &cs);
          const double xr = vr * cs - vi * sn, xi = vr * sn + vi * cs;
          vr = xr; vi = xi;
        }
        // polynomial argument u = f + W/2 - 1/2, f = first cell - position, along theta and phi
        ut_ = (T)((i0 - ut) + 0.5 * W - 0.5);
        up_ = (T)((j0u - up) + 0.5 * W - 0.5);
      }
    }
    for (int t = 0; t < 16; ++t) {
      const int src = 4 * t + row16;
      // every lane takes part in the shuffles (a shuffle reads nothing from an inactive lane)
      const int cl = __shfl(cell, src, 64);
      const double pr = __shfl(vr, src, 64), pi = __shfl(vi, src, 64);
      const T ua = __shfl(ut_, src, 64), ub = __shfl(up_, src, 64);
      if (cl < 0) continue;  // uniform over the 16-lane row
      const T* c = a.kpoly + (long)a.deg * W + jc;
      T kt = *c, kp = kt;
      for (int k = a.deg - 1; k >= 0; --k) {
        c -= W;
        kt = fma(kt, ua, *c);
        kp = fma(kp, ub, *c);
      }
      const double cr = pr * (double)kp, ci = pi * (double)kp;
      double* col = tile + 2 * (cl + jc);
      static_for<0, W>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const double kti = (double)row_bcast<i>(kt);
        if (own) {
          atomicAdd(&col[2 * i * tw], cr * kti);
          atomicAdd(&col[2 * i * tw + 1], ci * kti);
        }
      });
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ncell; i += blockDim.x) {
    const int rr = i / tw, cc = i - rr * tw;
    const int grow = r0 + rr;
    if (grow >= a.nrow) continue;
    const T vr = (T)tile[2 * i], vi = (T)tile[2 * i + 1];
    if (vr == T(0) && vi == T(0)) continue;
    const int gc = (c0 + cc) % a.nf;  // a tile can be wider than a small grid
    T* g = reinterpret_cast<T*>(a.grid + (long)grow * a.nf + gc);
    atomicAdd(&g[0], vr);
    atomicAdd(&g[1], vi);
  }
}

// Widths with a fully unrolled kernel; wider kernels use the runtime-width versions.
constexpr int kMaxUnrolledW = 16;

// ------------------------------------------------------------------ spreading by register row windows
// Widths kRowsMinW..kRowsMaxW.  A tile is tr rows x 16 SW columns of first cells (i0, j0), SW = 17 - W:
// 16-lane group g owns the 16 fine columns [g SW, g SW + 16) of the tile (+ halo), which hold the whole
// W-column footprint of every point whose first column lies in the group's slice [g SW, g SW + SW).  The
// points of the tile are counting-sorted in LDS by (slice, first row); each group then walks the rows
// r = 0, 1, ... of its slice, adding every point of row r into W rows of register accumulators (lane j:
// column g SW + j; the kernel row weights written by the group into an LDS slot and read back by all its
// lanes, the column weight shuffled in by the point's column offset).  Row r receives nothing after the
// points of row r, so its accumulator is added to the LDS accumulation rows once (16 lanes, one ds_add_f64
// pair) and reused as row r + W: a rotating window whose slot index r mod W is wave-uniform (the four groups
// of a wave take slices w, w + 4, w + 8, w + 12 and step through the rows together), so every register index
// is a compile-time constant of the W-fold unrolled row loop.  LDS atomics: two per fine cell and covering
// group instead of 2 W^2 per point.  Tiles with more points than the LDS record buffer holds are spread in
// several passes.  A workgroup walks several tiles down one column strip and keeps the accumulation rows in a
// ring, so the W - 1 halo rows below a tile carry over to the next instead of going to HBM with atomics.
constexpr int kRowsMinW = 8;
constexpr int kRowsMaxW = 16;
constexpr int kRowsGroups = 16;  // 16-lane groups per 256-thread workgroup (= slices per tile)
constexpr int kRowsPts = 3;      // points per thread in one pass of the row-window spreading (records <= 768)
constexpr int kRowsMinDeg = 3;   // kernel polynomial degrees with a register-resident Horner
constexpr int kRowsMaxDeg = 24;

template <int D, int DM, typename T>
__device__ __forceinline__ void horner2(const T (&cf)[DM + 1], T ut, T up, T& kt, T& kp) {
  kt = cf[D];
  kp = cf[D];
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    kt = fma(kt, ut, cf[k]);
    kp = fma(kp, up, cf[k]);
  }
}

// kt, kp = the polynomial of this lane's cell at ut, up; deg (<= DM) is uniform
template <int DM, typename T>
__device__ __forceinline__ void horner_any(const T (&cf)[DM + 1], int deg, T ut, T up, T& kt, T& kp) {
  switch (deg) {
    case 3: if constexpr (3 <= DM) horner2<3, DM>(cf, ut, up, kt, kp); break;
    case 4: if constexpr (4 <= DM) horner2<4, DM>(cf, ut, up, kt, kp); break;
    case 5: if constexpr (5 <= DM) horner2<5, DM>(cf, ut, up, kt, kp); break;
    case 6: if constexpr (6 <= DM) horner2<6, DM>(cf, ut, up, kt, kp); break;
    case 7: if constexpr (7 <= DM) horner2<7, DM>(cf, ut, up, kt, kp); break;
    case 8: if constexpr (8 <= DM) horner2<8, DM>(cf, ut, up, kt, kp); break;
    case 9: if constexpr (9 <= DM) horner2<9, DM>(cf, ut, up, kt, kp); break;
    case 10: if constexpr (10 <= DM) horner2<10, DM>(cf, ut, up, kt, kp); break;
    case 11: if constexpr (11 <= DM) horner2<11, DM>(cf, ut, up, kt, kp); break;
    case 12: if constexpr (12 <= DM) horner2<12, DM>(cf, ut, up, kt, kp); break;
    case 13: if constexpr (13 <= DM) horner2<13, DM>(cf, ut, up, kt, kp); break;
    case 14: if constexpr (14 <= DM) horner2<14, DM>(cf, ut, up, kt, kp); break;
    case 15: if constexpr (15 <= DM) horner2<15, DM>(cf, ut, up, kt, kp); break;
    case 16: if constexpr (16 <= DM) horner2<16, DM>(cf, ut, up, kt, kp); break;
    case 17: if constexpr (17 <= DM) horner2<17, DM>(cf, ut, up, kt, kp); break;
    case 18: if constexpr (18 <= DM) horner2<18, DM>(cf, ut, up, kt, kp); break;
    case 19: if constexpr (19 <= DM) horner2<19, DM>(cf, ut, up, kt, kp); break;
    case 20: if constexpr (20 <= DM) horner2<20, DM>(cf, ut, up, kt, kp); break;
    case 21: if constexpr (21 <= DM) horner2<21, DM>(cf, ut, up, kt, kp); break;
    case 22: if constexpr (22 <= DM) horner2<22, DM>(cf, ut, up, kt, kp); break;
    case 23: if constexpr (23 <= DM) horner2<23, DM>(cf, ut, up, kt, kp); break;
    default: horner2<DM, DM>(cf, ut, up, kt, kp); break;
  }
}

struct RowsLayout {
  int tr, tc, aw, cap, deg, nkey;
  int seg, ntr;  // tiles one workgroup walks down its column strip; tile rows of the grid
  size_t acc_off, val_off, uu_off, off_off, cnt_off, ktb_off, bytes;
};

template <typename T>
__host__ __device__ inline RowsLayout rows_layout(int w, int tr, int deg, int cap) {
  RowsLayout L;
  const int sw = 17 - w;
  L.tr = tr;
  L.tc = kRowsGroups * sw;
  L.aw = L.tc + w - 1;
  L.cap = cap;
  L.deg = deg;
  L.nkey = kRowsGroups * tr;
  L.seg = 1;
  L.ntr = 0;
  size_t o = 0;
  L.acc_off = o; o += sizeof(double2) * (size_t)(tr + w - 1) * L.aw;
  L.val_off = o; o += sizeof(double2) * (size_t)cap;
  L.uu_off = o;  o += sizeof(c2_t<T>) * (size_t)cap;
  L.off_off = o; o += (size_t)cap;
  o = (o + 15) & ~(size_t)15;
  L.cnt_off = o; o += sizeof(int) * (size_t)(2 * L.nkey + 1);
  o = (o + 15) & ~(size_t)15;
  L.ktb_off = o; o += sizeof(T) * 16 * kRowsGroups;
  L.bytes = o;
  return L;
}

template <typename T, int W, int DM>
__global__ __launch_bounds__(256) void k_spread_rows(SpreadArgs<T> a, RowsLayout lay) {
  static_assert(W >= kRowsMinW && W <= kRowsMaxW, "register row windows for 8 <= W <= 16");
  constexpr int SW = 17 - W;
  using T2 = c2_t<T>;
  extern __shared__ unsigned char smem_raw[];
  double* acc = reinterpret_cast<double*>(smem_raw + lay.acc_off);  // interleaved re, im; (tr + W - 1) x aw
  double2* val = reinterpret_cast<double2*>(smem_raw + lay.val_off);  // rotated value per point record
  T2* uu = reinterpret_cast<T2*>(smem_raw + lay.uu_off);               // its row and column polynomial arguments
  unsigned char* offs = smem_raw + lay.off_off;                        // its first column within the slice
  int* cnt = reinterpret_cast<int*>(smem_raw + lay.cnt_off);  // [nkey]: counts, then cursors
  int* start = cnt + lay.nkey;                                 // [nkey + 1]
  T* ktbuf = reinterpret_cast<T*>(smem_raw + lay.ktb_off);     // [16 groups][16] row weights of a point
  // one workgroup walks tiles t0 .. t1 - 1 of column strip tile_c; rows a tile has completed leave the ring
  // buffer acc (R = tr + W - 1 rows, slot = row - t0 tr mod R) while the W - 1 rows below carry over
  const int tile_c = blockIdx.x % a.ntc;
  const int t0 = (blockIdx.x / a.ntc) * lay.seg, t1 = min(t0 + lay.seg, lay.ntr);
  if (a.tile_start[t0 * a.ntc + tile_c] == a.tile_start[(t1 - 1) * a.ntc + tile_c + 1]) return;  // no points
  const int tr = lay.tr, aw = lay.aw, nkey = lay.nkey, deg = lay.deg;
  const int R = tr + W - 1;
  for (int i = threadIdx.x; i < 2 * R * aw; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int c0 = tile_c * a.tc;
  // cells no other workgroup touches: columns W - 1 .. tc - 1 of the strip (when there are two strips or
  // more), rows from W - 1 below the segment's first row (all rows of the first segment) up to the next
  // segment's first row
  const int excl_c1 = a.ntc >= 2 ? min(a.tc, a.nf - c0) : 0;
  const int excl_r0 = t0 == 0 ? 0 : t0 * tr + W - 1, excl_r1 = t1 * tr;
  const double scale = a.nf / (2.0 * M_PI);
  const double hw = 0.5 * W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, gsub = lane >> 4;
  const int slice = wave + 4 * gsub;  // the four groups of a wave own non-overlapping windows
  const int gbase = threadIdx.x & ~15;  // first thread of this lane's group
  T* ktb = ktbuf + 16 * (threadIdx.x >> 4);
  // the polynomial of footprint cell l16 (zero for l16 >= W and above the degree), in registers
  T cf[DM + 1];
#pragma unroll
  for (int k = 0; k <= DM; ++k) cf[k] = (l16 < W && k <= deg) ? a.kpoly[(long)k * W + l16] : T(0);
  // first cells of a binned point, relative to the tile; lr / lc outside the tile never happen
  auto cells = [&](int p, int r0, int& lr, int& lc, double& ut, double& up, int& i0, int& j0u) {
    const double2 tp = a.pts[p];
    ut = tp.x * scale;
    up = tp.y * scale;
    up -= a.nf * floor(up / a.nf);
    i0 = (int)ceil(ut - hw);
    j0u = (int)ceil(up - hw);
    const int j0 = j0u < 0 ? j0u + a.nf : j0u;
    lr = i0 - a.row_lo - r0;
    lc = j0 - c0;
  };
  // ring rows [rb, re) of tile-row origin r0 (slot base) onto the fine grid, then cleared
  auto flush_rows = [&](int r0, int base, int rb, int re) {
    for (int i = threadIdx.x; i < (re - rb) * aw; i += blockDim.x) {
      const int rr = rb + i / aw, cc = i % aw;
      int slot = base + rr;  // base < R, rr < R
      slot -= slot >= R ? R : 0;
      double* cell = acc + 2 * (slot * aw + cc);
      const double2 v = *reinterpret_cast<const double2*>(cell);
      *reinterpret_cast<double2*>(cell) = double2{0.0, 0.0};
      const int grow = r0 + rr;
      if (grow >= a.nrow) continue;
      const T vr = (T)v.x, vi = (T)v.y;
      if (vr == T(0) && vi == T(0)) continue;
      const int gc = (c0 + cc) % a.nf;  // a strip can be wider than a small grid
      T* g = reinterpret_cast<T*>(a.grid + (long)grow * a.nf + gc);
      if (grow >= excl_r0 && grow < excl_r1 && cc >= W - 1 && cc < excl_c1) {
        *reinterpret_cast<c2_t<T>*>(g) = c2_t<T>{vr, vi};
      } else {
        atomicAdd(&g[0], vr);
        atomicAdd(&g[1], vi);
      }
    }
  };
  for (int t = t0; t < t1; ++t) {
  const int tb = t * a.ntc + tile_c;
  const int s0 = a.tile_start[tb], s1 = a.tile_start[tb + 1];
  const int r0 = t * tr;
  const int base = ((t - t0) * tr) % R;  // ring slot of the tile's first row
  for (int cb = s0; cb < s1; cb += lay.cap) {
    const int n = min(lay.cap, s1 - cb);
    for (int i = threadIdx.x; i < nkey; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    // pass 1: each thread holds up to kRowsPts points of the pass: its key (slice, row), polynomial arguments,
    // column offset and (rotated) value, whose gathers are in flight across the count and the scan
    int keyv[kRowsPts], offv[kRowsPts];
    T2 uuv[kRowsPts];
    double vrv[kRowsPts], viv[kRowsPts];
#pragma unroll
    for (int k = 0; k < kRowsPts; ++k) {
      const int q = threadIdx.x + k * 256;
      keyv[k] = -1;
      offv[k] = 0;
      uuv[k] = T2{T(0), T(0)};
      vrv[k] = viv[k] = 0.0;
      if (q < n) {
        const int p = cb + q;
        int lr, lc, i0, j0u;
        double ut, up;
        cells(p, r0, lr, lc, ut, up, i0, j0u);
        if (lr >= 0 && lr < tr && lc >= 0 && lc < lay.tc) {
          keyv[k] = (lc / SW) * tr + lr;
          offv[k] = lc % SW;
          uuv[k] = T2{(T)((i0 - ut) + 0.5 * W - 0.5), (T)((j0u - up) + 0.5 * W - 0.5)};
          const int o = a.idx[p];
          vrv[k] = a.in_re[o];
          viv[k] = (a.spin && a.in_im) ? a.in_im[o] : 0.0;
          if (a.spin && a.gamma) {
            double sn, cs;
            sincos(-a.spin * a.gamma[o], &sn, &cs);
            const double xr = vrv[k] * cs - viv[k] * sn, xi = vrv[k] * sn + viv[k] * cs;
            vrv[k] = xr;
            viv[k] = xi;
          }
          atomicAdd(&cnt[keyv[k]], 1);
        }
      }
    }
    __syncthreads();
    // exclusive scan (one wave, nkey <= 16 x 32)
    if (wave == 0) {
      const int per = (nkey + 63) / 64;
      int sum = 0;
      for (int i = 0; i < per; ++i) {
        const int k = lane * per + i;
        sum += k < nkey ? cnt[k] : 0;
      }
      int incl = sum;
      for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
      }
      int run = incl - sum;
      for (int i = 0; i < per; ++i) {
        const int k = lane * per + i;
        if (k < nkey) {
          start[k] = run;
          run += cnt[k];
        }
      }
      if (lane == 63) start[nkey] = incl;
    }
    __syncthreads();
    // pass 2: records in sorted order (cnt counts down as the cursor)
#pragma unroll
    for (int k = 0; k < kRowsPts; ++k) {
      if (keyv[k] < 0) continue;
      const int pos = start[keyv[k]] + atomicSub(&cnt[keyv[k]], 1) - 1;
      val[pos] = double2{vrv[k], viv[k]};
      uu[pos] = uuv[k];
      offs[pos] = (unsigned char)offv[k];
    }
    __syncthreads();
    // row walk: register window of W rows x (re, im), slot of row r = r mod W
    double ar[W], ai[W];
#pragma unroll
    for (int i = 0; i < W; ++i) ar[i] = ai[i] = 0.0;
    const int kbase = slice * tr;
    double* acol = acc + 2 * (slice * SW + l16);
    for (int rb = 0; rb < tr + W - 1; rb += W) {
      static_for<0, W>([&](auto kk) {
        constexpr int K = decltype(kk)::value;
        const int r = rb + K;  // wave-uniform
        if (r < tr) {
          const int b = start[kbase + r], e = start[kbase + r + 1];
          for (int q = b; __any(q < e); ++q) {
            if (q < e) {  // uniform over the 16-lane group
              const double2 v = val[q];
              const T2 u = uu[q];
              const int off = offs[q];
              T kt, kp;
              horner_any<DM>(cf, deg, u.x, u.y, kt, kp);
              // the row weights reach the group's lanes through its LDS slot (LDS operations of one wave
              // complete in order; the fences keep the compiler from moving the reads above the write)
              ktb[l16] = kt;
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
              T ktv[W];
#pragma unroll
              for (int i = 0; i < W; ++i) ktv[i] = ktb[i];
              const int src = l16 - off;  // footprint cell of this lane's column
              const T kpc_ = __shfl(kp, gbase + (src & 15), 64);
              const double kpc = (src >= 0 && src < W) ? (double)kpc_ : 0.0;
              const double cr = v.x * kpc, ci = v.y * kpc;
              static_for<0, W>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                constexpr int slot = (K + i) % W;
                const double kti = (double)ktv[i];
                ar[slot] = fma(cr, kti, ar[slot]);
                ai[slot] = fma(ci, kti, ai[slot]);
              });
            }
          }
        }
        if (r < tr + W - 1) {  // row r is complete: add it to the ring and reuse the slot for row r + W
          if (ar[K] != 0.0 || ai[K] != 0.0) {
            int slot = base + r;
            slot -= slot >= R ? R : 0;
            atomicAdd(&acol[2 * slot * aw], ar[K]);
            atomicAdd(&acol[2 * slot * aw + 1], ai[K]);
          }
          ar[K] = 0.0;
          ai[K] = 0.0;
        }
      });
    }
    __syncthreads();  // records are rewritten by the next pass
  }
  // rows r0 .. r0 + tr - 1 get nothing from later tiles of the strip
  flush_rows(r0, base, 0, tr);
  __syncthreads();
  }
  // the W - 1 rows below the last tile are shared with the next segment
  flush_rows((t1 - 1) * tr, ((t1 - 1 - t0) * tr) % R, tr, R);
}

template <typename T>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  double* tile = reinterpret_cast<double*>(smem_raw);  // interleaved re, im, FP64 as in k_spread_w
  const int tb = blockIdx.x;
  const int s0 = a.tile_start[tb], s1 = a.tile_start[tb + 1];
  if (s0 == s1) return;
  const int tw = a.tc + a.w, th = a.tr + a.w;
  const int ncell = tw * th;
  for (int i = threadIdx.x; i < 2 * ncell; i += blockDim.x) tile[i] = 0.0;
  __syncthreads();
  const int tile_r = tb / a.ntc, tile_c = tb % a.ntc;
  const int r0 = tile_r * a.tr, c0 = tile_c * a.tc;
  const double scale = a.nf / (2.0 * M_PI);
  const double hw = 0.5 * a.w;
  const T inv = (T)(2.0 / a.w);
  for (int p = s0 + threadIdx.x; p < s1; p += blockDim.x) {
    const double2 tp = a.pts[p];
    const double ut = tp.x * scale;
    double up = tp.y * scale;
    up -= a.nf * floor(up / a.nf);
    const int i0 = (int)ceil(ut - hw);
    const int j0u = (int)ceil(up - hw);
    const int j0 = j0u < 0 ? j0u + a.nf : j0u;
    const int lr = i0 - a.row_lo - r0, lc = j0 - c0;
    if (lr < 0 || lr >= a.tr || lc < 0 || lc >= a.tc) continue;  // never for points binned by set_pts
    const int o = a.idx[p];
    double vr = a.in_re[o], vi = (a.spin && a.in_im) ? a.in_im[o] : 0.0;
    if (a.spin && a.gamma) {
      double s, c;
      sincos(-a.spin * a.gamma[o], &s, &c);
      double xr = vr * c - vi * s, xi = vr * s + vi * c;
      vr = xr; vi = xi;
    }
    T kp[kMaxW];
    const double fp = j0u - up;
    const double ft = i0 - ut;
#pragma unroll
    for (int j = 0; j < kMaxW; ++j)
      if (j < a.w) kp[j] = WorkT<T>::ker((T)(fp + j) * inv, a.beta);
    for (int i = 0; i < a.w; ++i) {
      const double kt = (double)WorkT<T>::ker((T)(ft + i) * inv, a.beta);
      const double cr = vr * kt, ci = vi * kt;
      double* trow = tile + 2 * ((lr + i) * tw + lc);
      for (int j = 0; j < a.w; ++j) {
        atomicAdd(&trow[2 * j], cr * (double)kp[j]);
        atomicAdd(&trow[2 * j + 1], ci * (double)kp[j]);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ncell; i += blockDim.x) {
    const int rr = i / tw, cc = i - rr * tw;
    const int grow = r0 + rr;
    if (grow >= a.nrow) continue;
    const T vr = (T)tile[2 * i], vi = (T)tile[2 * i + 1];
    if (vr == T(0) && vi == T(0)) continue;
    const int gc = (c0 + cc) % a.nf;  // a tile can be wider than a small grid
    T* g = reinterpret_cast<T*>(a.grid + (long)grow * a.nf + gc);
    atomicAdd(&g[0], vr);
    atomicAdd(&g[1], vi);
  }
}

// ------------------------------------------------------------------ point binning
struct BinArgs {
  int nf, w, row_lo, nrow, tr, tc, ntc, ntiles;
  long npts;
  const double* theta;
  const double* phi;
  int* key;
  int* cnt;
};

// Wave-aggregated binning: lanes holding consecutive points with the same tile form a run;
// the run's first lane issues one atomic for the whole run, and the run keeps its input order
// (ring-ordered lensing points mostly arrive tile by tile, so this is close to one atomic per wave).
struct RunInfo {
  int leader;  // lane index of this lane's run leader
  int len;     // run length (valid on the leader)
};

__device__ __forceinline__ RunInfo wave_runs(int key, bool valid) {
  const int lane = threadIdx.x & 63;
  const int kprev = __shfl_up(key, 1, 64);
  const int vprev = __shfl_up((int)valid, 1, 64);
  const bool cont = valid && lane > 0 && vprev && kprev == key;
  const bool lead = valid && !cont;
  const unsigned long long cm = __ballot(cont);
  const unsigned long long lm = __ballot(lead);
  RunInfo r;
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const unsigned long long lmask = lm & upto;
  r.leader = lmask ? 63 - __clzll((long long)lmask) : lane;
  // length: 1 + number of consecutive continuation bits above the leader
  unsigned long long above = (lane == 63) ? 0ull : (cm >> (lane + 1));
  r.len = 1 + (int)__builtin_ctzll(~above);
  if (r.len > 64 - lane) r.len = 64 - lane;
  return r;
}

// Tile of a point, or a.ntiles (the bin after the last tile, which no interpolation or spreading tile reads)
// for a point whose footprint leaves the formed rows or that is not finite.  The range checks come before
// any float-to-int conversion, which is undefined for NaN and out-of-range values.
__device__ __forceinline__ int bin_key(const BinArgs& a, double theta, double phi, bool& bad) {
  const double scale = a.nf / (2.0 * M_PI);
  const double hw = 0.5 * a.w;
  const double ut = theta * scale;
  double up = phi * scale;
  const bool finite = ut > -4.0 * a.nf && ut < 4.0 * a.nf && fabs(phi) < 1e15;  // false for NaN
  if (!finite) {
    bad = true;
    return a.ntiles;
  }
  up -= a.nf * floor(up / a.nf);
  const int q0 = (int)ceil(ut - hw) - a.row_lo;
  int j0 = (int)ceil(up - hw);
  if (j0 < 0) j0 += a.nf;
  if (j0 >= a.nf) j0 -= a.nf;
  bad = q0 < 0 || q0 + a.w > a.nrow;
  return bad ? a.ntiles : (q0 / a.tr) * a.ntc + (j0 / a.tc);
}

__global__ __launch_bounds__(256) void k_bin_count(BinArgs a, int* bad_count) {
  const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool valid = j < a.npts;
  bool bad = false;
  int k = 0;
  if (valid) {
    k = bin_key(a, a.theta[j], a.phi[j], bad);
    a.key[j] = k;
  }
  const unsigned long long bm = __ballot(valid && bad);
  if ((threadIdx.x & 63) == 0 && bm) atomicAdd(bad_count, __popcll(bm));
  const RunInfo r = wave_runs(k, valid);
  if (valid && r.leader == (int)(threadIdx.x & 63)) atomicAdd(&a.cnt[k], r.len);
}

__global__ __launch_bounds__(256) void k_bin_scatter(long npts, const double* __restrict__ theta,
                                                     const double* __restrict__ phi, const int* __restrict__ key,
                                                     const int* __restrict__ start, int* cursor,
                                                     double2* __restrict__ pts, int* __restrict__ idx) {
  const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool valid = j < npts;
  const int k = valid ? key[j] : 0;
  const RunInfo r = wave_runs(k, valid);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (valid && r.leader == lane) base = start[k] + atomicAdd(&cursor[k], r.len);
  base = __shfl(base, r.leader, 64);
  if (!valid) return;
  const int pos = base + (lane - r.leader);
  pts[pos] = double2{theta[j], phi[j]};
  idx[pos] = (int)j;
}

template <typename T>
__global__ void k_fill_zero(T* p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = T(0);
}

// host Gauss-Legendre nodes on [-1, 1] (Newton on P_n), for the kernel's Fourier transform
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      pp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}

template <typename T, int NT, int NMAX, typename K, typename A>
void launch_pass(K kern, int nblocks, size_t lds, const A& args, hipStream_t st) {
  NUSHT_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(NT), lds, st, args);
  NUSHT_LAUNCH_CHECK();
}

}  // namespace

// ------------------------------------------------------------------ parameters
template <typename T, int W>
void launch_interp(int w, dim3 grid, size_t lds, hipStream_t st, const InterpArgs<T>& a) {
  if constexpr (W > kMaxUnrolledW) {
    hipLaunchKernelGGL(k_interp<T>, grid, dim3(256), lds, st, a);
  } else {
    if (w == W) hipLaunchKernelGGL((k_interp_w<T, W>), grid, dim3(256), lds, st, a);
    else launch_interp<T, W + 1>(w, grid, lds, st, a);
  }
}

template <typename T, int W>
void launch_spread(int w, dim3 grid, size_t lds, hipStream_t st, const SpreadArgs<T>& a) {
  if constexpr (W > kMaxUnrolledW) {
    hipLaunchKernelGGL(k_spread<T>, grid, dim3(256), lds, st, a);
  } else {
    if (w == W) hipLaunchKernelGGL((k_spread_w<T, W>), grid, dim3(256), lds, st, a);
    else launch_spread<T, W + 1>(w, grid, lds, st, a);
  }
}

template <typename T, int W>
void launch_spread_rows(int w, dim3 grid, hipStream_t st, const SpreadArgs<T>& a, const RowsLayout& lay) {
  if constexpr (W > kRowsMaxW) {
    throw Error{kErrUnsupported, "nufft: no register row-window spreading for this width"};
  } else {
    if (w == W) {
      // polynomial coefficients in registers: 13 for degrees up to 12 (eps >= 1e-8), 25 above
      const void* k = lay.deg <= 12 ? (const void*)k_spread_rows<T, W, 12> : (const void*)k_spread_rows<T, W, 24>;
      NUSHT_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lay.bytes));
      if (lay.deg <= 12) hipLaunchKernelGGL((k_spread_rows<T, W, 12>), grid, dim3(256), lay.bytes, st, a, lay);
      else hipLaunchKernelGGL((k_spread_rows<T, W, 24>), grid, dim3(256), lay.bytes, st, a, lay);
    } else {
      launch_spread_rows<T, W + 1>(w, grid, st, a, lay);
    }
  }
}

NufftParams NufftParams::choose(int lmax, int spin, double eps, double sigma_hint) {
  NUSHT_REQUIRE(lmax >= 1, kErrArg, "nufft: lmax must be >= 1");
  NUSHT_REQUIRE(eps > 0 && eps < 1, kErrArg, "nufft: epsilon must be in (0, 1)");
  NufftParams p;
  p.lmax = lmax;
  p.spin = spin;
  p.eps = eps;
  p.single = eps >= 1e-6;
  const int maxlen = kLdsBytes / (p.single ? 8 : 16);
  int ntd = fft_good_size(2 * lmax + 2);
  while (ntd % 2) ntd = fft_good_size(ntd + 1);
  NUSHT_REQUIRE(ntd <= maxlen, kErrUnsupported, "nufft: lmax too large for the in-LDS theta transform");
  p.ntd = ntd;
  p.ntheta = ntd / 2 + 1;
  p.nm = 2 * lmax + 1;
  double sigma = sigma_hint > 0 ? sigma_hint : (eps < 1e-8 ? 2.0 : 1.25);
  int nf = fft_good_size((int)std::ceil(sigma * p.nm));
  while (nf % 2) nf = fft_good_size(nf + 1);
  if (nf > maxlen) {
    nf = maxlen;
    while (nf % 2 || !fft_supported(nf)) --nf;
  }
  const double sg = (double)nf / p.nm;
  NUSHT_REQUIRE(sg >= 1.2, kErrUnsupported, "nufft: fine grid too small for this lmax");
  int w = (int)std::ceil(std::log(1.0 / eps) / (M_PI * std::sqrt(1.0 - 1.0 / sg)));
  w = std::max(w, 2);
  NUSHT_REQUIRE(w <= kMaxW, kErrUnsupported, "nufft: kernel width exceeds the supported maximum");
  p.w = w;
  p.beta = 0.97 * M_PI * w * (1.0 - 1.0 / (2.0 * sg));
  // Small lmax: widen the grid until the rows touched by theta in [0, pi] (about nf / 2 + w + 3) form a window
  // that does not wrap around the torus, nf >= 2 w + 8 (the binning and the tile loads assume it).  The kernel
  // keeps the width and shape of the unwidened grid: the wider grid only lowers the aliasing error, while a
  // kernel narrowed to it would raise the truncation error.
  while (nf < 2 * w + 8 && nf < maxlen) nf = fft_good_size(nf + 2);
  while (nf % 2) nf = fft_good_size(nf + 1);
  p.nf = nf;
  p.sigma = (double)nf / p.nm;
  // fine rows touched by points with theta in [0, pi] (one cell of slack on both sides)
  int lo = -(w / 2) - 1;
  int hi = (int)std::ceil(0.5 * nf - 0.5 * w) + w;
  NUSHT_REQUIRE(hi - lo + 1 < nf, kErrUnsupported, "nufft: fine grid too small for the kernel width");
  p.row_lo = lo;
  p.nrow = hi - lo + 1;
  return p;
}

// ------------------------------------------------------------------ plan
// Piecewise-polynomial ES kernel.  A point at fractional offset t in [0, 1) from its first cell
// needs psi(z_i), z_i = (t + i - w/2) 2/w, i = 0..w-1.  Every cell gets its own polynomial
// psi(z_i) = sum_k c[k][i] u^k, u = t - 1/2, interpolated at degree + 1 Chebyshev nodes.  The two
// outer cells hold the square-root branch point at |z| = 1, where psi is already below
// e^{-beta (1 - sqrt(1 - (1 - 2/w)^2))} ~ eps, so their fit error stays below the target as well.
// The degree is the lowest in [2, 24] whose max error on a dense sample (evaluated by Horner in
// the working precision) over all cells is below `target`.
static void fit_kernel_poly(int w, double beta, double target, bool single, int& deg, double& err,
                            std::vector<double>& coef) {
  const int ni = w;  // every cell of the footprint, the two outer ones included (see ker_weights)
  deg = 0; err = 0; coef.clear();
  if (ni <= 0) return;
  auto psi = [&](long double z) -> long double {
    long double t = 1.0L - z * z;
    return t > 0 ? std::exp((long double)beta * (std::sqrt(t) - 1.0L)) : 0.0L;
  };
  double best_err = 1e300;
  for (int d = 2; d <= 24; ++d) {
    const int n = d + 1;
    std::vector<double> c((size_t)n * ni);
    std::vector<long double> un(n);
    for (int j = 0; j < n; ++j) un[j] = 0.5L * std::cos((long double)M_PI * (j + 0.5L) / n);
    for (int i = 0; i < ni; ++i) {
      // Vandermonde solve in long double (Gaussian elimination, partial pivoting)
      std::vector<long double> A((size_t)n * (n + 1));
      for (int r = 0; r < n; ++r) {
        long double pw = 1.0L;
        for (int k = 0; k < n; ++k) { A[(size_t)r * (n + 1) + k] = pw; pw *= un[r]; }
        A[(size_t)r * (n + 1) + n] = psi((un[r] + 0.5L + i - 0.5L * w) * 2.0L / w);
      }
      for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
          if (std::fabs(A[(size_t)r * (n + 1) + col]) > std::fabs(A[(size_t)piv * (n + 1) + col])) piv = r;
        if (piv != col)
          for (int k = 0; k <= n; ++k) std::swap(A[(size_t)col * (n + 1) + k], A[(size_t)piv * (n + 1) + k]);
        for (int r = col + 1; r < n; ++r) {
          const long double f = A[(size_t)r * (n + 1) + col] / A[(size_t)col * (n + 1) + col];
          for (int k = col; k <= n; ++k) A[(size_t)r * (n + 1) + k] -= f * A[(size_t)col * (n + 1) + k];
        }
      }
      std::vector<long double> x(n);
      for (int r = n - 1; r >= 0; --r) {
        long double v = A[(size_t)r * (n + 1) + n];
        for (int k = r + 1; k < n; ++k) v -= A[(size_t)r * (n + 1) + k] * x[k];
        x[r] = v / A[(size_t)r * (n + 1) + r];
      }
      for (int k = 0; k < n; ++k) c[(size_t)k * ni + i] = (double)x[k];
    }
    double e = 0;
    for (int i = 0; i < ni; ++i)
      for (int q = 0; q <= 512; ++q) {
        const double u = -0.5 + q / 512.0;
        double v;
        if (single) {
          float fv = (float)c[(size_t)d * ni + i];
          for (int k = d - 1; k >= 0; --k) fv = std::fma(fv, (float)u, (float)c[(size_t)k * ni + i]);
          v = fv;
        } else {
          v = c[(size_t)d * ni + i];
          for (int k = d - 1; k >= 0; --k) v = std::fma(v, u, c[(size_t)k * ni + i]);
        }
        e = std::max(e, (double)std::fabs((long double)v - psi((u + 0.5L + i - 0.5L * w) * 2.0L / w)));
      }
    if (e < best_err) { best_err = e; deg = d; coef = c; }
    if (e <= target) break;
  }
  err = best_err;
}

NufftPlan::NufftPlan(const NufftParams& p, long npts_max, hipStream_t st, int tile_rows)
    : p_(p), npts_max_(npts_max) {
  // deconvolution table: 1 / int_{-w/2}^{w/2} psi(2t/w) cos(2 pi k t / nf) dt
  std::vector<double> xq, wq;
  gauss_legendre(4 * p.w + 60, xq, wq);
  std::vector<double> dec(p.lmax + 1);
  for (int k = 0; k <= p.lmax; ++k) {
    double s = 0;
    for (size_t i = 0; i < xq.size(); ++i) {
      double t = 0.5 * p.w * xq[i];
      double z = xq[i];
      double psi = std::exp(p.beta * (std::sqrt(std::max(0.0, 1.0 - z * z)) - 1.0));
      s += wq[i] * psi * std::cos(2.0 * M_PI * k * t / p.nf);
    }
    dec[k] = 1.0 / (0.5 * p.w * s);
  }
  upload(deconv_, dec);
  bool fitted = false;
  {
    std::vector<double> c;
    const double target = std::max(p.eps / 100.0, p.single ? 3e-7 : 1e-15);
    fit_kernel_poly(p.w, p.beta, target, p.single, p_.poly_deg, p_.poly_err, c);
    fitted = !c.empty();
    if (c.empty()) c.push_back(0.0);
    if (p.single) {
      std::vector<float> cf(c.begin(), c.end());
      upload(kpoly_, cf);
    } else {
      upload(kpoly_, c);
    }
  }
  // FP32 DFS passes take radix-16 stages (one LDS pass fewer on 8192 and 10240); in FP64 the 16 complex values
  // per thread spill at 1024 threads, so those keep radix 8
  if (p.single) { ftd_f_.init(p.ntd, true); fnf_f_.init(p.nf, true); }
  else { ftd_d_.init(p.ntd); fnf_d_.init(p.nf); }
  const size_t cb = p.single ? 8 : 16;
  T_.alloc(cb * (size_t)p.nrow * p.nm);
  grid_.alloc(cb * (size_t)p.nrow * p.nf);
  // tiles: tile_r_ rows (default 8) x 16 (17 - w) columns for the register row-window spreading (8 <= w <= 16; the
  // interpolation takes the same tiles), 32 x 32 otherwise.  NUSHT_SPREAD_ROWS=0 keeps the LDS-atomic
  // spreading and its tiles; NUSHT_SPREAD_TR sets the tile height (rows of first cells).
  rows_spread_ = fitted && p.w >= kRowsMinW && p.w <= kRowsMaxW && p_.poly_deg >= kRowsMinDeg &&
                 p_.poly_deg <= kRowsMaxDeg &&
                 !(getenv("NUSHT_SPREAD_ROWS") && getenv("NUSHT_SPREAD_ROWS")[0] == '0');
  if (rows_spread_) {
    int tr = tile_rows > 0 ? tile_rows : 8;  // the measured best at lmax 2047 and 4095 (eps 1e-6, 1e-7)
    if (const char* e = getenv("NUSHT_SPREAD_TR"); tile_rows <= 0 && e && e[0]) tr = atoi(e);
    tr = std::max(1, std::min(32, tr));
    tile_r_ = tr;
    tile_c_ = kRowsGroups * (17 - p.w);
    // LDS: the accumulation tile plus point records (value, row and column arguments, column offset): three
    // workgroups per CU when 256 records fit in 52 KiB, two when they fit in 78 KiB, else one with up to
    // 156 KiB; the records left over in the chosen budget are used too (fewer passes over a tile)
    const size_t rec = sizeof(double2) + (p.single ? sizeof(float2) : sizeof(double2)) + 1;
    const RowsLayout base = p.single ? rows_layout<float>(p.w, tr, p_.poly_deg, 0)
                                     : rows_layout<double>(p.w, tr, p_.poly_deg, 0);
    const size_t budget = base.bytes + 256 * rec <= 52 * 1024   ? 52 * 1024
                          : base.bytes + 256 * rec <= 78 * 1024 ? 78 * 1024
                                                                : 156 * 1024;
    NUSHT_REQUIRE(base.bytes + 64 * rec <= budget, kErrUnsupported, "nufft: spreading tile too large for LDS");
    long cap = (long)((budget - base.bytes) / rec);
    cap = std::max(64L, std::min(256L * kRowsPts, cap - cap % 64));
    rows_cap_ = (int)cap;
  }
  ntile_r_ = ceil_div(p.nrow, tile_r_);
  ntile_c_ = ceil_div(p.nf, tile_c_);
  const int ntiles = ntile_r_ * ntile_c_;
  // the row-window spreading walks up to 16 tiles down a column strip per workgroup, keeping about 3000
  // workgroups (NUSHT_SPREAD_SEG sets the count)
  rows_seg_ = std::max(1, std::min(16, ntiles / 3072));
  if (const char* e = getenv("NUSHT_SPREAD_SEG"); e && e[0]) rows_seg_ = std::max(1, std::min(ntile_r_, atoi(e)));
  npts_max_ = std::max<long>(npts_max, 1);
  NUSHT_REQUIRE(npts_max_ < (1L << 31), kErrUnsupported, "nufft: too many points");
  pts_.alloc(sizeof(double2) * npts_max_);
  idx_.alloc(sizeof(int) * npts_max_);
  key_.alloc(sizeof(int) * npts_max_);
  tile_start_.alloc(sizeof(int) * (ntiles + 1));
  tile_cnt_.alloc(sizeof(int) * (ntiles + 2));
  cursor_.alloc(sizeof(int) * (ntiles + 1));
  NUSHT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp_bytes_, tile_cnt_.as<int>(), tile_start_.as<int>(),
                                             ntiles + 1, st));
  scan_tmp_.alloc(std::max<size_t>(scan_tmp_bytes_, 16));
}

size_t NufftPlan::bytes() const {
  return deconv_.bytes() + kpoly_.bytes() + T_.bytes() + grid_.bytes() + pts_.bytes() + idx_.bytes() + key_.bytes() +
         tile_start_.bytes() + tile_cnt_.bytes() + cursor_.bytes() + scan_tmp_.bytes();
}

void NufftPlan::set_pts(const double* theta, const double* phi, long npts, hipStream_t st, bool check) {
  NUSHT_REQUIRE(npts >= 0 && npts <= npts_max_, kErrArg, "set_pts: npts exceeds the plan's maximum");
  npts_ = npts;
  StageScope sc("set_pts", st);
  const int ntiles = ntile_r_ * ntile_c_;
  // tile_cnt_[ntiles + 1] is an error counter
  hipLaunchKernelGGL(k_fill_zero<int>, dim3(ceil_div(ntiles + 2, 256)), dim3(256), 0, st, tile_cnt_.as<int>(),
                     (long)ntiles + 2);
  hipLaunchKernelGGL(k_fill_zero<int>, dim3(ceil_div(ntiles + 1, 256)), dim3(256), 0, st, cursor_.as<int>(),
                     (long)ntiles + 1);
  NUSHT_LAUNCH_CHECK();
  if (npts == 0) {
    hipLaunchKernelGGL(k_fill_zero<int>, dim3(ceil_div(ntiles + 1, 256)), dim3(256), 0, st, tile_start_.as<int>(),
                       (long)ntiles + 1);
    return;
  }
  BinArgs b{p_.nf, p_.w, p_.row_lo, p_.nrow, tile_r_, tile_c_, ntile_c_, ntiles, npts, theta, phi,
            key_.as<int>(), tile_cnt_.as<int>()};
  hipLaunchKernelGGL(k_bin_count, dim3(ceil_div(npts, 256)), dim3(256), 0, st, b, tile_cnt_.as<int>() + ntiles + 1);
  NUSHT_LAUNCH_CHECK();
  size_t tb = scan_tmp_bytes_;
  NUSHT_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp_.get(), tb, tile_cnt_.as<int>(), tile_start_.as<int>(),
                                             ntiles + 1, st));
  hipLaunchKernelGGL(k_bin_scatter, dim3(ceil_div(npts, 256)), dim3(256), 0, st, npts, theta, phi, key_.as<int>(),
                     tile_start_.as<int>(), cursor_.as<int>(), pts_.as<double2>(), idx_.as<int>());
  NUSHT_LAUNCH_CHECK();
  if (!check) return;
  int bad = 0;
  NUSHT_HIP(hipMemcpyAsync(&bad, tile_cnt_.as<int>() + ntiles + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  NUSHT_HIP(hipStreamSynchronize(st));
  if (bad != 0) {
    npts_ = 0;  // the plan holds no valid points until the next successful set_pts
    throw Error{kErrArg, "set_pts: " + std::to_string(bad) + " point(s) with theta outside [0, pi] or NaN"};
  }
}

template <typename T>
static FftDesc<T> desc_of(const FftLength<float>& f, const FftLength<double>& d);
template <>
FftDesc<float> desc_of<float>(const FftLength<float>& f, const FftLength<double>&) { return f.desc(); }
template <>
FftDesc<double> desc_of<double>(const FftLength<float>&, const FftLength<double>& d) { return d.desc(); }

template <typename T, template <typename, int, int> class KT>
struct Pass;

template <typename T>
static constexpr int nmax_big() { return kLdsBytes / (int)sizeof(c2_t<T>); }

#define NUSHT_LAUNCH_SIZED(KERN, T, n, nblocks, lds, args, st)                                        \
  do {                                                                                               \
    if ((n) <= 2048) launch_pass<T, 256, 2048>(KERN<T, 256, 2048>, nblocks, lds, args, st);           \
    else launch_pass<T, 1024, nmax_big<T>()>(KERN<T, 1024, nmax_big<T>()>, nblocks, lds, args, st);     \
  } while (0)

template <typename T>
void NufftPlan::type2_impl(const double2* G, long sm, const double* gamma, double* ore, double* oim,
                           hipStream_t st) {
  using T2 = c2_t<T>;
  const NufftParams& p = p_;
  const FftDesc<T> ftd = desc_of<T>(ftd_f_, ftd_d_);
  const FftDesc<T> fnf = desc_of<T>(fnf_f_, fnf_d_);
  const int nmrows = p.spin ? 2 * p.lmax + 1 : p.lmax + 1;
  ThetaArgs<T> ta{p.lmax, p.spin, p.ntheta, p.ntd, p.nf, p.nm, p.row_lo, p.nrow, nmrows, G, sm,
                  deconv_.as<double>(), ftd, fnf, T_.as<T2>()};
  const size_t lds = sizeof(T2) * (size_t)std::max(p.ntd, p.nf);
  {
    StageScope sc("theta_pass_t2", st);
    NUSHT_LAUNCH_SIZED(k_theta2, T, std::max(p.ntd, p.nf), (nmrows + 1) / 2, lds, ta, st);  // column pairs
  }
  PhiArgs<T> pa{p.lmax, p.nf, p.nm, p.nrow, fnf, T_.as<T2>(), grid_.as<T2>()};
  {
    StageScope sc("phi_pass_t2", st);
    NUSHT_LAUNCH_SIZED(k_phi2, T, p.nf, p.nrow, sizeof(T2) * (size_t)p.nf, pa, st);
  }
  if (npts_ == 0) return;
  const int ntiles = ntile_r_ * ntile_c_;
  InterpArgs<T> ia{p.nf, p.w, p.row_lo, p.nrow, p.spin, tile_r_, tile_c_, ntile_c_, ntiles, (T)p.beta,
                   kpoly_.as<T>(), p.poly_deg, npts_, pts_.as<double2>(), idx_.as<int>(), tile_start_.as<int>(),
                   grid_.as<T2>(), gamma, ore, oim};
  StageScope sc("nufft_interp", st);
  // tile, then (templated widths) the class counts and the point order
  const size_t tlds = sizeof(T2) * (size_t)(tile_r_ + p.w) * (tile_c_ + p.w) + sizeof(int) * (16 + kInterpSortCap);
  launch_interp<T, 2>(p.w, dim3(ntiles + 1), tlds, st, ia);  // + 1: the workgroup for out-of-range points
  NUSHT_LAUNCH_CHECK();
}

template <typename T>
void NufftPlan::type1_impl(const double* ire, const double* iim, const double* gamma, double2* G, long sm,
                           hipStream_t st) {
  using T2 = c2_t<T>;
  const NufftParams& p = p_;
  const FftDesc<T> ftd = desc_of<T>(ftd_f_, ftd_d_);
  const FftDesc<T> fnf = desc_of<T>(fnf_f_, fnf_d_);
  const long ngrid = 2L * p.nrow * p.nf;
  {
  StageScope sc0("grid_zero", st);
  hipLaunchKernelGGL(k_fill_zero<T>, dim3(std::min<long>(ceil_div(ngrid, 256), 65536)), dim3(256), 0, st,
                     grid_.as<T>(), ngrid);
  NUSHT_LAUNCH_CHECK();
  }
  if (npts_ > 0) {
    StageScope sc1("nufft_spread", st);
    SpreadArgs<T> sa{p.nf, p.w, p.row_lo, p.nrow, p.spin, tile_r_, tile_c_, ntile_c_, (T)p.beta, kpoly_.as<T>(),
                     p.poly_deg, pts_.as<double2>(), idx_.as<int>(), tile_start_.as<int>(), ire, iim, gamma,
                     grid_.as<T2>()};
    if (rows_spread_) {
      RowsLayout lay = rows_layout<T>(p.w, tile_r_, p.poly_deg, rows_cap_);
      lay.seg = rows_seg_;
      lay.ntr = ntile_r_;
      launch_spread_rows<T, kRowsMinW>(p.w, dim3(ntile_c_ * (int)ceil_div(ntile_r_, rows_seg_)), st, sa, lay);
    } else {
      const size_t lds = sizeof(double2) * (size_t)(tile_r_ + p.w) * (tile_c_ + p.w);  // FP64 accumulation tile
      launch_spread<T, 2>(p.w, dim3(ntile_r_ * ntile_c_), lds, st, sa);
    }
    NUSHT_LAUNCH_CHECK();
  }
  PhiArgs<T> pa{p.lmax, p.nf, p.nm, p.nrow, fnf, T_.as<T2>(), grid_.as<T2>()};
  {
    StageScope sc("phi_pass_t1", st);
    NUSHT_LAUNCH_SIZED(k_phi1, T, p.nf, p.nrow, sizeof(T2) * (size_t)p.nf, pa, st);
  }
  const int nmrows = p.spin ? 2 * p.lmax + 1 : p.lmax + 1;
  ThetaArgs<T> ta{p.lmax, p.spin, p.ntheta, p.ntd, p.nf, p.nm, p.row_lo, p.nrow, nmrows, G, sm,
                  deconv_.as<double>(), ftd, fnf, T_.as<T2>()};
  const size_t lds = sizeof(T2) * (size_t)std::max(p.ntd, p.nf);
  StageScope sc("theta_pass_t1", st);
  NUSHT_LAUNCH_SIZED(k_theta1, T, std::max(p.ntd, p.nf), (nmrows + 1) / 2, lds, ta, st);  // column pairs
}

void NufftPlan::type2(const double2* G, long sm, const double* gamma, double* ore, double* oim, hipStream_t st) {
  if (p_.single) type2_impl<float>(G, sm, gamma, ore, oim, st);
  else type2_impl<double>(G, sm, gamma, ore, oim, st);
}

void NufftPlan::type1(const double* ire, const double* iim, const double* gamma, double2* G, long sm,
                      hipStream_t st) {
  if (p_.single) type1_impl<float>(ire, iim, gamma, G, sm, st);
  else type1_impl<double>(ire, iim, gamma, G, sm, st);
}

// ------------------------------------------------------------------ plan-time autotuning
namespace {

// Uniform points on the sphere from a counter-based hash (splitmix64 finaliser).
__device__ __forceinline__ double hash_unit(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * 0x1p-53;
}

__global__ void k_random_points(long n, double* __restrict__ theta, double* __restrict__ phi) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    theta[i] = acos(1.0 - 2.0 * hash_unit(2 * (unsigned long long)i));
    phi[i] = 2.0 * M_PI * hash_unit(2 * (unsigned long long)i + 1);
  }
}

}  // namespace

std::unique_ptr<NufftPlan> nufft_autotune(int lmax, int spin, double eps, long npts_max, double2* G, long sm,
                                          hipStream_t st, NufftTune& tune) {
  int dev = 0;
  NUSHT_HIP(hipGetDevice(&dev));
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, double, long>, std::pair<NufftTune, NufftParams>> cache;
  const auto key = std::make_tuple(dev, lmax, spin, eps, npts_max);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      tune = it->second.first;
      return std::make_unique<NufftPlan>(it->second.second, npts_max, st, tune.tile_rows);
    }
  }
  const long n = std::max<long>(npts_max, 1);
  DevBuf pts(sizeof(double) * 4 * (size_t)n);  // theta, phi, values re, im
  double* th = pts.as<double>();
  double* ph = th + n;
  double* vr = ph + n;
  double* vi = vr + n;
  hipLaunchKernelGGL(k_random_points, dim3(std::min<long>(ceil_div(n, 256), 65536)), dim3(256), 0, st, n, th, ph);
  NUSHT_LAUNCH_CHECK();
  NUSHT_HIP(hipMemsetAsync(vr, 0, sizeof(double) * 2 * (size_t)n, st));
  const int nmrows = spin ? 2 * lmax + 1 : lmax + 1;
  const NufftParams def = NufftParams::choose(lmax, spin, eps, 0.0);
  NUSHT_HIP(hipMemsetAsync(G, 0, sizeof(double2) * (size_t)sm * nmrows, st));
  hipEvent_t e0, e1, e2;
  NUSHT_HIP(hipEventCreate(&e0));
  NUSHT_HIP(hipEventCreate(&e1));
  NUSHT_HIP(hipEventCreate(&e2));
  const char* env = getenv("NUSHT_VERBOSE");
  const bool verbose = env && env[0] && env[0] != '0';
  std::vector<int> tried_nf;
  float best = 1e30f;
  NufftTune bt;
  std::unique_ptr<NufftPlan> best_plan;
  for (double sg : {def.sigma, 1.25, 1.5, 2.0}) {
    NufftParams p;
    try {
      p = NufftParams::choose(lmax, spin, eps, sg);
    } catch (const Error&) {
      continue;
    }
    if (std::find(tried_nf.begin(), tried_nf.end(), p.nf) != tried_nf.end()) continue;
    tried_nf.push_back(p.nf);
    for (int tr : {16, 8, 32}) {
      auto plan = std::make_unique<NufftPlan>(p, n, st, tr);
      const bool rows = plan->rows_spread();  // otherwise the tile height is fixed: time one plan
      plan->set_pts(th, ph, n, st, false);
      plan->type2(G, sm, nullptr, vr, spin ? vi : nullptr, st);  // warm-up
      plan->type1(vr, spin ? vi : nullptr, nullptr, G, sm, st);
      NUSHT_HIP(hipEventRecord(e0, st));
      for (int k = 0; k < 2; ++k) plan->type2(G, sm, nullptr, vr, spin ? vi : nullptr, st);
      NUSHT_HIP(hipEventRecord(e1, st));
      for (int k = 0; k < 2; ++k) plan->type1(vr, spin ? vi : nullptr, nullptr, G, sm, st);
      NUSHT_HIP(hipEventRecord(e2, st));
      NUSHT_HIP(hipEventSynchronize(e2));
      float t2 = 0, t1 = 0;
      NUSHT_HIP(hipEventElapsedTime(&t2, e0, e1));
      NUSHT_HIP(hipEventElapsedTime(&t1, e1, e2));
      t2 /= 2;
      t1 /= 2;
      ++bt.candidates;
      if (verbose)
        fprintf(stderr, "[nusht] nufft autotune lmax %d spin %d eps %g: sigma %.4f (nf %d, w %d, degree %d) tile %d x %d "
                        "(%s spreading): type 2 %.3f ms, type 1 %.3f ms\n", lmax, spin, eps, p.sigma, p.nf, p.w,
                plan->params().poly_deg, plan->tile_rows(), plan->tile_cols(), rows ? "row-window" : "LDS-atomic", t2,
                t1);
      if (t2 + t1 < best) {
        best = t2 + t1;
        bt.sigma = p.sigma;
        bt.tile_rows = rows ? tr : 0;
        bt.type2_ms = t2;
        bt.type1_ms = t1;
        best_plan = std::move(plan);
      }
      if (!rows) break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  NUSHT_REQUIRE(best_plan, kErrUnsupported, "nufft autotune: no candidate fits");
  tune = bt;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = {bt, best_plan->params()};
  return best_plan;
}

}  // namespace nusht
