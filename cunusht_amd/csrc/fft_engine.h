// In-LDS mixed-radix Stockham FFT for one row per workgroup (gfx950).
//
// Each stage reads every element of the row from LDS into registers, waits on a barrier,
// and writes the butterfly outputs back to their Stockham positions, so a single LDS buffer
// of n complex values suffices (n = 10240 in FP64 is exactly the 160 KiB LDS of a CU).
// Stage (radix R, ns = product of previous radices), butterfly j in [0, n/R):
//     v[r] = x[j + r n/R] * w_n^{(j mod ns) r n/(ns R)}
//     v    = DFT_R(v)
//     y[(j / ns) ns R + (j mod ns) + r ns] = v[r]
// which yields the natural-order forward DFT  X[k] = sum_x x e^{-2 pi i k x / n}.
// Inverse transforms are forward transforms of conjugated data (callers conjugate on load and store).
#pragma once

#include "common.h"

namespace nusht {

constexpr int kFftMaxStages = 10;

// Kernel-argument view of a planned length.
template <typename T>
struct FftDesc {
  int n;
  int nstages;
  int radix[kFftMaxStages];
  const c2_t<T>* tw;  // tw[k] = e^{-2 pi i k / n}, k in [0, n)
};

// ---------------------------------------------------------------- radix kernels (forward)
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    T2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T>
struct Dft<T, 3> {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    const T c1 = T(-0.5);
    const T s1 = T(-0.86602540378443864676372317075293618);  // -sin(2pi/3)
    T2 a = v[0], b = v[1], c = v[2];
    T2 t = cadd(b, c), d = csub(b, c);
    v[0] = cadd(a, t);
    T2 m = T2{a.x + c1 * t.x, a.y + c1 * t.y};
    T2 r = T2{-s1 * d.y, s1 * d.x};  // i*s1*d
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <typename T>
struct Dft<T, 4> {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    T2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    T2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    T2 dmi = T2{d.y, -d.x};  // -i*d
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, dmi);
    v[3] = csub(b, dmi);
  }
};

template <typename T>
struct Dft<T, 5> {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    const T c1 = T(0.30901699437494742410229341718281906);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718281906);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938214);   // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907277);   // sin(4pi/5)
    T2 a = v[0];
    T2 p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
    T2 p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
    v[0] = cadd(a, cadd(p1, p2));
    T2 r1 = T2{a.x + c1 * p1.x + c2 * p2.x, a.y + c1 * p1.y + c2 * p2.y};
    T2 r2 = T2{a.x + c2 * p1.x + c1 * p2.x, a.y + c2 * p1.y + c1 * p2.y};
    // forward: X_k = ... - i (s ...) m
    T2 i1 = T2{s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y};
    T2 i2 = T2{s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y};
    v[1] = T2{r1.x + i1.y, r1.y - i1.x};
    v[4] = T2{r1.x - i1.y, r1.y + i1.x};
    v[2] = T2{r2.x + i2.y, r2.y - i2.x};
    v[3] = T2{r2.x - i2.y, r2.y + i2.x};
  }
};

template <typename T>
struct Dft<T, 8> {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    const T h = T(0.70710678118654752440084436210484904);
    // radix-2 over (v[k], v[k+4])
    T2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
    T2 a1 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
    T2 a2 = cadd(v[2], v[6]), a6 = csub(v[2], v[6]);
    T2 a3 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
    // twiddles w8^k on the odd half: a5*w8, a6*(-i), a7*w8^3
    T2 b5 = T2{h * (a5.x + a5.y), h * (a5.y - a5.x)};
    T2 b6 = T2{a6.y, -a6.x};
    T2 b7 = T2{h * (a7.y - a7.x), -h * (a7.x + a7.y)};
    T2 e[4] = {a0, a1, a2, a3};
    T2 o[4] = {a4, b5, b6, b7};
    Dft<T, 4>::run(e);
    Dft<T, 4>::run(o);
    v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
    v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
  }
};

template <typename T>
struct Dft<T, 16> {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    // 4 x 4: inner DFT4 over stride-4 groups, twiddle w16^{q k}, outer DFT4
    T2 g[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int p = 0; p < 4; ++p) g[q][p] = v[q + 4 * p];
      Dft<T, 4>::run(g[q]);
    }
    const T c1 = T(0.92387953251128675612818318939678829), s1 = T(0.38268343236508977172845998403039887);
    const T h = T(0.70710678118654752440084436210484904);
    // w16^k for k = 1..9 (forward, e^{-2 pi i k/16})
    const T2 w1 = T2{c1, -s1}, w2 = T2{h, -h}, w3 = T2{s1, -c1};
    const T2 w4 = T2{T(0), T(-1)}, w6 = T2{-h, -h}, w9 = T2{-c1, s1};
    g[1][1] = cmul(g[1][1], w1); g[1][2] = cmul(g[1][2], w2); g[1][3] = cmul(g[1][3], w3);
    g[2][1] = cmul(g[2][1], w2); g[2][2] = cmul(g[2][2], w4); g[2][3] = cmul(g[2][3], w6);
    g[3][1] = cmul(g[3][1], w3); g[3][2] = cmul(g[3][2], w6); g[3][3] = cmul(g[3][3], w9);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T2 t[4] = {g[0][k], g[1][k], g[2][k], g[3][k]};
      Dft<T, 4>::run(t);
#pragma unroll
      for (int q = 0; q < 4; ++q) v[k + 4 * q] = t[q];
    }
  }
};

// Odd prime radix (7, 11, 13), symmetric form:
//   X_k = v0 + sum_{j=1}^{(R-1)/2} [ (v_j + v_{R-j}) cos(2 pi jk/R) - i (v_j - v_{R-j}) sin(2 pi jk/R) ]
// with cos/sin of 2 pi j / R from constant tables (index (j k) mod R).
template <int R> struct OddTab;
template <> struct OddTab<7> {
  static constexpr double c[6] = {6.23489801858733594386e-01, -2.22520933956314337365e-01, -9.00968867902419034976e-01, -9.00968867902419145999e-01, -2.22520933956314587165e-01, 6.23489801858733372342e-01};
  static constexpr double s[6] = {7.81831482468029803634e-01, 9.74927912181823619342e-01, 4.33883739117558231424e-01, -4.33883739117558009379e-01, -9.74927912181823619342e-01, -7.81831482468029914656e-01};
};
template <> struct OddTab<11> {
  static constexpr double c[10] = {8.41253532831181205509e-01, 4.15415013001886435084e-01, -1.42314838273285004799e-01, -6.54860733945284989588e-01, -9.59492973614497368651e-01, -9.59492973614497479673e-01, -6.54860733945285211632e-01, -1.42314838273285226844e-01, 4.15415013001886046506e-01, 8.41253532831181205509e-01};
  static constexpr double s[10] = {5.40640817455597555430e-01, 9.09631995354518330110e-01, 9.89821441880932795243e-01, 7.55749574354258268905e-01, 2.81732556841429671035e-01, -2.81732556841429393479e-01, -7.55749574354258157882e-01, -9.89821441880932684221e-01, -9.09631995354518552155e-01, -5.40640817455597444408e-01};
};
template <> struct OddTab<13> {
  static constexpr double c[12] = {8.85456025653209910509e-01, 5.68064746731155922888e-01, 1.20536680255323006006e-01, -3.54604887042535454889e-01, -7.48510748171101192305e-01, -9.70941817426052011797e-01, -9.70941817426052122819e-01, -7.48510748171101303328e-01, -3.54604887042535898978e-01, 1.20536680255323200295e-01, 5.68064746731154812664e-01, 8.85456025653210021531e-01};
  static constexpr double s[12] = {4.64723172043768506523e-01, 8.22983865893656352242e-01, 9.92708874098053972723e-01, 9.35016242685414833424e-01, 6.63122658240795193052e-01, 2.39315664287557683387e-01, -2.39315664287557433587e-01, -6.63122658240794971007e-01, -9.35016242685414722402e-01, -9.92708874098053972723e-01, -8.22983865893657018376e-01, -4.64723172043768395501e-01};
};

template <typename T, int R>
struct DftGeneric {
  using T2 = c2_t<T>;
  __device__ static void run(T2* v) {
    constexpr int H = (R - 1) / 2;
    T2 p[H], q[H];
    const T2 v0 = v[0];
    T2 sum = v0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      p[j - 1] = cadd(v[j], v[R - j]);
      q[j - 1] = csub(v[j], v[R - j]);
      sum = cadd(sum, p[j - 1]);
    }
    v[0] = sum;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      T2 A = v0, B{T(0), T(0)};
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const int idx = (j * k) % R;  // in [1, R-1]
        const T c = (T)OddTab<R>::c[idx - 1], s = (T)OddTab<R>::s[idx - 1];
        A = T2{A.x + c * p[j - 1].x, A.y + c * p[j - 1].y};
        B = T2{B.x + s * q[j - 1].x, B.y + s * q[j - 1].y};
      }
      // X_k = A - i B,  X_{R-k} = A + i B
      v[k] = T2{A.x + B.y, A.y - B.x};
      v[R - k] = T2{A.x - B.y, A.y + B.x};
    }
  }
};
template <typename T> struct Dft<T, 7> : DftGeneric<T, 7> {};
template <typename T> struct Dft<T, 11> : DftGeneric<T, 11> {};
template <typename T> struct Dft<T, 13> : DftGeneric<T, 13> {};

// ---------------------------------------------------------------- one stage
// NT threads; every thread handles up to NB butterflies held in registers.
// LDS bank spread for the stages in between: within each aligned block of 16 values, value i sits at
// i ^ ((i >> 4) & 7).  A 16-byte value spans 4 of the 64 banks, so 16 lanes are conflict-free when their
// values differ in the low 4 index bits; a stage's strided writes (stride R values when ns = 1, runs of ns
// values R ns apart later) otherwise pile onto a few bank groups (8-way for R = 8).  Among the XOR maps of
// bits 4..7 onto bits 0..3 this one gives the fewest conflicts over the stage patterns of the radix-8/4/5/3/2
// lengths used here (mean 1.13 lanes per bank group and 16-lane pass, worst 2; 2.0 and 8 without).  The
// first stage reads and the last stage writes the natural order, which are contiguous.  Used when n is a
// multiple of 16, so the permutation stays inside [0, n).
template <bool ON>
__device__ __forceinline__ int fft_swz(int i) {
  if constexpr (ON) return i ^ ((i >> 4) & 7);
  return i;
}

template <typename T, int R, int NT, int NMAX, bool SWZ_IN = false, bool SWZ_OUT = false>
__device__ __forceinline__ void fft_stage(c2_t<T>* buf, int n, int ns, const c2_t<T>* __restrict__ tw) {
  using T2 = c2_t<T>;
  constexpr int NB = (NMAX + R * NT - 1) / (R * NT);
  const int nb = n / R;
  const int stride = n / R;
  const int twstep = n / (ns * R);
  T2 v[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    int j = threadIdx.x + i * NT;
    if (j < nb) {
      int k = j % ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[fft_swz<SWZ_IN>(j + r * stride)];
      if (ns > 1) {
        // one table read per butterfly; w^r by repeated multiplication (error ~ R ulp)
        const T2 w1 = tw[k * twstep];
        T2 wr = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[i][r] = cmul(v[i][r], wr);
          if (r + 1 < R) wr = cmul(wr, w1);
        }
      }
      Dft<T, R>::run(v[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    int j = threadIdx.x + i * NT;
    if (j < nb) {
      int k = j % ns;
      int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[fft_swz<SWZ_OUT>(base + r * ns)] = v[i][r];
    }
  }
  __syncthreads();
}

// Forward FFT of buf[0..n) in LDS.  Caller must __syncthreads() after filling buf.
// GENERIC enables the odd radices 7 and 13 (needed only for externally imposed ring lengths);
// lengths chosen internally are 2-3-5 smooth and use radices 8, 4, 2, 3, 5.
// R16 enables radix-16 stages (lengths planned with FftLength::init(n, true)); they cost 64 VGPRs of values per
// butterfly in FP64, so only kernels with the register room (the 1024-thread DFS passes) turn them on.
template <typename T, int NT, int NMAX, bool GENERIC, bool SI, bool SO, bool R16 = false>
__device__ __forceinline__ void fft_stage_r(c2_t<T>* buf, const FftDesc<T>& d, int R, int ns) {
  if constexpr (R16) {
    if (R == 16) {
      fft_stage<T, 16, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw);
      return;
    }
  }
  switch (R) {
    case 2: fft_stage<T, 2, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw); break;
    case 3: fft_stage<T, 3, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw); break;
    case 4: fft_stage<T, 4, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw); break;
    case 5: fft_stage<T, 5, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw); break;
    case 8: fft_stage<T, 8, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw); break;
    default:
      if constexpr (GENERIC) {
        if (R == 7) fft_stage<T, 7, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw);
        else if (R == 13) fft_stage<T, 13, NT, NMAX, SI, SO>(buf, d.n, ns, d.tw);
      }
      break;
  }
}

// SWZ: lengths that are multiples of 16 keep the stages in between in the bank-spread order (fft_swz);
// it costs registers, so kernels that are register-bound leave it off.
template <typename T, int NT, int NMAX, bool GENERIC = false, bool SWZ = false, bool R16 = false>
__device__ void fft_lds(c2_t<T>* buf, const FftDesc<T>& d) {
  int ns = 1;
  if constexpr (SWZ) {
    if ((d.n & 15) == 0 && d.nstages > 1) {
      const int last = d.nstages - 1;
      fft_stage_r<T, NT, NMAX, GENERIC, false, true, R16>(buf, d, d.radix[0], ns);
      ns *= d.radix[0];
      for (int s = 1; s < last; ++s) {
        fft_stage_r<T, NT, NMAX, GENERIC, true, true, R16>(buf, d, d.radix[s], ns);
        ns *= d.radix[s];
      }
      fft_stage_r<T, NT, NMAX, GENERIC, true, false, R16>(buf, d, d.radix[last], ns);
      return;
    }
  }
  for (int s = 0; s < d.nstages; ++s) {
    fft_stage_r<T, NT, NMAX, GENERIC, false, false, R16>(buf, d, d.radix[s], ns);
    ns *= d.radix[s];
  }
}

template <typename T>
class FftLength {
 public:
  FftLength() = default;
  explicit FftLength(int n) { init(n); }
  void init(int n, bool radix16 = false);  // radix16: plan 16-point stages first (fft_lds<..., R16 = true> only)
  bool valid() const { return n_ > 0; }
  int n() const { return n_; }
  size_t bytes() const { return tw_.bytes(); }
  FftDesc<T> desc() const {
    FftDesc<T> d{};
    d.n = n_;
    d.nstages = (int)radices_.size();
    for (int i = 0; i < d.nstages; ++i) d.radix[i] = radices_[i];
    d.tw = tw_.as<c2_t<T>>();
    return d;
  }

 private:
  int n_ = 0;
  std::vector<int> radices_;
  DevBuf tw_;
};

// Factorization into supported radices (16 when radix16, then 8, 4, then 2, 3, 5, 7, 13).
// Returns an empty vector if the length has other prime factors.
std::vector<int> fft_factorize(int n, bool radix16 = false);
bool fft_supported(int n);
bool fft_needs_generic(int n);  // a radix 7, 11 or 13 stage is required
// Smallest 2-3-5 smooth length >= n (for choosing internal grid sizes).
int fft_good_size(int n);

}  // namespace nusht
