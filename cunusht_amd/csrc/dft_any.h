// DFT of any length up to kDftNMAX in LDS, one workgroup of kDftNT threads per transform.
//
// Each distinct length gets a description: a direct mixed-radix FFT when n factors into 2, 3, 5,
// 7, 13; Rader's algorithm when n is prime with a smooth n - 1; Bluestein's chirp-z algorithm when
// its padded length fits the LDS buffer; otherwise a split n = r b into r sub-transforms of length b
// (direct or Bluestein), X_k = sum_{j1} w_n^{j1 k} Y_{j1}[k mod b], with the ring values held in
// registers meanwhile.  Used by the ring SHT (rings of any length) and the 2D FFT (sides that are
// not smooth, e.g. the 2 ntheta - 2 rows of a doubled equiangular grid).
#pragma once

#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "common.h"
#include "fft_engine.h"

namespace nusht {

enum class DftMethod : int { Direct = 0, Rader = 1, Bluestein = 2, Split = 3 };
constexpr int kDftMethods = 4;
// 768 threads: a radix-13 stage over 8190 points (Rader for n = 8191) needs one butterfly per thread,
// and 768 rather than 640 threads halves its register spill.
constexpr int kDftNT = 768;
constexpr int kDftNMAX = 8320;
constexpr int kDftE = (kDftNMAX + kDftNT - 1) / kDftNT;  // values per thread held in registers

// Kernel-argument view of one length.
template <typename T>
struct DftLen {
  int n;              // transform length
  int method;         // DftMethod
  int mlen;           // FFT length run in LDS: n (direct), n - 1 (Rader), padded (Bluestein, split sub)
  int r, b;           // split: n = r * b, r sub-transforms of length b
  int sub_bluestein;  // split: sub-transforms by Bluestein (else a direct FFT of length b)
  FftDesc<T> fd;      // of length mlen
  const c2_t<T>* chirp;  // Bluestein: e^{-i pi j^2 / len}, len = n (or b for split)
  const c2_t<T>* kern;   // Bluestein: FFT of the wrapped conjugate chirp (mlen); Rader: FFT of the permuted roots
  const int* perm_in;    // Rader: g^q mod n
  const int* perm_out;   // Rader: g^-q mod n
  __host__ __device__ int lds_values() const { return n > mlen ? n : mlen; }
};

// Forward DFT of buf[0..n) by Bluestein's algorithm on the padded length L.mlen:
// X_k = c_k sum_j (x_j c_j) conj(c)_{k-j},  c_j = e^{-i pi j^2 / n}.
template <typename T>
__device__ __forceinline__ void bluestein_lds(c2_t<T>* buf, int n, const DftLen<T>& L) {
  using T2 = c2_t<T>;
  constexpr int NT = kDftNT;
  const int M = L.mlen;
  for (int j = threadIdx.x; j < M; j += NT) buf[j] = j < n ? cmul(buf[j], L.chirp[j]) : T2{T(0), T(0)};
  __syncthreads();
  fft_lds<T, NT, kDftNMAX, true>(buf, L.fd);
  for (int j = threadIdx.x; j < M; j += NT) buf[j] = cconj(cmul(buf[j], L.kern[j]));
  __syncthreads();
  fft_lds<T, NT, kDftNMAX, true>(buf, L.fd);
  const T inv = T(1) / T(M);
  for (int k = threadIdx.x; k < n; k += NT) buf[k] = cscale(cmul(L.chirp[k], cconj(buf[k])), inv);
  __syncthreads();
}

// In-place forward DFT of buf[0..L.n) in LDS (buf holds L.lds_values() entries).  All kDftNT threads call it.
template <int METHOD, typename T>
__device__ __forceinline__ void dft_lds(c2_t<T>* buf, const DftLen<T>& L) {
  using T2 = c2_t<T>;
  const int n = L.n;
  constexpr int NT = kDftNT;
  constexpr int E = kDftE;
  if constexpr (METHOD == (int)DftMethod::Direct) {
    fft_lds<T, NT, kDftNMAX, true>(buf, L.fd);
    return;
  }
  if constexpr (METHOD == (int)DftMethod::Bluestein) {
    bluestein_lds<T>(buf, n, L);
    return;
  }
  if constexpr (METHOD == (int)DftMethod::Split) {
    // n = r b, j = j1 + r j2:  X_k = sum_{j1} w_n^{j1 k} Y_{j1}[k mod b],  Y_{j1} = DFT_b(x_{j1 + r j2})
    const int r = L.r, b = L.b;
    T2 x[E], acc[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int q = threadIdx.x + i * NT;
      x[i] = q < n ? buf[q] : T2{T(0), T(0)};
      acc[i] = T2{T(0), T(0)};
    }
    __syncthreads();
    for (int j1 = 0; j1 < r; ++j1) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int q = threadIdx.x + i * NT;
        if (q < n && q % r == j1) buf[q / r] = x[i];
      }
      __syncthreads();
      if (L.sub_bluestein) {
        bluestein_lds<T>(buf, b, L);
      } else {
        fft_lds<T, NT, kDftNMAX, true>(buf, L.fd);
      }
#pragma unroll
      for (int i = 0; i < E; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < n) {
          const long e = (long)j1 * k % n;
          double sn, cs;
          sincospi(-2.0 * (double)e / (double)n, &sn, &cs);
          acc[i] = cadd(acc[i], cmul(T2{(T)cs, (T)sn}, buf[k % b]));
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int q = threadIdx.x + i * NT;
      if (q < n) buf[q] = acc[i];
    }
    __syncthreads();
    return;
  }
  if constexpr (METHOD != (int)DftMethod::Rader) return;
  // Rader (n prime): X[g^-p] = x0 + (a (*) b)_p with a_q = x[g^q], b_q = w^{g^-q}
  // The perm_in and kern reads are unconditional, from positions clamped to n1 - 1, so that each thread's E
  // reads of a table are in flight together (perm_out stays a guarded read: indices held in registers across
  // the last step spill).
  const int n1 = n - 1;
  const T2 x0 = buf[0];
  T2 v[E];
  {
    int pin[E];
#pragma unroll
    for (int i = 0; i < E; ++i) pin[i] = L.perm_in[min((int)threadIdx.x + i * NT, n1 - 1)];
#pragma unroll
    for (int i = 0; i < E; ++i) v[i] = buf[pin[i]];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int q = threadIdx.x + i * NT;
    if (q < n1) buf[q] = v[i];
  }
  __syncthreads();
  fft_lds<T, NT, kDftNMAX, true>(buf, L.fd);
  const T2 a0 = buf[0];
#pragma unroll
  for (int i = 0; i < E; ++i) v[i] = L.kern[min((int)threadIdx.x + i * NT, n1 - 1)];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int q = threadIdx.x + i * NT;
    if (q < n1) buf[q] = cconj(cmul(buf[q], v[i]));
  }
  __syncthreads();
  fft_lds<T, NT, kDftNMAX, true>(buf, L.fd);
  const T inv = T(1) / T(n1);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int p = threadIdx.x + i * NT;
    if (p < n1) v[i] = cadd(x0, cscale(cconj(buf[p]), inv));
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < E; ++i) {
    int p = threadIdx.x + i * NT;
    if (p < n1) buf[L.perm_out[p]] = v[i];
  }
  if (threadIdx.x == 0) buf[0] = cadd(x0, a0);
  __syncthreads();
}

// Host side: the descriptions of a set of lengths and their tables (device memory owned here).
template <typename T>
class DftSet {
 public:
  // index of length n (built on first request; throws kErrUnsupported for a length with no in-LDS DFT)
  int add(int n, hipStream_t st);
  const DftLen<T>& at(int i) const { return tabs_[i]->d; }
  int size() const { return (int)tabs_.size(); }
  size_t bytes() const;

 private:
  struct Tables {
    DftLen<T> d{};
    DevBuf chirp, kern, perm_in, perm_out;
  };
  const FftLength<T>& fft_of(int len);
  std::map<int, std::unique_ptr<FftLength<T>>> ffts_;  // by FFT length
  std::vector<std::unique_ptr<Tables>> tabs_;
  std::map<int, int> index_;  // transform length -> index
};

}  // namespace nusht
