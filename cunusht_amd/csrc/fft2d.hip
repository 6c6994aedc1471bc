// Batched in-place 2D complex FFT (the reference's _C2C / _iC2C through cuFFT,
// cunusht/deflection/GPU_nufft_transformer.py:271-287), on the in-LDS Stockham engine.
//
// Row pass: one workgroup per row of n1 values.  Column pass: one workgroup per strip of C adjacent
// columns, loaded row by row (C contiguous values per row, so the loads coalesce) into C column
// buffers in LDS and transformed one after the other by the whole workgroup.  Smooth sides (factors
// 2, 3, 5, 7, 13) use the Stockham engine directly; any other side up to 8320 uses the any-length DFT
// (Rader, Bluestein or a split into sub-transforms), as the reference's doubled grids need
// (2 ntheta - 2 is never smooth there).  The backward transform is conj(FFT(conj(x))), unnormalised,
// as cuFFT's inverse.  Tables are made once per length and kept for the life of the process.
#include "fft2d.h"

#include <map>
#include <memory>
#include <mutex>

#include "dft_any.h"
#include "fft_engine.h"
#include "timing.h"

namespace nusht {
namespace {

constexpr int kFft2dNT = 512;
template <typename T> struct Fft2dCap;                          // complex values that fit in the 160 KiB LDS
template <> struct Fft2dCap<double> { static constexpr int n = 10240; };
template <> struct Fft2dCap<float> { static constexpr int n = 20480; };

template <typename T>
__global__ __launch_bounds__(kFft2dNT) void k_fft_rows(c2_t<T>* data, int n0, int n1, FftDesc<T> fd, int backward) {
  extern __shared__ unsigned char smem_raw[];
  using T2 = c2_t<T>;
  T2* buf = reinterpret_cast<T2*>(smem_raw);
  T2* row = data + ((long)blockIdx.y * n0 + blockIdx.x) * n1;
  for (int k = threadIdx.x; k < n1; k += kFft2dNT) {
    const T2 v = row[k];
    buf[k] = backward ? cconj(v) : v;
  }
  __syncthreads();
  fft_lds<T, kFft2dNT, Fft2dCap<T>::n, true, true>(buf, fd);
  for (int k = threadIdx.x; k < n1; k += kFft2dNT) {
    const T2 v = buf[k];
    row[k] = backward ? cconj(v) : v;
  }
}

template <typename T>
__global__ __launch_bounds__(kFft2dNT) void k_fft_cols(c2_t<T>* data, int n0, int n1, int C, FftDesc<T> fd,
                                                       int backward) {
  extern __shared__ unsigned char smem_raw[];
  using T2 = c2_t<T>;
  T2* buf = reinterpret_cast<T2*>(smem_raw);  // C columns of n0 values
  T2* a = data + (long)blockIdx.y * n0 * n1;
  const int c0 = blockIdx.x * C;
  const int nc = min(C, n1 - c0);
  for (int i = threadIdx.x; i < n0 * nc; i += kFft2dNT) {
    const int r = i / nc, c = i - r * nc;
    const T2 v = a[(long)r * n1 + c0 + c];
    buf[c * n0 + r] = backward ? cconj(v) : v;
  }
  __syncthreads();
  for (int c = 0; c < nc; ++c) fft_lds<T, kFft2dNT, Fft2dCap<T>::n, true, true>(buf + c * n0, fd);
  for (int i = threadIdx.x; i < n0 * nc; i += kFft2dNT) {
    const int r = i / nc, c = i - r * nc;
    const T2 v = buf[c * n0 + r];
    a[(long)r * n1 + c0 + c] = backward ? cconj(v) : v;
  }
}

// Any-length variants: one kDftNT-thread workgroup per row / column strip, the length's description in the
// kernel arguments.
template <typename T, int METHOD>
__global__ __launch_bounds__(kDftNT) void k_dft_rows(c2_t<T>* data, int n0, int n1, DftLen<T> L, int backward) {
  extern __shared__ unsigned char smem_raw[];
  using T2 = c2_t<T>;
  T2* buf = reinterpret_cast<T2*>(smem_raw);
  T2* row = data + ((long)blockIdx.y * n0 + blockIdx.x) * n1;
  for (int k = threadIdx.x; k < n1; k += kDftNT) {
    const T2 v = row[k];
    buf[k] = backward ? cconj(v) : v;
  }
  __syncthreads();
  dft_lds<METHOD, T>(buf, L);
  for (int k = threadIdx.x; k < n1; k += kDftNT) {
    const T2 v = buf[k];
    row[k] = backward ? cconj(v) : v;
  }
}

template <typename T, int METHOD>
__global__ __launch_bounds__(kDftNT) void k_dft_cols(c2_t<T>* data, int n0, int n1, int C, DftLen<T> L,
                                                     int backward) {
  extern __shared__ unsigned char smem_raw[];
  using T2 = c2_t<T>;
  T2* buf = reinterpret_cast<T2*>(smem_raw);  // C column buffers of L.lds_values() entries
  const int ld = L.lds_values();
  T2* a = data + (long)blockIdx.y * n0 * n1;
  const int c0 = blockIdx.x * C;
  const int nc = min(C, n1 - c0);
  for (int i = threadIdx.x; i < n0 * nc; i += kDftNT) {
    const int r = i / nc, c = i - r * nc;
    const T2 v = a[(long)r * n1 + c0 + c];
    buf[c * ld + r] = backward ? cconj(v) : v;
  }
  __syncthreads();
  for (int c = 0; c < nc; ++c) dft_lds<METHOD, T>(buf + c * ld, L);
  for (int i = threadIdx.x; i < n0 * nc; i += kDftNT) {
    const int r = i / nc, c = i - r * nc;
    const T2 v = buf[c * ld + r];
    a[(long)r * n1 + c0 + c] = backward ? cconj(v) : v;
  }
}

template <typename T>
DftLen<T> any_length(int n, hipStream_t st) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<DftSet<T>>> sets;  // per device
  int dev = 0;
  NUSHT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto& set = sets[dev];
  if (!set) set = std::make_unique<DftSet<T>>();
  return set->at(set->add(n, st));
}

template <typename T, typename... A>
void launch_any(bool rows, int method, dim3 grid, size_t lds, hipStream_t st, A... args) {
  const void* f = nullptr;
  switch (method) {
    case 0: f = rows ? (const void*)k_dft_rows<T, 0> : (const void*)k_dft_cols<T, 0>; break;
    case 1: f = rows ? (const void*)k_dft_rows<T, 1> : (const void*)k_dft_cols<T, 1>; break;
    case 2: f = rows ? (const void*)k_dft_rows<T, 2> : (const void*)k_dft_cols<T, 2>; break;
    default: f = rows ? (const void*)k_dft_rows<T, 3> : (const void*)k_dft_cols<T, 3>; break;
  }
  NUSHT_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(c2_t<T>) * kDftNMAX)));
  void* argv[] = {(void*)&args...};
  NUSHT_HIP(hipLaunchKernel(f, grid, dim3(kDftNT), argv, lds, st));
}

template <typename T>
const FftLength<T>& cached_length(int n) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::unique_ptr<FftLength<T>>> cache;  // (device, n)
  int dev = 0;
  NUSHT_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto& slot = cache[{dev, n}];
  if (!slot) slot = std::make_unique<FftLength<T>>(n);
  return *slot;
}

template <typename T>
void fft2d_impl(c2_t<T>* data, int n0, int n1, int batch, int backward, hipStream_t st) {
  constexpr int cap = Fft2dCap<T>::n;
  NUSHT_REQUIRE(data && n0 >= 1 && n1 >= 1 && batch >= 1, kErrArg, "fft2d: bad arguments");
  const bool fast0 = fft_supported(n0) && n0 <= cap, fast1 = fft_supported(n1) && n1 <= cap;
  NUSHT_REQUIRE((fast0 || n0 <= kDftNMAX) && (fast1 || n1 <= kDftNMAX), kErrUnsupported,
                "fft2d: a side exceeds " + std::to_string(cap) + " (smooth) or " + std::to_string(kDftNMAX));
  NUSHT_REQUIRE(batch <= 65535 && n0 <= 65535, kErrUnsupported, "fft2d: too many rows or arrays for one launch");
  using T2 = c2_t<T>;
  // lengths first (table set-up may synchronize the stream once per new length)
  DftLen<T> L0{}, L1{};
  if (!fast0) L0 = any_length<T>(n0, st);
  if (!fast1) L1 = any_length<T>(n1, st);
  StageScope sc("fft2d", st);
  if (fast1) {
    const FftLength<T>& f1 = cached_length<T>(n1);
    NUSHT_HIP(hipFuncSetAttribute((const void*)k_fft_rows<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(T2) * cap)));
    hipLaunchKernelGGL(k_fft_rows<T>, dim3(n0, batch), dim3(kFft2dNT), sizeof(T2) * n1, st, data, n0, n1, f1.desc(),
                       backward);
  } else {
    launch_any<T>(true, L1.method, dim3(n0, batch), sizeof(T2) * L1.lds_values(), st, data, n0, n1, L1, backward);
  }
  NUSHT_LAUNCH_CHECK();
  if (fast0) {
    const FftLength<T>& f0 = cached_length<T>(n0);
    NUSHT_HIP(hipFuncSetAttribute((const void*)k_fft_cols<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(sizeof(T2) * cap)));
    const int C = std::max(1, std::min(16, cap / n0));
    hipLaunchKernelGGL(k_fft_cols<T>, dim3(ceil_div(n1, C), batch), dim3(kFft2dNT), sizeof(T2) * (size_t)C * n0, st,
                       data, n0, n1, C, f0.desc(), backward);
  } else {
    const int C = std::max(1, std::min(16, kDftNMAX / L0.lds_values()));
    launch_any<T>(false, L0.method, dim3(ceil_div(n1, C), batch), sizeof(T2) * (size_t)C * L0.lds_values(), st, data,
                  n0, n1, C, L0, backward);
  }
  NUSHT_LAUNCH_CHECK();
}

}  // namespace

void fft2d_c2c(double2* data, int n0, int n1, int batch, int backward, hipStream_t st) {
  fft2d_impl<double>(data, n0, n1, batch, backward, st);
}

void fft2d_c2c(float2* data, int n0, int n1, int batch, int backward, hipStream_t st) {
  fft2d_impl<float>(data, n0, n1, batch, backward, st);
}

}  // namespace nusht
