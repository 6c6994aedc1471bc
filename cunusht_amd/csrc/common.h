// Shared host/device utilities for the nuSHT HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace nusht {

// ---------------------------------------------------------------- errors
enum Status : int {
  kOk = 0,
  kErrArg = 1,
  kErrHip = 2,
  kErrUnsupported = 3,
  kErrState = 4,
  kErrNoMem = 5,
};

void set_error(const std::string& msg);
const char* last_error();

struct Error {
  int code;
  std::string msg;
};

#define NUSHT_HIP(call)                                                                 \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      throw ::nusht::Error{::nusht::kErrHip, std::string(#call) + ": " + hipGetErrorString(_e)}; \
  } while (0)

#define NUSHT_REQUIRE(cond, code, msg)                           \
  do {                                                           \
    if (!(cond)) throw ::nusht::Error{(code), std::string(msg)}; \
  } while (0)

#define NUSHT_LAUNCH_CHECK() NUSHT_HIP(hipGetLastError())

// ---------------------------------------------------------------- complex
template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using c2_t = typename cvec<T>::type;

template <typename T2>
__host__ __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T2>
__host__ __device__ __forceinline__ T2 cmulc(T2 a, T2 b) {  // a * conj(b)
  return T2{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T2>
__host__ __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return T2{a.x + b.x, a.y + b.y}; }
template <typename T2>
__host__ __device__ __forceinline__ T2 csub(T2 a, T2 b) { return T2{a.x - b.x, a.y - b.y}; }
template <typename T2>
__host__ __device__ __forceinline__ T2 cconj(T2 a) { return T2{a.x, -a.y}; }
template <typename T2, typename S>
__host__ __device__ __forceinline__ T2 cscale(T2 a, S s) { return T2{a.x * s, a.y * s}; }

// ---------------------------------------------------------------- device memory
// RAII device buffer; allocation happens only at plan creation, never on the hot path.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p_, bytes);
    if (e != hipSuccess) {
      p_ = nullptr;
      throw Error{kErrNoMem, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e)};
    }
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p_); }
  void* get() const { return p_; }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

template <typename T>
inline void upload(DevBuf& b, const std::vector<T>& v) {
  b.alloc(v.size() * sizeof(T));
  if (!v.empty()) NUSHT_HIP(hipMemcpy(b.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Number of XCDs on MI355X: consecutive workgroup ids are dispatched round-robin over them.
constexpr int kNumXcd = 8;

// Maps a linear block id to a work index so that work items i, i+1, ... land on the same XCD
// (and hence share that XCD's L2) while still being spread over all XCDs.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  int per = nblocks / kNumXcd;
  int rem = nblocks % kNumXcd;
  int xcd = bid % kNumXcd;
  int slot = bid / kNumXcd;
  // the first `rem` XCDs receive one extra block
  int base = xcd * per + (xcd < rem ? xcd : rem);
  return base + slot;
}

}  // namespace nusht
