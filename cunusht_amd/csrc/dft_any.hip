// Host side of the any-length DFT (see dft_any.h): method choice and tables per length.
#include "dft_any.h"

#include <cmath>

namespace nusht {
namespace {

// One-off table transform: out = FFT(in) of length fd.n.
template <typename T>
__global__ __launch_bounds__(kDftNT) void k_fft_table(FftDesc<T> fd, const c2_t<T>* in, c2_t<T>* out) {
  extern __shared__ unsigned char smem_raw[];
  c2_t<T>* buf = reinterpret_cast<c2_t<T>*>(smem_raw);
  for (int k = threadIdx.x; k < fd.n; k += kDftNT) buf[k] = in[k];
  __syncthreads();
  fft_lds<T, kDftNT, kDftNMAX, true>(buf, fd);
  for (int k = threadIdx.x; k < fd.n; k += kDftNT) out[k] = buf[k];
}

bool is_prime(long n) {
  if (n < 2) return false;
  for (long d = 2; d * d <= n; ++d)
    if (n % d == 0) return false;
  return true;
}

long powmod(long b, long e, long m) {
  long r = 1;
  b %= m;
  while (e > 0) {
    if (e & 1) r = r * b % m;
    b = b * b % m;
    e >>= 1;
  }
  return r;
}

long primitive_root(long n) {
  std::vector<long> fac;
  long m = n - 1;
  for (long d = 2; d * d <= m; ++d)
    if (m % d == 0) { fac.push_back(d); while (m % d == 0) m /= d; }
  if (m > 1) fac.push_back(m);
  for (long g = 2; g < n; ++g) {
    bool ok = true;
    for (long q : fac)
      if (powmod(g, (n - 1) / q, n) == 1) { ok = false; break; }
    if (ok) return g;
  }
  return -1;
}

// FFT of `in` (length fd.n) into a new device table
template <typename T>
void fft_table(const FftDesc<T>& fd, const std::vector<c2_t<T>>& in, DevBuf& out, hipStream_t st) {
  NUSHT_HIP(hipFuncSetAttribute((const void*)k_fft_table<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(c2_t<T>) * kDftNMAX)));
  DevBuf bt;
  upload(bt, in);
  out.alloc(sizeof(c2_t<T>) * fd.n);
  hipLaunchKernelGGL(k_fft_table<T>, dim3(1), dim3(kDftNT), sizeof(c2_t<T>) * fd.n, st, fd, bt.as<c2_t<T>>(),
                     out.as<c2_t<T>>());
  NUSHT_LAUNCH_CHECK();
  NUSHT_HIP(hipStreamSynchronize(st));
}

// Bluestein tables for a transform of length n padded to M: chirp c_j = e^{-i pi j^2 / n} and the FFT of
// the conjugate chirp wrapped onto the cyclic length M.
template <typename T>
void bluestein_tables(int n, const FftDesc<T>& fd, DevBuf& chirp, DevBuf& kern, hipStream_t st) {
  using T2 = c2_t<T>;
  const int M = fd.n;
  std::vector<T2> c(n), b(M, T2{T(0), T(0)});
  for (long j = 0; j < n; ++j) {
    long j2 = (j * j) % (2L * n);
    double ang = -M_PI * (double)j2 / n;
    c[j] = T2{(T)std::cos(ang), (T)std::sin(ang)};
    b[j] = T2{(T)std::cos(ang), (T)-std::sin(ang)};
    if (j > 0) b[M - j] = b[j];
  }
  upload(chirp, c);
  fft_table<T>(fd, b, kern, st);
}

}  // namespace

template <typename T>
const FftLength<T>& DftSet<T>::fft_of(int len) {
  auto it = ffts_.find(len);
  if (it == ffts_.end()) it = ffts_.emplace(len, std::make_unique<FftLength<T>>(len)).first;
  return *it->second;
}

template <typename T>
int DftSet<T>::add(int n, hipStream_t st) {
  auto it = index_.find(n);
  if (it != index_.end()) return it->second;
  NUSHT_REQUIRE(n >= 1 && n <= kDftNMAX, kErrUnsupported,
                "dft: length " + std::to_string(n) + " outside 1.." + std::to_string(kDftNMAX));
  auto t = std::make_unique<Tables>();
  DftLen<T>& d = t->d;
  d.n = n;
  if (fft_supported(n)) {
    d.method = (int)DftMethod::Direct;
    d.mlen = n;
  } else if (is_prime(n) && fft_supported(n - 1)) {
    d.method = (int)DftMethod::Rader;
    d.mlen = n - 1;
  } else if (fft_good_size(2 * n - 1) <= kDftNMAX) {
    d.method = (int)DftMethod::Bluestein;
    d.mlen = fft_good_size(2 * n - 1);
  } else {
    // n = r b with the fewest sub-transforms whose length b runs in LDS
    d.method = (int)DftMethod::Split;
    for (int r = 2; r <= n && d.r == 0; ++r) {
      if (n % r) continue;
      const int b = n / r;
      if (fft_supported(b)) {
        d.r = r; d.b = b; d.mlen = b; d.sub_bluestein = 0;
      } else if (fft_good_size(2 * b - 1) <= kDftNMAX) {
        d.r = r; d.b = b; d.mlen = fft_good_size(2 * b - 1); d.sub_bluestein = 1;
      }
    }
    NUSHT_REQUIRE(d.r > 0 && d.r < n, kErrUnsupported,
                  "dft: length " + std::to_string(n) + " not supported (prime with a non-smooth n - 1)");
    NUSHT_REQUIRE(d.r <= 64, kErrUnsupported, "dft: length " + std::to_string(n) + " needs too many sub-transforms");
  }
  d.fd = fft_of(d.mlen).desc();
  if (d.method == (int)DftMethod::Bluestein) {
    bluestein_tables<T>(n, d.fd, t->chirp, t->kern, st);
  } else if (d.method == (int)DftMethod::Split && d.sub_bluestein) {
    bluestein_tables<T>(d.b, d.fd, t->chirp, t->kern, st);
  } else if (d.method == (int)DftMethod::Rader) {
    using T2 = c2_t<T>;
    long gr = primitive_root(n);
    long ginv = powmod(gr, n - 2, n);
    std::vector<int> pin(n - 1), pout(n - 1);
    std::vector<T2> b(n - 1);
    long gi = 1, go = 1;
    for (int q = 0; q < n - 1; ++q) {
      pin[q] = (int)gi;
      pout[q] = (int)go;
      double ang = -2.0 * M_PI * (double)go / n;
      b[q] = T2{(T)std::cos(ang), (T)std::sin(ang)};
      gi = gi * gr % n;
      go = go * ginv % n;
    }
    upload(t->perm_in, pin);
    upload(t->perm_out, pout);
    fft_table<T>(d.fd, b, t->kern, st);
  }
  d.chirp = t->chirp.template as<c2_t<T>>();
  d.kern = t->kern.template as<c2_t<T>>();
  d.perm_in = t->perm_in.template as<int>();
  d.perm_out = t->perm_out.template as<int>();
  const int idx = (int)tabs_.size();
  tabs_.push_back(std::move(t));
  index_[n] = idx;
  return idx;
}

template <typename T>
size_t DftSet<T>::bytes() const {
  size_t b = 0;
  for (const auto& f : ffts_) b += f.second->bytes();
  for (const auto& t : tabs_) b += t->chirp.bytes() + t->kern.bytes() + t->perm_in.bytes() + t->perm_out.bytes();
  return b;
}

template class DftSet<double>;
template class DftSet<float>;

}  // namespace nusht
