// Host-side planning for the in-LDS FFT engine: radix factorization and twiddle tables.
#include "fft_engine.h"

#include <cmath>

namespace nusht {

std::vector<int> fft_factorize(int n, bool radix16) {
  std::vector<int> r;
  if (n < 1) return r;
  int m = n;
  if (radix16)
    while (m % 16 == 0) { r.push_back(16); m /= 16; }
  const int big[] = {8, 4};
  for (int b : big)
    while (m % b == 0) { r.push_back(b); m /= b; }
  const int small[] = {2, 3, 5, 7, 13};
  for (int p : small)
    while (m % p == 0) { r.push_back(p); m /= p; }
  if (m != 1) return {};
  if (r.empty()) r.push_back(1);  // n == 1
  return r;
}

bool fft_supported(int n) { return n >= 1 && !fft_factorize(n).empty(); }

bool fft_needs_generic(int n) {
  for (int r : fft_factorize(n))
    if (r == 7 || r == 13) return true;
  return false;
}

int fft_good_size(int n) {
  int k = n < 1 ? 1 : n;
  while (true) {
    // 2-3-5 smooth only, so that the generic O(R^2) radices stay out of internally chosen lengths
    int m = k;
    for (int p : {2, 3, 5})
      while (m % p == 0) m /= p;
    if (m == 1) return k;
    ++k;
  }
}

template <typename T>
void FftLength<T>::init(int n, bool radix16) {
  auto rad = fft_factorize(n, radix16);
  NUSHT_REQUIRE(!rad.empty(), kErrUnsupported, "fft: unsupported length " + std::to_string(n));
  NUSHT_REQUIRE((int)rad.size() <= kFftMaxStages, kErrUnsupported, "fft: too many stages");
  if (rad.size() == 1 && rad[0] == 1) rad.clear();
  n_ = n;
  radices_ = rad;
  std::vector<c2_t<T>> tw(n);
  for (int k = 0; k < n; ++k) {
    // exact angle reduction: e^{-2 pi i k/n} via sincospi on the reduced fraction
    long kk = k;
    double c, s;
    if (2 * kk > n) kk -= n;
    double a = -2.0 * (double)kk / n;
    s = std::sin(M_PI * a);
    c = std::cos(M_PI * a);
    if (4 * std::labs(kk) == n) c = 0.0;
    if (2 * std::labs(kk) == n) s = 0.0;
    tw[k] = c2_t<T>{(T)c, (T)s};
  }
  upload(tw_, tw);
}

template class FftLength<float>;
template class FftLength<double>;

}  // namespace nusht
