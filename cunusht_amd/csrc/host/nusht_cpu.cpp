// Host CPU backend core (C++17, OpenMP): spin-weighted Legendre synthesis and its adjoint on rings,
// ES-kernel interpolation and spreading on a window of rows of a fine torus grid, and the deflected
// pointing.  The FFTs around them run in Python (cunusht_amd/cpu_transformer.py).  Declared in
// include/nusht_cpu.h.
//
// Legendre: for each m >= 0 the normalised Wigner functions f_l = sqrt((2l+1)/4pi) d^l_{m,m'}(theta),
// m' = -s (the spin-s chain) and m' = +s (the spin -s chain), follow the three-term recursion in l
//   f_{l+1} = (alpha_l x - beta_l) f_l - gamma_l f_{l-1},  x = cos(theta),
//   alpha_l = F (l+1) / D1,  beta_l = F m m' / (l D1),  gamma_l = sqrt((2l+3)/(2l-1)) (l+1) D0 / (l D1),
//   F = sqrt((2l+1)(2l+3)),  D1 = sqrt(((l+1)^2 - m^2)((l+1)^2 - m'^2)),  D0 = sqrt((l^2 - m^2)(l^2 - m'^2)),
// vectorised over rings and started at l0 = max(m, |m'|), where Wigner's sum for d^{l0}_{m,m'} has a single
// term, evaluated in long double logarithms.  Values below 2^-300 are carried as (mantissa, 2^(600 k)),
// k < 0, and contribute once k reaches 0.
#include "../../../include/nusht_cpu.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
constexpr double kThresh = 0x1p300;
constexpr double kSmall = 0x1p-600;
constexpr double kZero = -1e9;  // scale exponent of a chain that is exactly zero (at a pole)

// log2 |d^j_{a,b}(theta)| + log2 sqrt((2j+1)/4pi) and the sign, j = max(|a|, |b|): the one surviving
// term of Wigner's sum, s = max(0, b - a).
void start_value(int j, int a, int b, double theta, long double& log2v, double& sgn) {
  const int s = std::max(0, b - a);
  const int pc = 2 * j + b - a - 2 * s, ps = a - b + 2 * s;
  const long double half = 0.5L * (long double)theta;
  const long double c = cosl(half), sn = sinl(half);
  if ((pc > 0 && c == 0.0L) || (ps > 0 && sn == 0.0L)) { sgn = 0.0; log2v = 0.0L; return; }
  long double t = 0.5L * (lgammal(j + a + 1.0L) + lgammal(j - a + 1.0L) + lgammal(j + b + 1.0L) +
                          lgammal(j - b + 1.0L)) -
                  (lgammal(j + b - s + 1.0L) + lgammal(s + 1.0L) + lgammal(a - b + s + 1.0L) +
                   lgammal(j - a - s + 1.0L));
  if (pc) t += pc * logl(c);
  if (ps) t += ps * logl(sn);
  t += 0.5L * logl((2.0L * j + 1.0L) / (4.0L * 3.14159265358979323846264338327950288L));
  log2v = t / 0.693147180559945309417232121458176568L;
  sgn = ((a - b + s) & 1) ? -1.0 : 1.0;
}

// One recursion chain (fixed m, m') over all rings.
struct Chain {
  int m = 0, mp = 0, l0 = 0;
  std::vector<double> cur, prev, k;  // mantissas and scale exponents (k <= 0)
  int nunder = 0;                    // rings whose value is still below range

  void start(int lmax, const std::vector<double>& theta, int m_, int mp_) {
    m = m_;
    mp = mp_;
    l0 = std::max(m, std::abs(mp));
    const size_t n = theta.size();
    cur.assign(n, 0.0);
    prev.assign(n, 0.0);
    k.assign(n, 0.0);
    nunder = 0;
    if (l0 > lmax) return;
    for (size_t r = 0; r < n; ++r) {
      long double l2;
      double sg;
      start_value(l0, m, mp, theta[r], l2, sg);
      if (sg == 0.0) { k[r] = kZero; continue; }
      int kk = 0;
      if (l2 < -300.0L) kk = -(int)ceill((-300.0L - l2) / 600.0L);
      cur[r] = sg * (double)exp2l(l2 - 600.0L * kk);
      k[r] = kk;
      if (kk < 0) ++nunder;
    }
  }

  // f_l -> f_{l+1}
  void step(int l, const double* x) {
    const size_t n = cur.size();
    double* c = cur.data();
    double* p = prev.data();
    if (l == 0) {  // m = m' = 0: f_1 = sqrt(3) x f_0
      const double s3 = std::sqrt(3.0);
      for (size_t r = 0; r < n; ++r) {
        const double nx = s3 * x[r] * c[r];
        p[r] = c[r];
        c[r] = nx;
      }
    } else {
      const double lp = l + 1.0, dm = m, dmp = mp;
      const double d1 = std::sqrt((lp * lp - dm * dm) * (lp * lp - dmp * dmp));
      const double d0 = std::sqrt(((double)l * l - dm * dm) * ((double)l * l - dmp * dmp));
      const double F = std::sqrt((2.0 * l + 3.0) * (2.0 * l + 1.0));
      const double al = F * lp / d1, be = F * dm * dmp / (l * d1);
      const double ga = std::sqrt((2.0 * l + 3.0) / (2.0 * l - 1.0)) * lp * d0 / (l * d1);
#pragma omp simd
      for (size_t r = 0; r < n; ++r) {
        const double nx = (al * x[r] - be) * c[r] - ga * p[r];
        p[r] = c[r];
        c[r] = nx;
      }
    }
    if (nunder) {
      nunder = 0;
      for (size_t r = 0; r < n; ++r) {
        if (k[r] < 0.0 && k[r] > kZero) {
          if (std::fabs(c[r]) > kThresh) {
            c[r] *= kSmall;
            p[r] *= kSmall;
            k[r] += 1.0;
          }
          if (k[r] < 0.0) ++nunder;
        }
      }
    }
  }

  // f at the current l (zero while below range or at a pole where it vanishes)
  double val(size_t r) const { return k[r] == 0.0 ? cur[r] : 0.0; }
};

inline long nalm_of(int lmax) { return (long)(lmax + 1) * (lmax + 2) / 2; }
inline long idx_of(int lmax, int l, int m) { return (long)m * (2 * lmax + 1 - m) / 2 + l; }

int threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

inline double psi(double z, double beta) {
  const double t = 1.0 - z * z;
  return t > 0.0 ? std::exp(beta * (std::sqrt(t) - 1.0)) : 0.0;
}

constexpr int kMaxW = 32;

// Footprint of one point: first row / column and the kernel weights in both directions.
struct Foot {
  int i0, j0;
  double kt[kMaxW], kp[kMaxW];
};

inline void footprint(double theta, double phi, int nf1, int nf2, int w, double beta, Foot& f) {
  const double ut = theta * (nf1 / (2.0 * M_PI));
  double up = phi * (nf2 / (2.0 * M_PI));
  up -= nf2 * std::floor(up / nf2);
  const double hw = 0.5 * w, inv = 2.0 / w;
  f.i0 = (int)std::ceil(ut - hw);
  f.j0 = (int)std::ceil(up - hw);
  for (int i = 0; i < w; ++i) {
    f.kt[i] = psi((f.i0 + i - ut) * inv, beta);
    f.kp[i] = psi((f.j0 + i - up) * inv, beta);
  }
}

inline int wrap(int c, int n) { return c >= n ? c - n : (c < 0 ? c + n : c); }

bool grid_args_ok(int nf1, int nf2, int row_lo, int nrow, int w, double beta) {
  if (nf1 < 2 || nf2 < 2 || w < 2 || w > kMaxW || w > nf2 || !(beta > 0.0)) return false;
  if (nrow < 1 || nrow > nf1) return false;
  if (nrow == nf1 && row_lo != 0) return false;
  return true;
}

// row of the window for torus row t, -1 if outside
inline int window_row(int t, int row_lo, int nrow, int nf1) {
  if (nrow == nf1) return ((t % nf1) + nf1) % nf1;
  const int r = t - row_lo;
  return (r >= 0 && r < nrow) ? r : -1;
}

void spread_point(const Foot& f, cd v, int nf1, int nf2, int row_lo, int nrow, int w, cd* grid) {
  for (int i = 0; i < w; ++i) {
    const int rr = window_row(f.i0 + i, row_lo, nrow, nf1);
    if (rr < 0) continue;
    cd* row = grid + (size_t)rr * nf2;
    const cd vi = v * f.kt[i];
    for (int j = 0; j < w; ++j) row[wrap(f.j0 + j, nf2)] += vi * f.kp[j];
  }
}

}  // namespace

extern "C" {

int nusht_cpu_version(void) { return 1; }

int nusht_cpu_synth_rings(int lmax, int spin, int ncomp, const double* alm, int nring, const double* theta,
                          double* G, int nthreads) {
  if (lmax < 0 || spin < 0 || spin > 3 || nring < 1 || !alm || !theta || !G) return 1;
  if (!(ncomp == 1 || (spin > 0 && ncomp == 2))) return 1;
  for (int r = 0; r < nring; ++r)
    if (!(theta[r] >= 0.0 && theta[r] <= M_PI)) return 1;
  const cd* a = reinterpret_cast<const cd*>(alm);
  cd* g = reinterpret_cast<cd*>(G);
  const long nalm = nalm_of(lmax);
  const int nrows = spin ? 2 * lmax + 1 : lmax + 1;
  const std::vector<double> th(theta, theta + nring);
  std::vector<double> x(nring);
  for (int r = 0; r < nring; ++r) x[r] = std::cos(th[r]);
  std::memset(G, 0, sizeof(cd) * (size_t)nrows * nring);
  const double ss = (spin & 1) ? -1.0 : 1.0;  // (-1)^s
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads(nthreads))
  for (int m = lmax; m >= 0; --m) {  // long chains first
    Chain u, v;
    std::vector<cd> gp(nring, 0.0), gm(nring, 0.0);
    u.start(lmax, th, m, -spin);
    if (spin) v.start(lmax, th, m, spin);
    for (int l = u.l0; l <= lmax; ++l) {
      const cd e = a[idx_of(lmax, l, m)];
      if (!spin) {
        for (int r = 0; r < nring; ++r) gp[r] += e * u.val(r);
      } else {
        // Q + iU = -sum (E + iB) sY_lm with sY_lm = (-1)^s f(m' = -s); the m < 0 rows take
        // -(conj E + i conj B) f(m' = +s)
        const cd b = ncomp > 1 ? a[nalm + idx_of(lmax, l, m)] : cd(0.0);
        const cd ep = ss * (e + cd(0.0, 1.0) * b);
        const cd em = std::conj(e) + cd(0.0, 1.0) * std::conj(b);
        for (int r = 0; r < nring; ++r) {
          gp[r] -= ep * u.val(r);
          if (m > 0) gm[r] -= em * v.val(r);
        }
      }
      if (l < lmax) {
        u.step(l, x.data());
        if (spin) v.step(l, x.data());
      }
    }
    cd* rowp = g + (size_t)(spin ? lmax + m : m) * nring;
    for (int r = 0; r < nring; ++r) rowp[r] = gp[r];
    if (spin && m > 0) {
      cd* rowm = g + (size_t)(lmax - m) * nring;
      for (int r = 0; r < nring; ++r) rowm[r] = gm[r];
    }
  }
  return 0;
}

int nusht_cpu_adjoint_rings(int lmax, int spin, int ncomp, const double* G, int nring, const double* theta,
                            double* alm, int nthreads) {
  if (lmax < 0 || spin < 0 || spin > 3 || nring < 1 || !alm || !theta || !G) return 1;
  if (!(ncomp == 1 || (spin > 0 && ncomp == 2))) return 1;
  for (int r = 0; r < nring; ++r)
    if (!(theta[r] >= 0.0 && theta[r] <= M_PI)) return 1;
  const cd* g = reinterpret_cast<const cd*>(G);
  cd* a = reinterpret_cast<cd*>(alm);
  const long nalm = nalm_of(lmax);
  const std::vector<double> th(theta, theta + nring);
  std::vector<double> x(nring);
  for (int r = 0; r < nring; ++r) x[r] = std::cos(th[r]);
  std::memset(alm, 0, sizeof(cd) * (size_t)nalm * ncomp);
  const double ss = (spin & 1) ? -1.0 : 1.0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads(nthreads))
  for (int m = lmax; m >= 0; --m) {
    Chain u, v;
    u.start(lmax, th, m, -spin);
    if (spin) v.start(lmax, th, m, spin);
    const cd* rowp = g + (size_t)(spin ? lmax + m : m) * nring;
    const cd* rowm = spin ? g + (size_t)(lmax - m) * nring : nullptr;
    for (int l = u.l0; l <= lmax; ++l) {
      if (!spin) {
        cd acc = 0.0;
        for (int r = 0; r < nring; ++r) acc += rowp[r] * u.val(r);
        a[idx_of(lmax, l, m)] = m == 0 ? cd(acc.real(), 0.0) : acc;
      } else {
        // X = (-1)^s sum G'_m f(m' = -s), Y = sum conj(G'_{-m}) f(m' = +s):
        // E' = -(X + Y) / 2, B' = i (X - Y) / 2
        cd X = 0.0, Y = 0.0;
        for (int r = 0; r < nring; ++r) {
          X += rowp[r] * u.val(r);
          Y += std::conj(rowm[r]) * v.val(r);
        }
        X *= ss;
        a[idx_of(lmax, l, m)] = -0.5 * (X + Y);
        if (ncomp > 1) a[nalm + idx_of(lmax, l, m)] = cd(0.0, 0.5) * (X - Y);
      }
      if (l < lmax) {
        u.step(l, x.data());
        if (spin) v.step(l, x.data());
      }
    }
  }
  return 0;
}

int nusht_cpu_interp(int nf1, int nf2, int row_lo, int nrow, int w, double beta, const double* grid, int64_t npts,
                     const double* theta, const double* phi, double* out, int nthreads) {
  if (!grid_args_ok(nf1, nf2, row_lo, nrow, w, beta) || npts < 0 || !grid || !theta || !phi || !out) return 1;
  const cd* gr = reinterpret_cast<const cd*>(grid);
  cd* o = reinterpret_cast<cd*>(out);
#pragma omp parallel for schedule(static) num_threads(threads(nthreads))
  for (int64_t p = 0; p < npts; ++p) {
    Foot f;
    footprint(theta[p], phi[p], nf1, nf2, w, beta, f);
    cd s = 0.0;
    for (int i = 0; i < w; ++i) {
      const int rr = window_row(f.i0 + i, row_lo, nrow, nf1);
      if (rr < 0) continue;
      const cd* row = gr + (size_t)rr * nf2;
      cd rs = 0.0;
      for (int j = 0; j < w; ++j) rs += row[wrap(f.j0 + j, nf2)] * f.kp[j];
      s += rs * f.kt[i];
    }
    o[p] = s;
  }
  return 0;
}

int nusht_cpu_spread(int nf1, int nf2, int row_lo, int nrow, int w, double beta, const double* in, int64_t npts,
                     const double* theta, const double* phi, double* grid, int nthreads) {
  if (!grid_args_ok(nf1, nf2, row_lo, nrow, w, beta) || npts < 0 || !grid || !theta || !phi || !in) return 1;
  const cd* v = reinterpret_cast<const cd*>(in);
  cd* gr = reinterpret_cast<cd*>(grid);
  const int nth = threads(nthreads);
  if (nrow == nf1 || nth == 1) {  // wrapping window (small grids) or one thread: in order
    for (int64_t p = 0; p < npts; ++p) {
      Foot f;
      footprint(theta[p], phi[p], nf1, nf2, w, beta, f);
      spread_point(f, v[p], nf1, nf2, row_lo, nrow, w, gr);
    }
    return 0;
  }
  // Bands of H >= w window rows, points bucketed by their first row; a point of band b writes rows
  // [b H, (b + 2) H), so the bands of one parity never share a row and run in parallel without atomics.
  const int H = std::max(w, (nrow + 4 * nth - 1) / (4 * nth));
  const int nb = (nrow + H - 1) / H + 1;
  std::vector<int> band(npts);
  std::vector<int64_t> start(nb + 1, 0);
  for (int64_t p = 0; p < npts; ++p) {
    const double ut = theta[p] * (nf1 / (2.0 * M_PI));
    const int r0 = (int)std::ceil(ut - 0.5 * w) - row_lo;
    const int b = std::min(std::max(r0, 0) / H, nb - 1);
    band[p] = b;
    ++start[b + 1];
  }
  for (int b = 0; b < nb; ++b) start[b + 1] += start[b];
  std::vector<int64_t> order(npts);
  {
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (int64_t p = 0; p < npts; ++p) order[pos[band[p]]++] = p;
  }
  for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int b = parity; b < nb; b += 2) {
      for (int64_t q = start[b]; q < start[b + 1]; ++q) {
        const int64_t p = order[q];
        Foot f;
        footprint(theta[p], phi[p], nf1, nf2, w, beta, f);
        spread_point(f, v[p], nf1, nf2, row_lo, nrow, w, gr);
      }
    }
  }
  return 0;
}

int nusht_cpu_pointing(int64_t n, const double* theta, const double* phi, const double* dt, const double* dp,
                       double* tp, double* pp, double* gamma, int nthreads) {
  if (n < 0 || !theta || !phi || !dt || !dp || !tp || !pp) return 1;
#pragma omp parallel for schedule(static) num_threads(threads(nthreads))
  for (int64_t i = 0; i < n; ++i) {
    const double t = theta[i], a2 = dt[i] * dt[i] + dp[i] * dp[i], a = std::sqrt(a2);
    double sa, ca;  // sin(a) / a and cos(a), by their series for small a
    if (a < 1e-3) {
      sa = 1.0 - a2 / 6.0 * (1.0 - a2 / 20.0 * (1.0 - a2 / 42.0));
      ca = 1.0 - a2 / 2.0 * (1.0 - a2 / 12.0 * (1.0 - a2 / 30.0 * (1.0 - a2 / 56.0)));
    } else {
      sa = std::sin(a) / a;
      ca = std::cos(a);
    }
    const double st = std::sin(t), ct = std::cos(t);
    // unit vector of the end point of the geodesic of length a leaving (t, 0) in direction (dt, dp)
    const double n1 = st * ca + ct * dt[i] * sa, n2 = -dp[i] * sa, n3 = ct * ca - st * dt[i] * sa;
    tp[i] = std::atan2(std::sqrt(n1 * n1 + n2 * n2), n3);
    const double q = phi[i] - std::atan2(n2, n1);
    pp[i] = q >= 2.0 * M_PI ? q - 2.0 * M_PI : q;  // as the GPU pointing kernel: in [-a, 2 pi)
    if (gamma) gamma[i] = std::atan2(dp[i], dt[i]) - std::atan2(dp[i], a2 * sa * ct / st + dt[i] * ca);
  }
  return 0;
}

}  // extern "C"
