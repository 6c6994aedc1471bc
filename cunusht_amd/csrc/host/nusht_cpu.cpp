// Host CPU backend core (C++17, OpenMP): spin-weighted Legendre synthesis and its adjoint on rings,
// ES-kernel interpolation and spreading on a window of rows of a fine torus grid (kernel weights from
// per-cell polynomials, a phase e^{i (c phi + s gamma)} folded in), and the deflected pointing.  The FFTs around them run in Python (cunusht_amd/cpu_transformer.py).  Declared in
// include/nusht_cpu.h.
//
// Legendre: for each m >= 0 the normalised Wigner functions f_l = sqrt((2l+1)/4pi) d^l_{m,m'}(theta),
// m' = -s (the spin-s chain) and m' = +s (the spin -s chain), follow the three-term recursion in l
//   f_{l+1} = (alpha_l x - beta_l) f_l - gamma_l f_{l-1},  x = cos(theta),
//   alpha_l = F (l+1) / D1,  beta_l = F m m' / (l D1),  gamma_l = sqrt((2l+3)/(2l-1)) (l+1) D0 / (l D1),
//   F = sqrt((2l+1)(2l+3)),  D1 = sqrt(((l+1)^2 - m^2)((l+1)^2 - m'^2)),  D0 = sqrt((l^2 - m^2)(l^2 - m'^2)),
// run on blocks of 64 rings (the block's state stays in L1) and started at l0 = max(m, |m'|), where Wigner's
// sum for d^{l0}_{m,m'} has a single term, evaluated in long double logarithms (the factorials once per chain,
// log cos(theta/2) and log sin(theta/2) once per ring).  Values below 2^-300 are carried as
// (mantissa, 2^(600 k)), k < 0, and contribute once k reaches 0.
#include "../../../include/nusht_cpu.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
constexpr double kThresh = 0x1p300;
constexpr double kSmall = 0x1p-600;

// Recursion coefficients and start constants of one chain (fixed m, m').
struct ChainCoef {
  int m = 0, mp = 0, l0 = 0;
  int pc = 0, ps = 0;     // powers of cos(theta/2) and sin(theta/2) in d^{l0}_{m,m'}
  long double lconst = 0;  // theta-independent part of log f_{l0}: factorials and sqrt((2 l0 + 1) / 4pi)
  double sgn = 1.0;
  std::vector<double> al, be, ga;  // index l - l0, l0 <= l < lmax

  void prepare(int lmax, int m_, int mp_) {
    m = m_;
    mp = mp_;
    l0 = std::max(m, std::abs(mp));
    if (l0 > lmax) return;
    // Wigner's sum for d^{j}_{a,b}, j = max(|a|, |b|), has the single term s = max(0, b - a)
    const int j = l0, a = m, b = mp, s = std::max(0, b - a);
    pc = 2 * j + b - a - 2 * s;
    ps = a - b + 2 * s;
    lconst = 0.5L * (lgammal(j + a + 1.0L) + lgammal(j - a + 1.0L) + lgammal(j + b + 1.0L) + lgammal(j - b + 1.0L)) -
             (lgammal(j + b - s + 1.0L) + lgammal(s + 1.0L) + lgammal(a - b + s + 1.0L) + lgammal(j - a - s + 1.0L)) +
             0.5L * logl((2.0L * j + 1.0L) / (4.0L * 3.14159265358979323846264338327950288L));
    sgn = ((a - b + s) & 1) ? -1.0 : 1.0;
    const int n = lmax - l0;
    al.assign(n, 0.0);
    be.assign(n, 0.0);
    ga.assign(n, 0.0);
    for (int l = l0; l < lmax; ++l) {
      if (l == 0) {  // m = m' = 0: f_1 = sqrt(3) x f_0
        al[0] = std::sqrt(3.0);
        continue;
      }
      const double lp = l + 1.0, dm = m, dmp = mp;
      const double d1 = std::sqrt((lp * lp - dm * dm) * (lp * lp - dmp * dmp));
      const double d0 = std::sqrt(((double)l * l - dm * dm) * ((double)l * l - dmp * dmp));
      const double F = std::sqrt((2.0 * l + 3.0) * (2.0 * l + 1.0));
      al[l - l0] = F * lp / d1;
      be[l - l0] = F * dm * dmp / (l * d1);
      ga[l - l0] = std::sqrt((2.0 * l + 3.0) / (2.0 * l - 1.0)) * lp * d0 / (l * d1);
    }
  }
};

// log cos(theta/2), log sin(theta/2) of every ring (long double), -inf where they vanish.
struct RingLogs {
  std::vector<long double> lc, ls;
  std::vector<double> x;
  explicit RingLogs(const double* theta, int n) : lc(n), ls(n), x(n) {
    for (int r = 0; r < n; ++r) {
      const long double h = 0.5L * (long double)theta[r];
      const long double c = cosl(h), s = sinl(h);
      lc[r] = c > 0.0L ? logl(c) : -INFINITY;
      ls[r] = s > 0.0L ? logl(s) : -INFINITY;
      x[r] = std::cos(theta[r]);
    }
  }
};

constexpr int kRB = 64;  // rings per block: a block's recursion state stays in L1

// Recursion state of one chain on a block of rings.  Values below 2^-300 are held as mantissa * 2^(600 k),
// k < 0; `live` is 1 once k = 0 (0 before, and for a chain that vanishes identically at a pole).
struct BlockChain {
  double c[kRB], p[kRB], live[kRB];
  int k[kRB];
  int nunder = 0;

  void start(const ChainCoef& cf, const RingLogs& rl, int r0, int nb) {
    nunder = 0;
    for (int i = 0; i < kRB; ++i) {
      c[i] = p[i] = live[i] = 0.0;
      k[i] = 0;
    }
    for (int i = 0; i < nb; ++i) {
      const int r = r0 + i;
      if ((cf.pc > 0 && rl.lc[r] == -INFINITY) || (cf.ps > 0 && rl.ls[r] == -INFINITY)) {
        k[i] = INT32_MIN / 2;  // exactly zero for every l
        continue;
      }
      long double t = cf.lconst;
      if (cf.pc) t += cf.pc * rl.lc[r];
      if (cf.ps) t += cf.ps * rl.ls[r];
      const long double l2 = t / 0.693147180559945309417232121458176568L;
      int kk = 0;
      if (l2 < -300.0L) kk = -(int)ceill((-300.0L - l2) / 600.0L);
      c[i] = cf.sgn * (double)exp2l(l2 - 600.0L * kk);
      k[i] = kk;
      if (kk < 0) ++nunder;
      else live[i] = 1.0;
    }
  }

  // f_l -> f_{l+1}
  inline void step(const ChainCoef& cf, int l, const double* x) {
    const double al = cf.al[l - cf.l0], be = cf.be[l - cf.l0], ga = cf.ga[l - cf.l0];
#pragma omp simd
    for (int i = 0; i < kRB; ++i) {
      const double nx = (al * x[i] - be) * c[i] - ga * p[i];
      p[i] = c[i];
      c[i] = nx;
    }
    if (nunder) {
      nunder = 0;
      for (int i = 0; i < kRB; ++i) {
        if (k[i] < 0 && k[i] > INT32_MIN / 4) {
          if (std::fabs(c[i]) > kThresh) {
            c[i] *= kSmall;
            p[i] *= kSmall;
            if (++k[i] == 0) live[i] = 1.0;
          }
          if (k[i] < 0) ++nunder;
        }
      }
    }
  }
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

inline long double psi_l(long double z, long double beta) {
  const long double t = 1.0L - z * z;
  return t > 0.0L ? expl(beta * (sqrtl(t) - 1.0L)) : 0.0L;
}

constexpr int kMaxW = 32;
constexpr int kMaxDeg = 24;

// Piecewise-polynomial ES kernel: a point at fractional offset t in [0, 1) from its first cell needs
// psi(z_i), z_i = (t + i - w/2) 2/w, i = 0..w-1; each cell gets its own polynomial in u = t - 1/2,
// interpolated at deg + 1 Chebyshev nodes (Vandermonde solve in long double), coefficient k of cell i at
// c[k * kMaxW + i].  deg is the lowest in [2, kMaxDeg] at which every cell but the two outer ones (which hold
// the square-root branch point of psi at |z| = 1) has max error <= tol on a dense sample; cells still above
// tol at that degree are evaluated exactly.
struct EsPoly {
  int w = 0, deg = 0;
  double err = 0.0;           // max error of the polynomial cells
  std::vector<double> c;
  std::vector<int> exact;     // cells evaluated exactly
};

EsPoly fit_es(int w, double beta, double tol) {
  EsPoly out;
  out.w = w;
  for (int d = 2; d <= kMaxDeg; ++d) {
    const int n = d + 1;
    std::vector<double> c((size_t)n * kMaxW, 0.0);
    std::vector<double> cell_err(w, 0.0);
    std::vector<long double> un(n);
    for (int j = 0; j < n; ++j) un[j] = 0.5L * cosl(3.14159265358979323846264338327950288L * (j + 0.5L) / n);
    for (int i = 0; i < w; ++i) {
      std::vector<long double> A((size_t)n * (n + 1));
      for (int r = 0; r < n; ++r) {
        long double pw = 1.0L;
        for (int k = 0; k < n; ++k) {
          A[(size_t)r * (n + 1) + k] = pw;
          pw *= un[r];
        }
        A[(size_t)r * (n + 1) + n] = psi_l((un[r] + 0.5L + i - 0.5L * w) * 2.0L / w, beta);
      }
      for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
          if (fabsl(A[(size_t)r * (n + 1) + col]) > fabsl(A[(size_t)piv * (n + 1) + col])) piv = r;
        if (piv != col)
          for (int k = 0; k <= n; ++k) std::swap(A[(size_t)col * (n + 1) + k], A[(size_t)piv * (n + 1) + k]);
        for (int r = col + 1; r < n; ++r) {
          const long double f = A[(size_t)r * (n + 1) + col] / A[(size_t)col * (n + 1) + col];
          for (int k = col; k <= n; ++k) A[(size_t)r * (n + 1) + k] -= f * A[(size_t)col * (n + 1) + k];
        }
      }
      std::vector<long double> xs(n);
      for (int r = n - 1; r >= 0; --r) {
        long double v = A[(size_t)r * (n + 1) + n];
        for (int k = r + 1; k < n; ++k) v -= A[(size_t)r * (n + 1) + k] * xs[k];
        xs[r] = v / A[(size_t)r * (n + 1) + r];
      }
      for (int k = 0; k < n; ++k) c[(size_t)k * kMaxW + i] = (double)xs[k];
      for (int q = 0; q <= 256; ++q) {
        const double u = -0.5 + q / 256.0;
        double v = c[(size_t)d * kMaxW + i];
        for (int k = d - 1; k >= 0; --k) v = v * u + c[(size_t)k * kMaxW + i];
        cell_err[i] = std::max(cell_err[i],
                               (double)fabsl((long double)v - psi_l((u + 0.5L + i - 0.5L * w) * 2.0L / w, beta)));
      }
    }
    bool inner_ok = true;
    for (int i = 1; i + 1 < w; ++i) inner_ok = inner_ok && cell_err[i] <= tol;
    if (inner_ok || d == kMaxDeg) {
      out.deg = d;
      out.c = c;
      out.err = 0.0;
      out.exact.clear();
      for (int i = 0; i < w; ++i) {
        if (cell_err[i] > tol) out.exact.push_back(i);
        else out.err = std::max(out.err, cell_err[i]);
      }
      break;
    }
  }
  return out;
}

// kernel fits are cached per (w, beta, tol): a call fits once, later calls reuse it
const EsPoly& es_poly(int w, double beta, double tol) {
  static std::mutex mu;
  static std::map<std::tuple<int, double, double>, EsPoly> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_tuple(w, beta, tol);
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, fit_es(w, beta, tol)).first;
  return it->second;
}

// Footprint of one point: first row / column and the kernel weights in both directions.
struct Foot {
  int i0, j0;
  double kt[kMaxW], kp[kMaxW];
};

// kernel values of all cells at offset u - 1/2 (cells w.. up to a multiple of 4 hold zero polynomials)
inline void horner(const EsPoly& k, double u, double* out) {
  const int d = k.deg, w4 = (k.w + 3) & ~3;
  const double* c = k.c.data();
#pragma omp simd
  for (int i = 0; i < w4; ++i) out[i] = c[(size_t)d * kMaxW + i];
  for (int q = d - 1; q >= 0; --q) {
#pragma omp simd
    for (int i = 0; i < w4; ++i) out[i] = out[i] * u + c[(size_t)q * kMaxW + i];
  }
}

inline void footprint(double theta, double phi, const nusht_cpu_grid& g, const EsPoly& k, Foot& f) {
  const double ut = theta * (g.nf1 / (2.0 * M_PI));
  double up = phi * (g.nf2 / (2.0 * M_PI));
  up -= g.nf2 * std::floor(up / g.nf2);
  const double hw = 0.5 * g.w;
  f.i0 = (int)std::ceil(ut - hw);
  f.j0 = (int)std::ceil(up - hw);
  // offsets t = i0 - (u - w/2) in [0, 1); the polynomials take u = t - 1/2
  const double t1 = f.i0 - (ut - hw), t2 = f.j0 - (up - hw);
  horner(k, t1 - 0.5, f.kt);
  horner(k, t2 - 0.5, f.kp);
  for (int i : k.exact) {
    f.kt[i] = psi((t1 + i - hw) * (2.0 / g.w), g.beta);
    f.kp[i] = psi((t2 + i - hw) * (2.0 / g.w), g.beta);
  }
}

bool grid_ok(const nusht_cpu_grid* g) {
  if (!g) return false;
  if (g->nf1 < 2 || g->nf2 < 2 || g->w < 2 || g->w > kMaxW || g->w > g->nf2 || !(g->beta > 0.0)) return false;
  if (!(g->ktol > 0.0)) return false;
  if (g->nrow < 1 || g->nrow > g->nf1) return false;
  if (g->nrow == g->nf1 && g->row_lo != 0) return false;
  return true;
}

// row of the window for torus row t, -1 if outside
inline int window_row(int t, const nusht_cpu_grid& g) {
  if (g.nrow == g.nf1) return ((t % g.nf1) + g.nf1) % g.nf1;
  const int r = t - g.row_lo;
  return (r >= 0 && r < g.nrow) ? r : -1;
}

inline int wrap(int c, int n) { return c >= n ? c - n : (c < 0 ? c + n : c); }

// e^{i (mshift phi + spin gamma)}
inline cd phase(double phi, const double* gamma, int64_t p, int mshift, int spin) {
  double a = mshift * phi;
  if (gamma && spin) a += spin * gamma[p];
  if (a == 0.0) return cd(1.0, 0.0);
  return cd(std::cos(a), std::sin(a));
}

cd interp_point(const Foot& f, const nusht_cpu_grid& g, const double* grid) {
  const int w = g.w, w2 = 2 * w;
  if (f.j0 >= 0 && f.j0 + w <= g.nf2) {
    // sum_j kp_j sum_i kt_i row_i[j0 + j], the inner sums over contiguous (re, im) pairs
    double col[2 * kMaxW] = {0};
    for (int i = 0; i < w; ++i) {
      const int rr = window_row(f.i0 + i, g);
      if (rr < 0) continue;
      const double* row = grid + 2 * ((size_t)rr * g.nf2 + f.j0);
      const double kt = f.kt[i];
#pragma omp simd
      for (int q = 0; q < w2; ++q) col[q] += kt * row[q];
    }
    double re = 0.0, im = 0.0;
#pragma omp simd reduction(+ : re, im)
    for (int j = 0; j < w; ++j) {
      re += f.kp[j] * col[2 * j];
      im += f.kp[j] * col[2 * j + 1];
    }
    return cd(re, im);
  }
  const cd* gr = reinterpret_cast<const cd*>(grid);
  cd s = 0.0;
  for (int i = 0; i < w; ++i) {
    const int rr = window_row(f.i0 + i, g);
    if (rr < 0) continue;
    const cd* row = gr + (size_t)rr * g.nf2;
    cd rs = 0.0;
    for (int j = 0; j < w; ++j) rs += row[wrap(f.j0 + j, g.nf2)] * f.kp[j];
    s += rs * f.kt[i];
  }
  return s;
}

void spread_point(const Foot& f, cd v, const nusht_cpu_grid& g, double* grid) {
  const int w = g.w, w2 = 2 * w;
  if (f.j0 >= 0 && f.j0 + w <= g.nf2) {
    double vk[2 * kMaxW];
    for (int j = 0; j < w; ++j) {
      vk[2 * j] = v.real() * f.kp[j];
      vk[2 * j + 1] = v.imag() * f.kp[j];
    }
    for (int i = 0; i < w; ++i) {
      const int rr = window_row(f.i0 + i, g);
      if (rr < 0) continue;
      double* row = grid + 2 * ((size_t)rr * g.nf2 + f.j0);
      const double kt = f.kt[i];
#pragma omp simd
      for (int q = 0; q < w2; ++q) row[q] += kt * vk[q];
    }
    return;
  }
  cd* gr = reinterpret_cast<cd*>(grid);
  for (int i = 0; i < w; ++i) {
    const int rr = window_row(f.i0 + i, g);
    if (rr < 0) continue;
    cd* row = gr + (size_t)rr * g.nf2;
    const cd vi = v * f.kt[i];
    for (int j = 0; j < w; ++j) row[wrap(f.j0 + j, g.nf2)] += vi * f.kp[j];
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
  const RingLogs rl(theta, nring);
  std::memset(G, 0, sizeof(cd) * (size_t)nrows * nring);
  const double ss = (spin & 1) ? -1.0 : 1.0;  // (-1)^s
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads(nthreads))
  for (int m = lmax; m >= 0; --m) {  // long chains first
    ChainCoef cu, cv;
    cu.prepare(lmax, m, -spin);
    if (cu.l0 > lmax) continue;
    if (spin) cv.prepare(lmax, m, spin);
    // per l: spin 0 takes a_lm; spin s takes ep = (-1)^s (E + iB) for the m row and
    // em = conj E + i conj B for the -m row (Q + iU = -sum (E + iB) sY_lm, sY_lm = (-1)^s f(m' = -s),
    // the m < 0 rows use f(m' = +s))
    const int nl = lmax - cu.l0 + 1;
    std::vector<double> epr(nl), epi(nl), emr(nl), emi(nl);
    for (int l = cu.l0; l <= lmax; ++l) {
      const cd e = a[idx_of(lmax, l, m)];
      if (!spin) {
        epr[l - cu.l0] = e.real();
        epi[l - cu.l0] = e.imag();
      } else {
        const cd b = ncomp > 1 ? a[nalm + idx_of(lmax, l, m)] : cd(0.0);
        const cd ep = -ss * (e + cd(0.0, 1.0) * b);
        const cd em = -(std::conj(e) + cd(0.0, 1.0) * std::conj(b));
        epr[l - cu.l0] = ep.real();
        epi[l - cu.l0] = ep.imag();
        emr[l - cu.l0] = em.real();
        emi[l - cu.l0] = em.imag();
      }
    }
    cd* rowp = g + (size_t)(spin ? lmax + m : m) * nring;
    cd* rowm = g + (size_t)(lmax - m) * nring;
    BlockChain u, v;
    for (int r0 = 0; r0 < nring; r0 += kRB) {
      const int nb = std::min(kRB, nring - r0);
      double x[kRB] = {0};
      for (int i = 0; i < nb; ++i) x[i] = rl.x[r0 + i];
      double gpr[kRB] = {0}, gpi[kRB] = {0}, gmr[kRB] = {0}, gmi[kRB] = {0};
      u.start(cu, rl, r0, nb);
      if (spin) v.start(cv, rl, r0, nb);
      for (int l = cu.l0; l <= lmax; ++l) {
        const double er = epr[l - cu.l0], ei = epi[l - cu.l0];
#pragma omp simd
        for (int i = 0; i < kRB; ++i) {
          const double f = u.c[i] * u.live[i];
          gpr[i] += er * f;
          gpi[i] += ei * f;
        }
        if (spin && m > 0) {
          const double fr = emr[l - cu.l0], fi = emi[l - cu.l0];
#pragma omp simd
          for (int i = 0; i < kRB; ++i) {
            const double f = v.c[i] * v.live[i];
            gmr[i] += fr * f;
            gmi[i] += fi * f;
          }
        }
        if (l < lmax) {
          u.step(cu, l, x);
          if (spin && m > 0) v.step(cv, l, x);
        }
      }
      for (int i = 0; i < nb; ++i) rowp[r0 + i] = cd(gpr[i], gpi[i]);
      if (spin && m > 0)
        for (int i = 0; i < nb; ++i) rowm[r0 + i] = cd(gmr[i], gmi[i]);
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
  const RingLogs rl(theta, nring);
  std::memset(alm, 0, sizeof(cd) * (size_t)nalm * ncomp);
  const double ss = (spin & 1) ? -1.0 : 1.0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads(nthreads))
  for (int m = lmax; m >= 0; --m) {
    ChainCoef cu, cv;
    cu.prepare(lmax, m, -spin);
    if (cu.l0 > lmax) continue;
    if (spin) cv.prepare(lmax, m, spin);
    const int nl = lmax - cu.l0 + 1;
    // X_l = sum_r G'_m f(m' = -s), Y_l = sum_r conj(G'_{-m}) f(m' = +s)
    std::vector<double> xr(nl, 0.0), xi(nl, 0.0), yr(nl, 0.0), yi(nl, 0.0);
    const cd* rowp = g + (size_t)(spin ? lmax + m : m) * nring;
    const cd* rowm = g + (size_t)(lmax - m) * nring;
    BlockChain u, v;
    for (int r0 = 0; r0 < nring; r0 += kRB) {
      const int nb = std::min(kRB, nring - r0);
      double x[kRB] = {0}, pr[kRB] = {0}, pi[kRB] = {0}, qr[kRB] = {0}, qi[kRB] = {0};
      for (int i = 0; i < nb; ++i) {
        x[i] = rl.x[r0 + i];
        pr[i] = rowp[r0 + i].real();
        pi[i] = rowp[r0 + i].imag();
        if (spin) {
          qr[i] = rowm[r0 + i].real();
          qi[i] = -rowm[r0 + i].imag();
        }
      }
      u.start(cu, rl, r0, nb);
      if (spin) v.start(cv, rl, r0, nb);
      for (int l = cu.l0; l <= lmax; ++l) {
        double sr = 0.0, si = 0.0;
#pragma omp simd reduction(+ : sr, si)
        for (int i = 0; i < kRB; ++i) {
          const double f = u.c[i] * u.live[i];
          sr += pr[i] * f;
          si += pi[i] * f;
        }
        xr[l - cu.l0] += sr;
        xi[l - cu.l0] += si;
        if (spin) {
          double tr = 0.0, ti = 0.0;
#pragma omp simd reduction(+ : tr, ti)
          for (int i = 0; i < kRB; ++i) {
            const double f = v.c[i] * v.live[i];
            tr += qr[i] * f;
            ti += qi[i] * f;
          }
          yr[l - cu.l0] += tr;
          yi[l - cu.l0] += ti;
        }
        if (l < lmax) {
          u.step(cu, l, x);
          if (spin) v.step(cv, l, x);
        }
      }
    }
    for (int l = cu.l0; l <= lmax; ++l) {
      const int q = l - cu.l0;
      if (!spin) {
        a[idx_of(lmax, l, m)] = cd(xr[q], m == 0 ? 0.0 : xi[q]);
      } else {
        // E' = -(X + Y) / 2, B' = i (X - Y) / 2 with X = (-1)^s sum G'_m f(m' = -s)
        const cd X = ss * cd(xr[q], xi[q]), Y(yr[q], yi[q]);
        a[idx_of(lmax, l, m)] = -0.5 * (X + Y);
        if (ncomp > 1) a[nalm + idx_of(lmax, l, m)] = cd(0.0, 0.5) * (X - Y);
      }
    }
  }
  return 0;
}

int nusht_cpu_es_kernel_info(int w, double beta, double ktol, int* degree, double* max_err) {
  if (w < 2 || w > kMaxW || !(beta > 0.0) || !(ktol > 0.0)) return 1;
  const EsPoly& k = es_poly(w, beta, ktol);
  if (degree) *degree = k.deg;
  if (max_err) *max_err = k.err;
  return 0;
}

int nusht_cpu_interp(const nusht_cpu_grid* g, const double* grid, int64_t npts, const double* theta, const double* phi,
                     int mshift, const double* gamma, int spin, double* out, int nthreads) {
  if (!grid_ok(g) || npts < 0 || !grid || !theta || !phi || !out) return 1;
  const EsPoly& k = es_poly(g->w, g->beta, g->ktol);
  cd* o = reinterpret_cast<cd*>(out);
#pragma omp parallel for schedule(static) num_threads(threads(nthreads))
  for (int64_t p = 0; p < npts; ++p) {
    Foot f;
    footprint(theta[p], phi[p], *g, k, f);
    o[p] = phase(phi[p], gamma, p, mshift, spin) * interp_point(f, *g, grid);
  }
  return 0;
}

int nusht_cpu_spread(const nusht_cpu_grid* g, const double* in, int64_t npts, const double* theta, const double* phi,
                     int mshift, const double* gamma, int spin, double* grid, int nthreads) {
  if (!grid_ok(g) || npts < 0 || !grid || !theta || !phi || !in) return 1;
  const EsPoly& k = es_poly(g->w, g->beta, g->ktol);
  const cd* v = reinterpret_cast<const cd*>(in);
  const int nth = threads(nthreads);
  const int w = g->w, nrow = g->nrow;
  auto one = [&](int64_t p) {
    Foot f;
    footprint(theta[p], phi[p], *g, k, f);
    spread_point(f, std::conj(phase(phi[p], gamma, p, mshift, spin)) * v[p], *g, grid);
  };
  if (nrow == g->nf1 || nth == 1) {  // wrapping window (small grids) or one thread: in order
    for (int64_t p = 0; p < npts; ++p) one(p);
    return 0;
  }
  // Bands of H >= w window rows, points bucketed by their first row; a point of band b writes rows
  // [b H, (b + 2) H), so the bands of one parity never share a row and run in parallel without atomics.
  const int H = std::max(w, (nrow + 4 * nth - 1) / (4 * nth));
  const int nb = (nrow + H - 1) / H + 1;
  std::vector<int> band(npts);
  std::vector<int64_t> start(nb + 1, 0);
  for (int64_t p = 0; p < npts; ++p) {
    const double ut = theta[p] * (g->nf1 / (2.0 * M_PI));
    const int r0 = (int)std::ceil(ut - 0.5 * w) - g->row_lo;
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
    for (int b = parity; b < nb; b += 2)
      for (int64_t q = start[b]; q < start[b + 1]; ++q) one(order[q]);
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
