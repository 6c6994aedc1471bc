// Spin-weighted Legendre ring transforms on gfx950 (FP64).  See legendre.h for the conventions.
//
// Replaces the reference's SHTns synth / analys / synth_grad calls
// (cunusht/sht/GPU_sht_transformer.py:56-96) with per-m three-term recursions:
//   synthesis : one wave per (m, block of 64 ring pairs); coefficients are wave-uniform
//               and arrive through the scalar cache; the l-loop runs in two phases, first
//               with underflow rescaling while any lane is still below 2^-300, then plain.
//   adjoint   : one workgroup per m holding all ring pairs (4 per lane); per-l sums over
//               lanes are formed with a butterfly transpose-reduce in registers, summed over
//               waves with LDS atomics, then written once.
#include "legendre.h"

#include "timing.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <cmath>

namespace nusht {
namespace {

__host__ __device__ inline int l0_of(int m, int s) { return m > s ? m : s; }

constexpr double kThresh = 0x1p300;   // rescale when a scaled value exceeds this
constexpr double kSmall = 0x1p-600;   // one scale step
constexpr double kSeedLo = 0x1p-300;

// ------------------------------------------------------------------ tables
__global__ void k_rec_tables(int lmax, int s, const long* __restrict__ mofs, double2* __restrict__ rec,
                             double* __restrict__ cnorm) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m > lmax) return;
  int l0 = l0_of(m, s);
  if (l0 > lmax) return;
  long base = mofs[m];
  double cprev = 1.0, ccur = 1.0;
  cnorm[base] = 1.0;
  for (int l = l0; l < lmax; ++l) {
    double alpha, b, gamma;
    if (l == 0) {
      alpha = sqrt(3.0); b = 0.0; gamma = 0.0;
    } else {
      double d1 = sqrt(((l + 1.0) * (l + 1) - (double)m * m) * ((l + 1.0) * (l + 1) - (double)s * s));
      double d0 = sqrt(((double)l * l - (double)m * m) * ((double)l * l - (double)s * s));
      double f = sqrt((2.0 * l + 3) * (2.0 * l + 1));
      alpha = f * (l + 1) / d1;
      b = f * (double)m * s / (l * d1);
      gamma = sqrt((2.0 * l + 3) / (2.0 * l - 1)) * (l + 1) * d0 / (l * d1);
    }
    double cnext = (l == l0) ? 1.0 : gamma * cprev;
    rec[base + (l - l0)] = double2{alpha * ccur / cnext, b * ccur / cnext};
    cnorm[base + (l - l0) + 1] = cnext;
    cprev = ccur;
    ccur = cnext;
  }
}

__device__ inline void renorm(double& v, double& k) {
  while (v != 0.0 && fabs(v) < kSeedLo) { v *= 0x1p600; k -= 1.0; }
}

// Starting values at l0 = max(m, s): u = N_l0 d^l0_{m,-s}, v = N_l0 d^l0_{m,+s}, with scale exponents.
__global__ __launch_bounds__(64) void k_seeds(int lmax, int s, int nslot, int sstride,
                                              const double* __restrict__ slot_theta, double2* __restrict__ st_mu,
                                              double2* __restrict__ st_p, double2* __restrict__ st_k) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslot) return;
  const double th = slot_theta[i];
  const double sh = sin(0.5 * th), ch = cos(0.5 * th);
  const double ns = sqrt((2.0 * s + 1) / (4.0 * M_PI));
  double u = 0, v = 0, ku = 0, kv = 0;
  // m < s: closed form at l0 = s
  for (int m = 0; m < s && m <= lmax; ++m) {
    double lg = 0.5 * (lgamma(2.0 * s + 1) - lgamma(s + m + 1.0) - lgamma(s - m + 1.0));
    double bin = exp(lg);
    double uu = ns * bin * pow(ch, (double)(s - m)) * pow(sh, (double)(s + m));
    if ((m + s) & 1) uu = -uu;
    double vv = ns * bin * pow(ch, (double)(s + m)) * pow(sh, (double)(s - m));
    const long q = (long)m * sstride + i;
    st_mu[q] = double2{uu, vv};
    st_p[q] = double2{0.0, 0.0};
    st_k[q] = double2{0.0, 0.0};
  }
  if (s > lmax) return;
  // m = s: u = N_s sin^{2s}(th/2), v = N_s cos^{2s}(th/2); then the product recursion in m.
  u = ns; v = ns;
  for (int j = 0; j < 2 * s; ++j) { u *= sh; v *= ch; }
  renorm(u, ku);
  renorm(v, kv);
  {
    const long q = (long)s * sstride + i;
    st_mu[q] = double2{u, v};
    st_p[q] = double2{0.0, 0.0};
    st_k[q] = double2{ku, kv};
  }
  for (int m = s + 1; m <= lmax; ++m) {
    double f = -sqrt(2.0 * m * (2.0 * m + 1) / ((double)m * m - (double)s * s)) * sh * ch;
    u *= f;
    v *= f;
    renorm(u, ku);
    renorm(v, kv);
    const long q = (long)m * sstride + i;
    st_mu[q] = double2{u, v};
    st_p[q] = double2{0.0, 0.0};
    st_k[q] = double2{ku, kv};
  }
}

// Advance each (m, 64-slot group) through the l-range in which every value of the group is
// below 2^-300, in steps of kStartAlign, and store the state at the first l that may contribute.
// Scale exponents never decrease, so a group that is still all-under after a step of kStartAlign
// was all-under at every l inside it; otherwise the step is undone.
constexpr int kStartAlign = 8;

template <bool SP>
__global__ __launch_bounds__(64) void k_advance(int lmax, int s, int nslot, int sstride, int ngroups,
                                                const double* __restrict__ slot_theta, const long* __restrict__ mofs,
                                                const double2* __restrict__ rec_all, double2* __restrict__ st_mu,
                                                double2* __restrict__ st_p, double2* __restrict__ st_k,
                                                int* __restrict__ start) {
  const int m = blockIdx.x / ngroups;
  const int g = blockIdx.x - m * ngroups;
  const int l0 = l0_of(m, s);
  if (l0 > lmax) return;
  const int slot = g * 64 + (int)threadIdx.x;
  const bool valid = slot < nslot;
  const long q = (long)m * sstride + slot;
  double x = 0, mu = 0, mv = 0, pu = 0, pv = 0, ku = 0, kv = 0;
  if (valid) {
    x = cos(slot_theta[slot]);
    const double2 a = st_mu[q], k = st_k[q];
    mu = a.x; ku = k.x;
    if (SP) { mv = a.y; kv = k.y; }
  }
  auto dead = [&]() { return !__any(valid && (ku == 0.0 || (SP && kv == 0.0))); };
  const double2* __restrict__ rec = rec_all + mofs[m];
  int lc = l0;
  while (lc + kStartAlign - 1 <= lmax && dead()) {
    const double smu = mu, smv = mv, spu = pu, spv = pv, sku = ku, skv = kv;
#pragma unroll
    for (int i = 0; i < kStartAlign; ++i) {
      const double2 rc = rec[lc - l0 + i];
      const double t = rc.x * x;
      const double nu = fma(t + rc.y, mu, -pu);
      pu = mu; mu = nu;
      if (SP) {
        const double nv = fma(t - rc.y, mv, -pv);
        pv = mv; mv = nv;
      }
      if (ku < 0.0 && fabs(mu) > kThresh) { mu *= kSmall; pu *= kSmall; ku += 1.0; }
      if (SP && kv < 0.0 && fabs(mv) > kThresh) { mv *= kSmall; pv *= kSmall; kv += 1.0; }
    }
    if (dead()) { lc += kStartAlign; continue; }
    mu = smu; mv = smv; pu = spu; pv = spv; ku = sku; kv = skv;
    break;
  }
  if (valid && lc > l0) {
    st_mu[q] = double2{mu, mv};
    st_p[q] = double2{pu, pv};
    st_k[q] = double2{ku, kv};
  }
  if (threadIdx.x == 0) start[(long)m * ngroups + g] = lc;
}

// ------------------------------------------------------------------ synthesis prep
// coef layout per map b, (m, l):  spin 0: coef[(mofs[m]+l-l0)*NB + b] = a_lm c_l fl_l
//                                 spin s: coef[((mofs[m]+l-l0)*NB + b)*2 + {0,1}] = (W', Wt')
//                                 spin s, E only (ncomp 1): coef[(mofs[m]+l-l0)*NB + b] = W'
//   W' = -(-1)^s c_l (E + iB),  Wt' = -c_l (conj E + i conj B);  with B = 0, Wt' = (-1)^s conj(W').
__global__ void k_prep(int lmax, int s, int nmaps, int nbpad, int ncomp, const double2* __restrict__ alm, long stride_map,
                       const double* __restrict__ fl, const long* __restrict__ mofs,
                       const double* __restrict__ cnorm, double2* __restrict__ coef) {
  const int m = blockIdx.y;
  const int l0 = l0_of(m, s);
  if (l0 > lmax) return;
  const long nalm = (long)(lmax + 1) * (lmax + 2) / 2;
  const long mbase = (long)m * (2 * lmax + 1 - m) / 2;
  const double sg = (s & 1) ? 1.0 : -1.0;  // -(-1)^s
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= lmax - l0; j += gridDim.x * blockDim.x) {
    const int l = l0 + j;
    const long q = mofs[m] + j;
    double c = cnorm[q];
    if (fl) c *= fl[l];
    for (int b = 0; b < nbpad; ++b) {
      double2 e{0.0, 0.0}, bb{0.0, 0.0};
      if (b < nmaps) {
        const double2* ab = alm + b * stride_map;
        e = ab[mbase + l];
        if (s != 0 && ncomp > 1) bb = ab[nalm + mbase + l];
      }
      if (s == 0) {
        coef[q * nbpad + b] = double2{e.x * c, e.y * c};
      } else if (ncomp == 1) {
        coef[q * nbpad + b] = double2{sg * c * e.x, sg * c * e.y};
      } else {
        double2 w{e.x - bb.y, e.y + bb.x};
        double2 wt{e.x + bb.y, -e.y + bb.x};
        coef[(q * nbpad + b) * 2] = double2{sg * c * w.x, sg * c * w.y};
        coef[(q * nbpad + b) * 2 + 1] = double2{-c * wt.x, -c * wt.y};
      }
    }
  }
}

// ------------------------------------------------------------------ synthesis kernel
template <int NB>
struct SynAcc0 {
  double2 a[2][NB];
};

__device__ __forceinline__ void fma2(double2& acc, double2 c, double w) {
  acc.x = fma(c.x, w, acc.x);
  acc.y = fma(c.y, w, acc.y);
}

struct SynthArgs {
  int lmax, s, nslot, sstride, nsb, nmaps;
  const double* slot_theta;
  const int* slot_north;
  const int* slot_south;
  const long* mofs;
  const double2* rec;
  const double2* st_mu;
  const double2* st_p;
  const double2* st_k;
  const int* start;
  const double2* coef;
  double2* G;
  long sm, sr, smap;
};

// One wave per (m, group of 64 ring pairs), starting at the group's first contributing l.
// The l-loop runs in chunks of CH so that the wave-uniform coefficients of a chunk are fetched
// together (one scalar-cache round trip per chunk instead of per l).  Tables are padded by CH
// entries so a chunk never reads past the end.
constexpr int kSynCH = 8;  // largest chunk of any variant (table padding)

// GR (spin s, E only): the coefficient stream holds W' alone.  Since Wt' = (-1)^s conj(W') and the
// recursion values are real, the -m sums are conjugates of the +m ones: the kernel accumulates
// sum W' mu_u and sum W' mu_v split by the parity of l - l0 (half the contraction of the E + B case) and
// forms the four ring coefficients at the end.
template <bool SP, int NB, int CH, int WPB, bool GR = false>
__global__ __launch_bounds__(64 * WPB) void k_synth(SynthArgs a) {
  static_assert(SP || !GR, "the gradient-only contraction is a spin-s case");
  constexpr int R = 1;
  constexpr int K = (SP && !GR) ? 2 * NB : NB;
  // Blocks are dealt to the XCDs round-robin (block b runs on XCD b % 8).  XCD x takes
  // m = x, x + 8, x + 16, ... with all blocks of one m in a row, so each m's coefficient stream
  // is read through one XCD's L2, and every XCD still starts with its longest m.  The WPB waves
  // of a block take WPB adjacent slot groups of the same m and read the same coefficients, so
  // all but the leading wave find them in the CU's scalar cache.
  const int blk = blockIdx.x;
  const int j = blk / kNumXcd;
  const int m = (blk - j * kNumXcd) + kNumXcd * (j / a.nsb);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int sb = (j % a.nsb) * WPB + wave;
  const int L = a.lmax;
  if (m > L) return;
  const int l0 = l0_of(m, a.s);
  // the launch may carry empty slot groups (odd block count per m, see launch_synth)
  if (l0 > L || sb * 64 >= a.nslot) return;
  const int lane = threadIdx.x & 63;
  const int lstart = a.start[(long)m * ((a.nslot + 63) >> 6) + sb];

  double x[R], mu_u[R], mu_v[R], pu[R], pv[R], ku[R], kv[R];
  int slot[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    slot[r] = sb * 64 + lane;
    x[r] = mu_u[r] = mu_v[r] = pu[r] = pv[r] = ku[r] = kv[r] = 0.0;
    if (slot[r] < a.nslot) {
      const long q = (long)m * a.sstride + slot[r];
      x[r] = cos(a.slot_theta[slot[r]]);
      const double2 sm = a.st_mu[q], sp = a.st_p[q], sk = a.st_k[q];
      mu_u[r] = sm.x; pu[r] = sp.x; ku[r] = sk.x;
      if (SP) { mu_v[r] = sm.y; pv[r] = sp.y; kv[r] = sk.y; }
    }
  }

  const long base = a.mofs[m];
  const double2* __restrict__ rec = a.rec + base;
  const double2* __restrict__ cf = a.coef + base * K;

  // spin 0: A0/A1 = parity-split sums; spin s: PN (+m north), MN (-m north), A (+m south), MS (-m south)
  double2 A0[R][NB], A1[R][NB], PN[R][NB], MN[R][NB], MS0[R][NB], MS1[R][NB];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      A0[r][b] = A1[r][b] = PN[r][b] = MN[r][b] = MS0[r][b] = MS1[r][b] = double2{0.0, 0.0};
    }

  struct Buf {
    double2 rc[CH];
    double2 cw[CH][K];
  };
  // Wave-uniform coefficients of one chunk of CH consecutive l (scalar loads).  The loop below
  // issues the loads of chunk c + 1 before it computes chunk c; the tables are padded by 2 CH
  // entries for the prefetch past the last chunk.
  auto load = [&](int lc, Buf& b) {
    const int j0 = lc - l0;
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      b.rc[i] = rec[j0 + i];
#pragma unroll
      for (int k = 0; k < K; ++k) b.cw[i][k] = cf[(long)(j0 + i) * K + k];
    }
  };
  // One chunk (the first l at even l - l0).  SLOW: some lane may still be below the
  // representable range (masked contributions, rescaling); FULL: all CH l <= L.
  auto chunk = [&](int lc, const Buf& bf, auto slow_tag, auto full_tag) {
    constexpr bool SLOW = decltype(slow_tag)::value;
    constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      if (!FULL && lc + i > L) break;
      const bool odd = (i & 1) != 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        double wu = mu_u[r], wv = mu_v[r];
        if (SLOW) { wu = ku[r] == 0.0 ? wu : 0.0; wv = kv[r] == 0.0 ? wv : 0.0; }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          if (!SP) {
            if (odd) fma2(A1[r][b], bf.cw[i][b], wu); else fma2(A0[r][b], bf.cw[i][b], wu);
          } else if (GR) {
            // MS0 / MS1 hold sum W' mu_u by parity here
            const double2 w = bf.cw[i][b];
            if (odd) { fma2(MS1[r][b], w, wu); fma2(A1[r][b], w, wv); }
            else { fma2(MS0[r][b], w, wu); fma2(A0[r][b], w, wv); }
          } else {
            const double2 w = bf.cw[i][2 * b], wt = bf.cw[i][2 * b + 1];
            fma2(PN[r][b], w, wu);
            fma2(MN[r][b], wt, wv);
            if (odd) { fma2(A1[r][b], w, wv); fma2(MS1[r][b], wt, wu); }
            else { fma2(A0[r][b], w, wv); fma2(MS0[r][b], wt, wu); }
          }
        }
      }
      // the step past l = lmax reads a padded table entry; its result is never used
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double t = bf.rc[i].x * x[r];
        const double nu = fma(t + bf.rc[i].y, mu_u[r], -pu[r]);
        pu[r] = mu_u[r]; mu_u[r] = nu;
        if (SP) {
          const double nv = fma(t - bf.rc[i].y, mu_v[r], -pv[r]);
          pv[r] = mu_v[r]; mu_v[r] = nv;
        }
        if (SLOW) {
          if (ku[r] < 0.0 && fabs(mu_u[r]) > kThresh) { mu_u[r] *= kSmall; pu[r] *= kSmall; ku[r] += 1.0; }
          if (SP && kv[r] < 0.0 && fabs(mu_v[r]) > kThresh) { mu_v[r] *= kSmall; pv[r] *= kSmall; kv[r] += 1.0; }
        }
      }
    }
  };
  auto any_under = [&]() {
    bool under = false;
#pragma unroll
    for (int r = 0; r < R; ++r) under |= (ku[r] < 0.0) || (kv[r] < 0.0);
    return __any(under);
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  int lc = lstart;
  Buf cur, nxt;
  load(lc, cur);
  // slow phase: while some lane of the wave is still below 2^-300
  while (lc <= L && any_under()) {
    load(lc + CH, nxt);
    if (lc + CH - 1 <= L) chunk(lc, cur, T_{}, T_{});
    else chunk(lc, cur, T_{}, F_{});
    cur = nxt;
    lc += CH;
  }
  // fast phase, two chunks per iteration so that the two buffers alternate without copies (the loads reach at
  // most CH entries past lmax, as in the single-chunk step)
  while (lc + 2 * CH - 1 <= L) {
    load(lc + CH, nxt);
    chunk(lc, cur, F_{}, T_{});
    load(lc + 2 * CH, cur);
    chunk(lc + CH, nxt, F_{}, T_{});
    lc += 2 * CH;
  }
  if (lc + CH - 1 <= L) {
    load(lc + CH, nxt);
    chunk(lc, cur, F_{}, T_{});
    cur = nxt;
    lc += CH;
  }
  if (lc <= L) chunk(lc, cur, F_{}, F_{});

  const double sg0 = ((l0 + m) & 1) ? -1.0 : 1.0;  // (-1)^{l+m} for even (l - l0)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (slot[r] >= a.nslot) continue;
    const int rn = a.slot_north[slot[r]];
    const int rs = a.slot_south[slot[r]];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b >= a.nmaps) break;
      double2* Gb = a.G + b * a.smap;
      if (!SP) {
        double2 gn = cadd(A0[r][b], A1[r][b]);
        double2 gs = double2{sg0 * (A0[r][b].x - A1[r][b].x), sg0 * (A0[r][b].y - A1[r][b].y)};
        Gb[(long)m * a.sm + (long)rn * a.sr] = gn;
        if (rs >= 0) Gb[(long)m * a.sm + (long)rs * a.sr] = gs;
      } else if (GR) {
        // north +m: sum W' mu_u; north -m: (-1)^s conj(sum W' mu_v); south +m: sg0 (A0 - A1);
        // south -m: sg0 (-1)^s conj(even - odd of sum W' mu_u)
        const double ss = (a.s & 1) ? -1.0 : 1.0;
        const double2 pn = cadd(MS0[r][b], MS1[r][b]);
        const double2 av = cadd(A0[r][b], A1[r][b]);
        const double2 pS = double2{sg0 * (A0[r][b].x - A1[r][b].x), sg0 * (A0[r][b].y - A1[r][b].y)};
        const double2 mN = double2{ss * av.x, -ss * av.y};
        const double2 mS = double2{sg0 * ss * (MS0[r][b].x - MS1[r][b].x), -sg0 * ss * (MS0[r][b].y - MS1[r][b].y)};
        const long rowp = (long)(L + m) * a.sm, rowm = (long)(L - m) * a.sm;
        Gb[rowp + (long)rn * a.sr] = pn;
        if (rs >= 0) Gb[rowp + (long)rs * a.sr] = pS;
        if (m > 0) {
          Gb[rowm + (long)rn * a.sr] = mN;
          if (rs >= 0) Gb[rowm + (long)rs * a.sr] = mS;
        }
      } else {
        double2 pS = double2{sg0 * (A0[r][b].x - A1[r][b].x), sg0 * (A0[r][b].y - A1[r][b].y)};
        double2 mS = double2{sg0 * (MS0[r][b].x - MS1[r][b].x), sg0 * (MS0[r][b].y - MS1[r][b].y)};
        const long rowp = (long)(L + m) * a.sm, rowm = (long)(L - m) * a.sm;
        Gb[rowp + (long)rn * a.sr] = PN[r][b];
        if (rs >= 0) Gb[rowp + (long)rs * a.sr] = pS;
        if (m > 0) {
          Gb[rowm + (long)rn * a.sr] = MN[r][b];
          if (rs >= 0) Gb[rowm + (long)rs * a.sr] = mS;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ synthesis on the matrix cores
// For 4 maps per launch the contraction over l is a GEMM per m: per chunk of 4 l,
// (64 ring pairs x 4 l) x (4 l x 16 columns), done with v_mfma_f64_16x16x4_f64 while the VALU
// runs the recursion.  Columns, 4 per map b, with s_l = (-1)^{l+m}:
//   spin s, u chain: Re W_b, Im W_b, Re s_l Wt_b, Im s_l Wt_b    (north +m, south -m)
//           v chain: Re Wt_b, Im Wt_b, Re s_l W_b, Im s_l W_b    (north -m, south +m)
//   spin 0, u chain: Re a_b, Im a_b, Re s_l a_b, Im s_l a_b      (north, south)
// A launch takes 4 NS maps (NS = 1 or 2 groups of 16 columns; the recursion and the A tiles serve both groups).
// The table holds, per m and per block q of 4 l (l = l0 + 4q + k), the values once: spin s
// [k][b][Re W, Im W, Re Wt, Im Wt] (64 NS doubles), spin 0 [k][b][Re a, Im a] (32 NS), b < 4 NS.  Lane j + 16k of
// group g's B operand (column j = 4b' + c of map b = 4g + b' at l0 + 4q + k) reads entry (k, b, c) for the u
// chain and (k, b, c ^ 2) for the v chain (spin s), (k, b, c & 1) for spin 0, and negates it when c >= 2 and
// l + m is odd; blocks start at l0 + 4q, so that sign is fixed per lane.  Entries past lmax are zero.

// Operand layouts of v_mfma_f64_16x16x4_f64 on gfx950 (64 lanes):
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in register i / 4 of lane j + 16 (i % 4).

// One thread per (block q of 4 l, k, map b): map b's entry at l = l0 + 4q + k, one 32-byte store (16 for spin 0;
// consecutive threads write consecutive entries).  Block row y serves m = y and m = lmax - y, whose tables together
// have about lmax + 2 rows of l, so no block row is idle.
template <int NS>
__global__ void k_prep4(int lmax, int s, int nmaps, int ncomp, AlmPtrs8 alm,
                        const double* __restrict__ fl, const long* __restrict__ mofs, const long* __restrict__ mblk,
                        const double* __restrict__ cnorm, double* __restrict__ coef) {
  const int nc = s ? 2 : 1;
  const long nalm = (long)(lmax + 1) * (lmax + 2) / 2;
  const double sg = (s & 1) ? 1.0 : -1.0;  // -(-1)^s
  for (int pass = 0; pass < 2; ++pass) {
    const int m = pass == 0 ? (int)blockIdx.y : lmax - (int)blockIdx.y;
    if (pass == 1 && m <= (int)blockIdx.y) break;  // the middle m has no partner
    const int l0 = l0_of(m, s);
    if (l0 > lmax) continue;
    const int nl = lmax - l0 + 1;
    constexpr int NM = 4 * NS, LNM = NS == 1 ? 2 : 3;  // maps per launch and log2
    const long nthr = (long)((nl + 3) >> 2) * 4 * NM;
    const long mbase = (long)m * (2 * lmax + 1 - m) / 2;
    double* out = coef + mblk[m] * nc * 32 * NS;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nthr; t += (long)gridDim.x * blockDim.x) {
      const int b = (int)(t & (NM - 1)), k = (int)((t >> LNM) & 3);
      const long q = t >> (LNM + 2);
      const int jj = (int)(4 * q + k);
      const int l = l0 + jj;
      double2 w{0.0, 0.0}, wt{0.0, 0.0};  // spin 0: w = a_lm c_l
      if (l <= lmax && b < nmaps) {
        double cl = cnorm[mofs[m] + jj];
        if (fl) cl *= fl[l];
        const double2* ab = alm.p[0];  // selects, not an indexed read of the argument array
#pragma unroll
        for (int u = 1; u < NM; ++u) ab = b == u ? alm.p[u] : ab;
        const double2 ev = ab[mbase + l];
        if (s == 0) {
          w = double2{cl * ev.x, cl * ev.y};
        } else {
          const double2 bv = ncomp > 1 ? ab[nalm + mbase + l] : double2{0.0, 0.0};
          w = double2{sg * cl * (ev.x - bv.y), sg * cl * (ev.y + bv.x)};
          wt = double2{-cl * (ev.x + bv.y), -cl * (-ev.y + bv.x)};
        }
      }
      double2* o = reinterpret_cast<double2*>(out + q * nc * 32 * NS + (k * NM + b) * nc * 2);
      o[0] = w;
      if (s != 0) o[1] = wt;
    }
  }
}

typedef double v4d_t __attribute__((ext_vector_type(4)));

// 4 x 4 transpose of (register k, row g of 16 lanes): afterwards register g holds in lane i + 16k
// what register k held in lane 16g + i (the A-operand tile of ring pairs 16g .. 16g + 15).
__device__ __forceinline__ void transpose_rows(double (&u)[4]) {
  unsigned lo[4], hi[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lo[k] = (unsigned)__double2loint(u[k]);
    hi[k] = (unsigned)__double2hiint(u[k]);
  }
  auto sw32 = [](unsigned& x, unsigned& y) {
    auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);  // x upper half <-> y lower half
    x = r[0]; y = r[1];
  };
  auto sw16 = [](unsigned& x, unsigned& y) {
    auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);  // x odd rows <-> y even rows
    x = r[0]; y = r[1];
  };
  sw32(lo[0], lo[2]); sw32(lo[1], lo[3]); sw16(lo[0], lo[1]); sw16(lo[2], lo[3]);
  sw32(hi[0], hi[2]); sw32(hi[1], hi[3]); sw16(hi[0], hi[1]); sw16(hi[2], hi[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) u[k] = __hiloint2double((int)hi[k], (int)lo[k]);
}

template <bool SP, int NS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NS == 1 ? 3 : 2))) void k_synth_mfma(
    SynthArgs a, const double* __restrict__ coef4, const long* __restrict__ mblk) {
  constexpr int NC = SP ? 2 : 1;
  // same block -> (m, slot group) mapping as k_synth with one wave per block
  const int blk = blockIdx.x;
  const int j = blk / kNumXcd;
  const int m = (blk - j * kNumXcd) + kNumXcd * (j / a.nsb);
  const int sb = j % a.nsb;
  const int L = a.lmax;
  if (m > L) return;
  const int l0 = l0_of(m, a.s);
  if (l0 > L || sb * 64 >= a.nslot) return;
  const int lane = threadIdx.x & 63;
  const int lstart = a.start[(long)m * ((a.nslot + 63) >> 6) + sb];

  const int slot = sb * 64 + lane;
  double x = 0, mu_u = 0, mu_v = 0, pu = 0, pv = 0, ku = 0, kv = 0;
  if (slot < a.nslot) {
    const long q = (long)m * a.sstride + slot;
    x = cos(a.slot_theta[slot]);
    const double2 sm = a.st_mu[q], sp = a.st_p[q], sk = a.st_k[q];
    mu_u = sm.x; pu = sp.x; ku = sk.x;
    if (SP) { mu_v = sm.y; pv = sp.y; kv = sk.y; }
  }
  const double2* __restrict__ rec = a.rec + a.mofs[m];
  const double* __restrict__ cb = coef4 + mblk[m] * NC * 32 * NS;
  // this lane's B entries in a block for group 0 (u chain, v chain; group g adds 16 g, 8 g for spin 0) and their
  // sign bit: column c = lane & 3 >= 2 at odd l + m, l = l0 + 4q + (lane >> 4)
  const int c4 = lane & 3;
  const int eu = SP ? (lane >> 4) * 16 * NS + (lane & 15) : (lane >> 4) * 8 * NS + ((lane & 15) >> 2) * 2 + (c4 & 1);
  const int ev = eu ^ 2;
  const unsigned sgnb = (c4 >= 2 && ((l0 + (lane >> 4) + m) & 1)) ? 0x80000000u : 0u;
  auto sgn = [&](double v) {
    return __hiloint2double(__double2hiint(v) ^ (int)sgnb, __double2loint(v));
  };

  v4d_t Du[NS][4], Dv[NS][4];
#pragma unroll
  for (int h = 0; h < NS; ++h)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      Du[h][g] = v4d_t{0.0, 0.0, 0.0, 0.0};
      Dv[h][g] = v4d_t{0.0, 0.0, 0.0, 0.0};
    }
  struct Buf {
    double2 rc[4];
    double bu[NS], bv[NS];
  };
  auto load = [&](int lc, Buf& b) {
    const int jl = lc - l0;
#pragma unroll
    for (int i = 0; i < 4; ++i) b.rc[i] = rec[jl + 1 + i];  // the steps from l + 1 (state one step ahead)
    const double* blkp = cb + (long)(jl >> 2) * NC * 32 * NS;
#pragma unroll
    for (int h = 0; h < NS; ++h) {
      b.bu[h] = blkp[eu + h * (SP ? 16 : 8)];
      if (SP) b.bv[h] = blkp[ev + h * 16];
    }
  };
  // GRP: the map groups a chunk accumulates (-1 all, else that group alone)
  auto chunk = [&](int lc, const Buf& bf, auto slow_tag, auto full_tag, auto grp_tag) {
    constexpr bool SLOW = decltype(slow_tag)::value;
    constexpr bool FULL = decltype(full_tag)::value;
    constexpr int GRP = decltype(grp_tag)::value;
    double U[4], V[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (!FULL && lc + i > L) { U[i] = 0.0; V[i] = 0.0; continue; }  // zero weight past lmax
      double wu = pu, wv = pv;  // the value at lc + i
      if (SLOW) { wu = ku == 0.0 ? wu : 0.0; wv = kv == 0.0 ? wv : 0.0; }
      U[i] = wu; V[i] = wv;
      const double t = bf.rc[i].x * x;
      const double nu = fma(t + bf.rc[i].y, mu_u, -pu);
      pu = mu_u; mu_u = nu;
      if (SP) {
        const double nv = fma(t - bf.rc[i].y, mu_v, -pv);
        pv = mu_v; mu_v = nv;
      }
      if (SLOW) {
        if (ku < 0.0 && fabs(mu_u) > kThresh) { mu_u *= kSmall; pu *= kSmall; ku += 1.0; }
        if (SP && kv < 0.0 && fabs(mu_v) > kThresh) { mu_v *= kSmall; pv *= kSmall; kv += 1.0; }
      }
    }
    transpose_rows(U);
#pragma unroll
    for (int h = 0; h < NS; ++h) {
      if (GRP >= 0 && h != GRP) continue;
      const double bu = sgn(bf.bu[h]);
#pragma unroll
      for (int g = 0; g < 4; ++g) Du[h][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(U[g], bu, Du[h][g], 0, 0, 0);
    }
    if (SP) {
      transpose_rows(V);
#pragma unroll
      for (int h = 0; h < NS; ++h) {
        if (GRP >= 0 && h != GRP) continue;
        const double bv = sgn(bf.bv[h]);
#pragma unroll
        for (int g = 0; g < 4; ++g) Dv[h][g] = __builtin_amdgcn_mfma_f64_16x16x4f64(V[g], bv, Dv[h][g], 0, 0, 0);
      }
    }
  };
  auto any_under = [&]() { return __any((ku < 0.0) || (kv < 0.0)); };
  using T_ = std::true_type;
  using F_ = std::false_type;
  // The l-loop runs one recursion step ahead: at the first l of a chunk it holds (mu, p) = (mu_{l+1}, mu_l), and
  // the chunk's four A values are p before each step, so none of them is still part of the recursion state when
  // the transposes overwrite them in place (no register copies).  The step to l + 1 is made here; a step past
  // lmax reads a padded or neighbouring table entry and its result is never used.
  {
    const double2 rc0 = rec[lstart - l0];
    const double t = rc0.x * x;
    const double nu = fma(t + rc0.y, mu_u, -pu);
    pu = mu_u; mu_u = nu;
    if (SP) {
      const double nv = fma(t - rc0.y, mu_v, -pv);
      pv = mu_v; mu_v = nv;
    }
    if (ku < 0.0 && fabs(mu_u) > kThresh) { mu_u *= kSmall; pu *= kSmall; ku += 1.0; }
    if (SP && kv < 0.0 && fabs(mu_v) > kThresh) { mu_v *= kSmall; pv *= kSmall; kv += 1.0; }
  }
  using GA_ = std::integral_constant<int, -1>;
  int lc = lstart;
  Buf cur, nxt;
  load(lc, cur);
  // slow phase (some lane below 2^-300); the next chunk's coefficients are loaded before the current chunk is
  // computed.  With two map groups it runs once per group, the recursion restarting from its saved state for the
  // second: a group's accumulators are then either still zero or untouched in the loop, and the loop keeps only
  // the other group's in registers (both together with the rescaling state exceed the two-wave register budget).
  // The second pass runs the same number of chunks, since the phase's end depends on the recursion alone.
  auto slow_phase = [&](auto grp_tag, int lend) {
    while (lend < 0 ? (lc <= L && any_under()) : lc < lend) {
      load(lc + 4, nxt);
      if (lc + 3 <= L) chunk(lc, cur, T_{}, T_{}, grp_tag);
      else chunk(lc, cur, T_{}, F_{}, grp_tag);
      cur = nxt;
      lc += 4;
    }
  };
  if constexpr (NS == 1) {
    slow_phase(GA_{}, -1);
  } else {
    const double s_mu_u = mu_u, s_pu = pu, s_mu_v = mu_v, s_pv = pv, s_ku = ku, s_kv = kv;
    slow_phase(std::integral_constant<int, 0>{}, -1);
    const int lend = lc;
    if (lend > lstart) {
      mu_u = s_mu_u; pu = s_pu; mu_v = s_mu_v; pv = s_pv; ku = s_ku; kv = s_kv;
      lc = lstart;
      load(lc, cur);
      slow_phase(std::integral_constant<int, 1>{}, lend);
    }
  }
  // two chunks per iteration, so the two buffers alternate without copies (loads reach at most one 4-l block
  // past lmax, as the single-chunk step does)
  while (lc + 7 <= L) {
    load(lc + 4, nxt);
    chunk(lc, cur, F_{}, T_{}, GA_{});
    load(lc + 8, cur);
    chunk(lc + 4, nxt, F_{}, T_{}, GA_{});
    lc += 8;
  }
  if (lc + 3 <= L) {
    load(lc + 4, nxt);
    chunk(lc, cur, F_{}, T_{}, GA_{});
    cur = nxt;
    lc += 4;
  }
  if (lc <= L) chunk(lc, cur, F_{}, F_{}, GA_{});

  // epilogue: lane = jc + 16 qq holds column jc of group h (map 4 h + jc / 4) for ring pairs 16 g + 4 r + qq
  // (register r)
  const int jc = lane & 15, qq = lane >> 4;
  const int c = jc & 3;
  const int part = c & 1;
#pragma unroll
  for (int h = 0; h < NS; ++h) {
    const int b = 4 * h + (jc >> 2);
    if (b >= a.nmaps) continue;
    double* Gd = reinterpret_cast<double*>(a.G + b * a.smap);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int sl = sb * 64 + 16 * g + 4 * r + qq;
        if (sl >= a.nslot) continue;
        const int rn = a.slot_north[sl];
        const int rs = a.slot_south[sl];
        const double du = Du[h][g][r];
        if (!SP) {
          if (c < 2) Gd[((long)m * a.sm + (long)rn * a.sr) * 2 + part] = du;
          else if (rs >= 0) Gd[((long)m * a.sm + (long)rs * a.sr) * 2 + part] = du;
        } else {
          const double dv = Dv[h][g][r];
          const long rowp = (long)(L + m) * a.sm, rowm = (long)(L - m) * a.sm;
          if (c < 2) {
            Gd[(rowp + (long)rn * a.sr) * 2 + part] = du;                        // north +m
            if (m > 0) Gd[(rowm + (long)rn * a.sr) * 2 + part] = dv;            // north -m
          } else if (rs >= 0) {
            Gd[(rowp + (long)rs * a.sr) * 2 + part] = dv;                        // south +m
            if (m > 0) Gd[(rowm + (long)rs * a.sr) * 2 + part] = du;            // south -m
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ adjoint kernel
template <bool SP> struct AdjCfg {
  static constexpr int R = SP ? 3 : 4;          // ring pairs per lane per pass
  static constexpr int MaxThreads = 768;
};

// Sum of K complex values per lane over the 64 lanes of a wave (butterfly transpose-reduce).
// Each step exchanges half of the remaining values with lane ^ off and keeps the other half,
// so after log2(K) steps every lane holds one fully lane-summed element; the remaining steps
// are plain all-reduce butterflies.  Lane ln ends up with element reduce_index<K>(ln).
// The exchanges are VALU lane moves, not LDS permutes: lane ^ 32 and lane ^ 16 by the gfx950 permlane
// swaps (v_permlane32_swap: lanes 32..63 of the first operand with lanes 0..31 of the second;
// v_permlane16_swap: the odd 16-lane rows of the first with the even rows of the second; with both operands
// the value itself, the lane's partner value is in one of the two results), lane ^ 8 by a DPP rotation by 8
// inside each 16-lane row.  The closing all-reduce over the lanes that hold one element (16 for K = 4, 8 for
// K = 8) adds the row rotation by 8 (16 lanes only), the half-row mirror and the quad swaps 2 and 1.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
constexpr int kDppRowRor8 = 0x128, kDppRowHalfMirror = 0x141, kDppQuadXor2 = 0x4E, kDppQuadXor1 = 0xB1;

// the value of lane ^ OFF
template <int OFF>
__device__ __forceinline__ double lane_xor(double v) {
  if constexpr (OFF == 8) {
    return dpp_mov<kDppRowRor8>(v);
  } else {
    static_assert(OFF == 16 || OFF == 32, "lane exchanges at offsets 32, 16 and 8");
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const bool first = ((threadIdx.x & 63) & OFF) == 0;
    if constexpr (OFF == 32) {
      const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
      const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
      return __hiloint2double((int)(first ? rh[1] : rh[0]), (int)(first ? rl[1] : rl[0]));
    } else {
      const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
      const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
      return __hiloint2double((int)(first ? rh[1] : rh[0]), (int)(first ? rl[1] : rl[0]));
    }
  }
}

// a + (a of lane ^ OFF) in the lanes with (lane & OFF) == 0, b + (b of lane ^ OFF) in the others, for OFF = 32
// or 16: one swap of the register pair per half does it.  permlane32_swap(x, y) leaves x = (own x | y of lane
// - 32) and y = (x of lane + 32 | own y) over the two halves, permlane16_swap the same over the row pairs, so
// x + y is the wanted sum in every lane.
template <int OFF>
__device__ __forceinline__ double pair_sum(double a, double b) {
  const unsigned alo = (unsigned)__double2loint(a), ahi = (unsigned)__double2hiint(a);
  const unsigned blo = (unsigned)__double2loint(b), bhi = (unsigned)__double2hiint(b);
  static_assert(OFF == 16 || OFF == 32, "pair sums at offsets 32 and 16");
  if constexpr (OFF == 32) {
    const auto rl = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    return __hiloint2double((int)rh[0], (int)rl[0]) + __hiloint2double((int)rh[1], (int)rl[1]);
  } else {
    const auto rl = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    return __hiloint2double((int)rh[0], (int)rl[0]) + __hiloint2double((int)rh[1], (int)rl[1]);
  }
}

template <int K, int k, int off>
__device__ __forceinline__ void tr_step(double2 (&v)[K], int ln) {
  if constexpr (k > 1) {
    constexpr int half = k / 2;
    const bool h = (ln & off) != 0;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      if constexpr (off >= 16) {
        v[i] = double2{pair_sum<off>(v[i].x, v[i + half].x), pair_sum<off>(v[i].y, v[i + half].y)};
      } else {
        const double2 keep = h ? v[i + half] : v[i];
        const double2 send = h ? v[i] : v[i + half];
        v[i] = double2{keep.x + lane_xor<off>(send.x), keep.y + lane_xor<off>(send.y)};
      }
    }
    tr_step<K, half, off / 2>(v, ln);
  } else {
    double x = v[0].x, y = v[0].y;
    if constexpr (off == 8) {
      x += dpp_mov<kDppRowRor8>(x);
      y += dpp_mov<kDppRowRor8>(y);
    }
    x += dpp_mov<kDppRowHalfMirror>(x);
    y += dpp_mov<kDppRowHalfMirror>(y);
    x += dpp_mov<kDppQuadXor2>(x);
    y += dpp_mov<kDppQuadXor2>(y);
    x += dpp_mov<kDppQuadXor1>(x);
    y += dpp_mov<kDppQuadXor1>(y);
    v[0] = double2{x, y};
  }
}

template <int K>
__device__ __forceinline__ double2 transpose_reduce(double2 (&v)[K]) {
  static_assert(K == 4 || K == 8, "the lane exchanges cover K = 4 and K = 8");
  tr_step<K, K, 32>(v, (int)(threadIdx.x & 63));
  return v[0];
}

template <int K>
__device__ __forceinline__ int reduce_index() {
  const int ln = threadIdx.x & 63;
  int idx = 0, k = K, off = 32;
#pragma unroll
  for (int step = 0; step < 6; ++step) {
    if (k > 1) {
      int half = k / 2;
      if (ln & off) idx += half;
      k = half;
    }
    off >>= 1;
  }
  return idx;
}

struct AdjArgs {
  int lmax, s, nslot, sstride, ngroups, ncomp;
  const double* slot_theta;
  const int* slot_north;
  const int* slot_south;
  const long* mofs;
  const double2* rec;
  const double* cnorm;
  const double2* st_mu;
  const double2* st_p;
  const double2* st_k;
  const int* start;
  const double2* G;
  long sm, sr;
  double2* alm;
};

template <bool SP>
__global__ __launch_bounds__(AdjCfg<SP>::MaxThreads) void k_adjoint(AdjArgs a) {
  constexpr int kAdjR = AdjCfg<SP>::R;
  extern __shared__ double2 lacc[];
  const int m = blockIdx.x;
  const int L = a.lmax;
  const int l0 = l0_of(m, a.s);
  const long nalm = (long)(L + 1) * (L + 2) / 2;
  const long mbase = (long)m * (2 * L + 1 - m) / 2;
  if (l0 > L) {
    for (int l = m + threadIdx.x; l <= L; l += blockDim.x) {
      for (int c = 0; c < a.ncomp; ++c) a.alm[c * nalm + mbase + l] = double2{0.0, 0.0};
    }
    return;
  }
  const int nl = L - l0 + 1;
  constexpr int NA = SP ? 2 : 1;
  for (int i = threadIdx.x; i < nl * NA; i += blockDim.x) lacc[i] = double2{0.0, 0.0};

  __syncthreads();
  const long base = a.mofs[m];
  const double2* __restrict__ rec = a.rec + base;
  const double sgA = ((l0 + m) & 1) ? -1.0 : 1.0;  // (-1)^{l+m} at l = l0
  constexpr int K = SP ? 4 : 8;
  const int ridx = reduce_index<K>();
  const bool writer = ((threadIdx.x & 63) & (64 / K - 1)) == 0;
  const int nthr = blockDim.x;
  const int cap = nthr * kAdjR;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;

  for (int pass0 = 0; pass0 < a.nslot; pass0 += cap) {
  // per-lane ring pairs: the wave takes kAdjR adjacent groups of 64 ring pairs; group r joins
  // the l-loop at its first contributing l, lst[r]
  double x[kAdjR], mu[kAdjR], pm[kAdjR], ku[kAdjR];
  double nv_[kAdjR], pv[kAdjR], kv[kAdjR];
  double2 P0[kAdjR], P1[kAdjR], Q0[kAdjR], Q1[kAdjR];  // spin 0: E/O in P0/P1; spin s: PN, PS, QN, QS
  int lst[kAdjR];
  int lmin = L + 1;
#pragma unroll
  for (int r = 0; r < kAdjR; ++r) {
    const int g = (pass0 >> 6) + wave * kAdjR + r;
    const int slot = g * 64 + lane;
    lst[r] = g < a.ngroups ? a.start[(long)m * a.ngroups + g] : L + 1;
    lmin = min(lmin, lst[r]);
    x[r] = mu[r] = pm[r] = ku[r] = nv_[r] = pv[r] = kv[r] = 0.0;
    P0[r] = P1[r] = Q0[r] = Q1[r] = double2{0.0, 0.0};
    if (slot < a.nslot) {
      const long q = (long)m * a.sstride + slot;
      x[r] = cos(a.slot_theta[slot]);
      const double2 sm = a.st_mu[q], sp = a.st_p[q], sk = a.st_k[q];
      mu[r] = sm.x; pm[r] = sp.x; ku[r] = sk.x;
      if (SP) { nv_[r] = sm.y; pv[r] = sp.y; kv[r] = sk.y; }
      const int rn = a.slot_north[slot];
      int rs = a.slot_south[slot];
      if (rs == rn) rs = -1;
      // the south-ring sign (-1)^{l+m} is sgA at even chunk positions and -sgA at odd ones (chunks start at
      // l0 plus multiples of 4): sgA is folded in here, the alternation into the l-loop's unrolled FMAs
      if (!SP) {
        double2 gn = a.G[(long)m * a.sm + (long)rn * a.sr];
        double2 gs = rs >= 0 ? a.G[(long)m * a.sm + (long)rs * a.sr] : double2{0.0, 0.0};
        P0[r] = sgA > 0 ? cadd(gn, gs) : csub(gn, gs);  // taken at even positions
        P1[r] = sgA > 0 ? csub(gn, gs) : cadd(gn, gs);  // taken at odd positions
      } else {
        const long rowp = (long)(L + m) * a.sm, rowm = (long)(L - m) * a.sm;
        P0[r] = a.G[rowp + (long)rn * a.sr];
        P1[r] = rs >= 0 ? a.G[rowp + (long)rs * a.sr] : double2{0.0, 0.0};
        Q0[r] = cconj(a.G[rowm + (long)rn * a.sr]);
        Q1[r] = rs >= 0 ? cconj(a.G[rowm + (long)rs * a.sr]) : double2{0.0, 0.0};
        P1[r] = double2{sgA * P1[r].x, sgA * P1[r].y};
        Q1[r] = double2{sgA * Q1[r].x, sgA * Q1[r].y};
      }
    }
  }
  // A group that never contributes (start past lmax) is emptied, so that the branch-free chunks below may
  // carry it along; lall is the l from which every other group of the wave has joined.
  int lall = l0;
#pragma unroll
  for (int r = 0; r < kAdjR; ++r) {
    if (lst[r] > L) {
      x[r] = mu[r] = pm[r] = ku[r] = nv_[r] = pv[r] = kv[r] = 0.0;
      P0[r] = P1[r] = Q0[r] = Q1[r] = double2{0.0, 0.0};
    } else {
      lall = max(lall, lst[r]);
    }
  }
  // One chunk of K l values.  FAST: every group has joined, no lane is in the scaled range and every l of the
  // chunk is below lmax, so the chunk runs without branches; otherwise each group, scale and bound is checked.
  auto chunk = [&](int lc, auto fast_tag) {
    constexpr bool FAST = decltype(fast_tag)::value;
    double2 X[K], Y[K];
#pragma unroll
    for (int i = 0; i < K; ++i) X[i] = double2{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < K; ++i) Y[i] = double2{0.0, 0.0};
    bool slow = false;
    bool act[kAdjR];
#pragma unroll
    for (int r = 0; r < kAdjR; ++r) {
      act[r] = FAST || lc >= lst[r];
      if (!FAST) slow |= act[r] && ((ku[r] < 0.0) || (kv[r] < 0.0));
    }
    if (!FAST) slow = __any(slow);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int l = lc + i;
      if (FAST || l <= L) {
        const double2 rc = (FAST || l < L) ? rec[l - l0] : double2{0.0, 0.0};
#pragma unroll
        for (int r = 0; r < kAdjR; ++r) {
          if (!act[r]) continue;
          double wu = mu[r], wv = nv_[r];
          if (!FAST && slow) { wu = ku[r] == 0.0 ? wu : 0.0; wv = kv[r] == 0.0 ? wv : 0.0; }
          if (!SP) {
            fma2(X[i], (i & 1) ? P1[r] : P0[r], wu);
          } else {
            const double sv = (i & 1) ? -wv : wv, su = (i & 1) ? -wu : wu;  // sign is a compile-time constant
            fma2(X[i], P0[r], wu);
            fma2(X[i], P1[r], sv);
            fma2(Y[i], Q0[r], wv);
            fma2(Y[i], Q1[r], su);
          }
          if (FAST || l < L) {
            double t = rc.x * x[r];
            double nu = fma(t + rc.y, mu[r], -pm[r]);
            pm[r] = mu[r]; mu[r] = nu;
            if (SP) { double nn = fma(t - rc.y, nv_[r], -pv[r]); pv[r] = nv_[r]; nv_[r] = nn; }
            if (!FAST && slow) {
              if (ku[r] < 0.0 && fabs(mu[r]) > kThresh) { mu[r] *= kSmall; pm[r] *= kSmall; ku[r] += 1.0; }
              if (SP && kv[r] < 0.0 && fabs(nv_[r]) > kThresh) { nv_[r] *= kSmall; pv[r] *= kSmall; kv[r] += 1.0; }
            }
          }
        }
      }
    }
    double2 xs = transpose_reduce<K>(X);
    if (writer) {
      const int l = lc + ridx;
      if (l <= L) {
        atomicAdd(&lacc[(l - l0) * NA].x, xs.x);
        atomicAdd(&lacc[(l - l0) * NA].y, xs.y);
      }
    }
    if constexpr (SP) {
      double2 ys = transpose_reduce<K>(Y);
      if (writer) {
        const int l = lc + ridx;
        if (l <= L) {
          atomicAdd(&lacc[(l - l0) * NA + 1].x, ys.x);
          atomicAdd(&lacc[(l - l0) * NA + 1].y, ys.y);
        }
      }
    }
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  // group starts are l0 plus multiples of 8, so they fall on chunk boundaries; a lane leaves the scaled range
  // for good once its value is unscaled, so the checked phase ends at the first chunk that qualifies
  int lc = lmin;
  for (; lc <= L; lc += K) {
    bool scaled = false;
#pragma unroll
    for (int r = 0; r < kAdjR; ++r) scaled |= (ku[r] < 0.0) || (kv[r] < 0.0);
    if (lc >= lall && lc + K - 1 < L && !__any(scaled)) break;
    chunk(lc, F_{});
  }
  for (; lc + K - 1 < L; lc += K) chunk(lc, T_{});
  for (; lc <= L; lc += K) chunk(lc, F_{});
  }  // pass
  __syncthreads();

  // write out (also zeros for m <= l < l0)
  const double sgs = (a.s & 1) ? -1.0 : 1.0;
  for (int l = m + threadIdx.x; l <= L; l += blockDim.x) {
    double2 e{0.0, 0.0}, bb{0.0, 0.0};
    if (l >= l0) {
      const double c = a.cnorm[base + (l - l0)];
      if (!SP) {
        e = lacc[l - l0];
        e = double2{c * e.x, m == 0 ? 0.0 : c * e.y};
      } else {
        double2 X = lacc[(l - l0) * 2], Y = lacc[(l - l0) * 2 + 1];
        X = double2{sgs * c * X.x, sgs * c * X.y};
        Y = double2{c * Y.x, c * Y.y};
        e = double2{-0.5 * (X.x + Y.x), -0.5 * (X.y + Y.y)};
        bb = double2{-0.5 * (X.y - Y.y), 0.5 * (X.x - Y.x)};
      }
    }
    a.alm[mbase + l] = e;
    if (SP && a.ncomp > 1) a.alm[nalm + mbase + l] = bb;
  }
}

}  // namespace

// ------------------------------------------------------------------ host
void LegendreTables::build(int L, int s, const std::vector<double>& theta, hipStream_t st) {
  NUSHT_REQUIRE(L >= 0 && s >= 0 && s <= 3, kErrArg, "legendre: bad lmax/spin");
  if (const char* e = getenv("NUSHT_LEGENDRE_MFMA"); e && e[0] == '0') mfma = false;
  lmax = L;
  spin = s;
  nring = (int)theta.size();
  n_prep = "legendre_prep_" + tag + "_s" + std::to_string(s);
  n_synth = "legendre_synth_" + tag + "_s" + std::to_string(s);
  n_adj = "legendre_adjoint_" + tag + "_s" + std::to_string(s);
  // pair rings symmetric about the equator
  std::vector<int> order(nring);
  for (int i = 0; i < nring; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return theta[a] < theta[b]; });
  std::vector<char> used(nring, 0);
  std::vector<double> sth;
  std::vector<int> sn, ss;
  int lo = 0, hi = nring - 1;
  while (lo <= hi) {
    int a = order[lo], b = order[hi];
    if (lo == hi) {
      sth.push_back(theta[a]); sn.push_back(a); ss.push_back(a);
      break;
    }
    if (std::fabs(theta[a] + theta[b] - M_PI) < 1e-13) {
      sth.push_back(theta[a]); sn.push_back(a); ss.push_back(b);
      ++lo; --hi;
    } else if (theta[a] + theta[b] < M_PI) {
      sth.push_back(theta[a]); sn.push_back(a); ss.push_back(-1);
      ++lo;
    } else {
      sth.push_back(theta[b]); sn.push_back(b); ss.push_back(-1);
      --hi;
    }
  }
  // an equator ring paired with itself: keep it as north only
  for (size_t i = 0; i < sn.size(); ++i)
    if (ss[i] == sn[i]) ss[i] = -1;
  nslot = (int)sth.size();
  upload(slot_theta, sth);
  upload(slot_north, sn);
  upload(slot_south, ss);

  std::vector<long> mo(L + 2, 0);
  for (int m = 0; m <= L; ++m) {
    int l0 = l0_of(m, s);
    mo[m + 1] = mo[m] + (l0 <= L ? (L - l0 + 1) : 0);
  }
  nlm = mo[L + 1];
  upload(mofs, mo);
  std::vector<long> mb(L + 1, 0);
  nblk = 0;
  for (int m = 0; m <= L; ++m) {
    mb[m] = nblk;
    nblk += ceil_div((long)(mo[m + 1] - mo[m]), 4L);
  }
  upload(mblk, mb);
  // padded so that a chunked, prefetching l-loop may read up to 2 kSynCH entries past the last one
  rec.alloc(sizeof(double2) * (size_t)(std::max<long>(nlm, 1) + 2 * kSynCH));
  cnorm.alloc(sizeof(double) * (size_t)(std::max<long>(nlm, 1) + 2 * kSynCH));
  sstride = nslot;
  ngroups = ceil_div(nslot, 64);
  const size_t nst = (size_t)(L + 1) * sstride;
  st_mu.alloc(sizeof(double2) * nst);
  st_p.alloc(sizeof(double2) * nst);
  st_k.alloc(sizeof(double2) * nst);
  start.alloc(sizeof(int) * (size_t)(L + 1) * ngroups);
  hipLaunchKernelGGL(k_rec_tables, dim3(ceil_div(L + 1, 128)), dim3(128), 0, st, L, s, mofs.as<long>(),
                     rec.as<double2>(), cnorm.as<double>());
  NUSHT_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_seeds, dim3(ceil_div(nslot, 64)), dim3(64), 0, st, L, s, nslot, sstride,
                     slot_theta.as<double>(), st_mu.as<double2>(), st_p.as<double2>(), st_k.as<double2>());
  NUSHT_LAUNCH_CHECK();
  const dim3 grid((unsigned)((long)ngroups * (L + 1)));
  if (s > 0)
    hipLaunchKernelGGL(k_advance<true>, grid, dim3(64), 0, st, L, s, nslot, sstride, ngroups,
                       slot_theta.as<double>(), mofs.as<long>(), rec.as<double2>(), st_mu.as<double2>(),
                       st_p.as<double2>(), st_k.as<double2>(), start.as<int>());
  else
    hipLaunchKernelGGL(k_advance<false>, grid, dim3(64), 0, st, L, s, nslot, sstride, ngroups,
                       slot_theta.as<double>(), mofs.as<long>(), rec.as<double2>(), st_mu.as<double2>(),
                       st_p.as<double2>(), st_k.as<double2>(), start.as<int>());
  NUSHT_LAUNCH_CHECK();
  // work actually run by a synthesis (reported with the plan; synchronizes once)
  std::vector<int> hs((size_t)(L + 1) * ngroups);
  NUSHT_HIP(hipMemcpyAsync(hs.data(), start.get(), sizeof(int) * hs.size(), hipMemcpyDeviceToHost, st));
  NUSHT_HIP(hipStreamSynchronize(st));
  pair_steps = 0;
  for (int m = 0; m <= L; ++m) {
    if (l0_of(m, s) > L) continue;
    for (int g = 0; g < ngroups; ++g)
      pair_steps += (long)(L + 1 - hs[(size_t)m * ngroups + g]) * std::min(64, nslot - 64 * g);
  }
}

static int padded_maps(int nmaps) { return nmaps <= 1 ? 1 : nmaps <= 2 ? 2 : nmaps <= 4 ? 4 : 8; }

size_t legendre_coef_bytes(const LegendreTables& t, int nmaps) {
  // padded by 2 kSynCH entries (chunked, prefetching reads past the last l of the last m)
  const size_t vec = sizeof(double2) * (size_t)(std::max<long>(t.nlm, 1) + 2 * kSynCH) * padded_maps(nmaps) *
                     (t.spin ? 2 : 1);
  if (nmaps < 3) return vec;
  // matrix-core layout: 32 doubles per chain, map group and block of 4 l, 2 blocks of padding (prefetch)
  const int ns = nmaps >= kLegendreMfma8Min ? 2 : 1;
  const size_t mf = sizeof(double) * 32 * ns * (t.spin ? 2 : 1) * (size_t)(t.nblk + 2);
  return std::max(vec, mf);
}

// Synthesis variants (l-chunk CH, waves per block WPB); the autotuner picks one per table.
struct SynVariant {
  int CH, WPB;
};
static const SynVariant kSynVariants[] = {{2, 1}, {2, 4}, {4, 4}, {2, 8}};
constexpr int kNumSynVariants = 4;

template <bool SP, int NB, bool GR = false>
static void launch_synth(const SynthArgs& a, int variant, hipStream_t st) {
  // Odd number of blocks per m: blocks are handed to XCDs and shader engines round-robin, and
  // with a power-of-two count every engine would see the same slot groups for every m, so the
  // polar groups (long underflow phases) would pile up on a few engines.
  const int wpb = kSynVariants[variant].WPB;
  const int nsb = ceil_div(ceil_div(a.nslot, 64), wpb) | 1;
  const int nblocks = nsb * ceil_div(a.lmax + 1, kNumXcd) * kNumXcd;
  SynthArgs b = a;
  b.nsb = nsb;
  const dim3 blk(64 * wpb);
  switch (variant) {
    case 0: hipLaunchKernelGGL((k_synth<SP, NB, 2, 1, GR>), dim3(nblocks), blk, 0, st, b); break;
    case 1: hipLaunchKernelGGL((k_synth<SP, NB, 2, 4, GR>), dim3(nblocks), blk, 0, st, b); break;
    case 2: hipLaunchKernelGGL((k_synth<SP, NB, 4, 4, GR>), dim3(nblocks), blk, 0, st, b); break;
    default: hipLaunchKernelGGL((k_synth<SP, NB, 2, 8, GR>), dim3(nblocks), blk, 0, st, b); break;
  }
  NUSHT_LAUNCH_CHECK();
}

static void legendre_synth_chunk(const LegendreTables& t, int nmaps, int ncomp, const double2* alm,
                                 long alm_stride_map, const double* fl, double2* coef_ws, double2* G, long sm, long sr,
                                 long smap, hipStream_t st);

static void legendre_synth_mfma(const LegendreTables& t, int nmaps, int ncomp, const AlmPtrs8& alm,
                                const double* fl, double2* coef_ws, double2* G, long sm, long sr, long smap,
                                hipStream_t st);

static AlmPtrs8 strided_maps(const double2* alm, long stride, int nmaps) {
  AlmPtrs8 a{};
  for (int b = 0; b < 8; ++b) a.p[b] = alm + (b < nmaps ? b : 0) * stride;
  return a;
}

// maps of the next matrix-core launch for `left` remaining maps (>= 3); NUSHT_LEGENDRE_MFMA8=0 keeps 4 per launch
static int mfma_maps(int left) {
  const char* e = getenv("NUSHT_LEGENDRE_MFMA8");
  const bool eight = !(e && e[0] == '0');
  return eight && left >= kLegendreMfma8Min ? std::min(8, left) : std::min(kLegendreMfmaMaps, left);
}

void legendre_synth(const LegendreTables& t, int nmaps, int ncomp, const double2* alm, long alm_stride_map,
                    const double* fl, double2* coef_ws, double2* G, long sm, long sr, long smap, hipStream_t st) {
  NUSHT_REQUIRE(nmaps >= 1 && nmaps <= 8, kErrArg, "legendre_synth: nmaps must be in [1, 8]");
  const int ch = std::max(1, std::min(t.chunk, 2));
  int b = 0;
  while (b < nmaps) {
    const int left = nmaps - b;
    if (t.mfma && left >= 3) {
      const int nb = mfma_maps(left);
      legendre_synth_mfma(t, nb, ncomp, strided_maps(alm + b * alm_stride_map, alm_stride_map, nb), fl, coef_ws,
                          G + b * smap, sm, sr, smap, st);
      b += nb;
    } else {
      const int nb = std::min(ch, left);
      legendre_synth_chunk(t, nb, ncomp, alm + b * alm_stride_map, alm_stride_map, fl, coef_ws, G + b * smap, sm,
                           sr, smap, st);
      b += nb;
    }
  }
}

void legendre_synth_maps(const LegendreTables& t, int nmaps, int ncomp, const double2* const* alm, double2* coef_ws,
                         double2* G, long sm, long sr, long smap, hipStream_t st) {
  NUSHT_REQUIRE(nmaps >= 1 && nmaps <= 8, kErrArg, "legendre_synth_maps: nmaps must be in [1, 8]");
  int b = 0;
  while (b < nmaps) {
    const int left = nmaps - b;
    if (t.mfma && left >= 3) {
      const int nb = mfma_maps(left);
      AlmPtrs8 ap{};
      for (int k = 0; k < 8; ++k) ap.p[k] = alm[b + (k < nb ? k : 0)];
      legendre_synth_mfma(t, nb, ncomp, ap, nullptr, coef_ws, G + b * smap, sm, sr, smap, st);
      b += nb;
    } else {
      legendre_synth(t, 1, ncomp, alm[b], 0, nullptr, coef_ws, G + b * smap, sm, sr, smap, st);
      b += 1;
    }
  }
}

static void legendre_synth_mfma(const LegendreTables& t, int nmaps, int ncomp, const AlmPtrs8& alm,
                                const double* fl, double2* coef_ws, double2* G, long sm, long sr, long smap,
                                hipStream_t st) {
  NUSHT_REQUIRE(nmaps >= 1 && nmaps <= 8, kErrArg, "matrix-core synthesis: 1 to 8 maps per launch");
  const int L = t.lmax;
  const int ns = nmaps > kLegendreMfmaMaps ? 2 : 1;  // 16-column map groups
  double* coef4 = reinterpret_cast<double*>(coef_ws);
  {
    StageScope sc(t.n_prep.c_str(), st);
    // block row y covers m = y and m = L - y; its x extent is sized for m = 0 (the longest table)
    dim3 grid(ceil_div((L + 4) / 4 * 16 * ns, 256), L / 2 + 1);
    if (ns == 1)
      hipLaunchKernelGGL(k_prep4<1>, grid, dim3(256), 0, st, L, t.spin, nmaps, ncomp, alm, fl, t.mofs.as<long>(),
                         t.mblk.as<long>(), t.cnorm.as<double>(), coef4);
    else
      hipLaunchKernelGGL(k_prep4<2>, grid, dim3(256), 0, st, L, t.spin, nmaps, ncomp, alm, fl, t.mofs.as<long>(),
                         t.mblk.as<long>(), t.cnorm.as<double>(), coef4);
    NUSHT_LAUNCH_CHECK();
  }
  SynthArgs a;
  a.lmax = L; a.s = t.spin; a.nslot = t.nslot; a.sstride = t.sstride; a.nmaps = nmaps;
  a.nsb = ceil_div(a.nslot, 64) | 1;
  a.slot_theta = t.slot_theta.as<double>();
  a.slot_north = t.slot_north.as<int>();
  a.slot_south = t.slot_south.as<int>();
  a.mofs = t.mofs.as<long>();
  a.rec = t.rec.as<double2>();
  a.st_mu = t.st_mu.as<double2>();
  a.st_p = t.st_p.as<double2>();
  a.st_k = t.st_k.as<double2>();
  a.start = t.start.as<int>();
  a.coef = coef_ws;
  a.G = G; a.sm = sm; a.sr = sr; a.smap = smap;
  const int nblocks = a.nsb * ceil_div(L + 1, kNumXcd) * kNumXcd;
  StageScope sc(t.n_synth.c_str(), st);
  if (t.spin > 0 && ns == 1)
    hipLaunchKernelGGL((k_synth_mfma<true, 1>), dim3(nblocks), dim3(64), 0, st, a, coef4, t.mblk.as<long>());
  else if (t.spin > 0)
    hipLaunchKernelGGL((k_synth_mfma<true, 2>), dim3(nblocks), dim3(64), 0, st, a, coef4, t.mblk.as<long>());
  else if (ns == 1)
    hipLaunchKernelGGL((k_synth_mfma<false, 1>), dim3(nblocks), dim3(64), 0, st, a, coef4, t.mblk.as<long>());
  else
    hipLaunchKernelGGL((k_synth_mfma<false, 2>), dim3(nblocks), dim3(64), 0, st, a, coef4, t.mblk.as<long>());
  NUSHT_LAUNCH_CHECK();
}

static void legendre_synth_chunk(const LegendreTables& t, int nmaps, int ncomp, const double2* alm,
                                 long alm_stride_map, const double* fl, double2* coef_ws, double2* G, long sm, long sr,
                                 long smap, hipStream_t st) {
  const int L = t.lmax;
  const int nbpad = padded_maps(nmaps);
  {
    StageScope sc(t.n_prep.c_str(), st);
    dim3 grid(ceil_div(L + 1, 256), L + 1);
    hipLaunchKernelGGL(k_prep, grid, dim3(256), 0, st, L, t.spin, nmaps, nbpad, ncomp, alm, alm_stride_map, fl,
                       t.mofs.as<long>(), t.cnorm.as<double>(), coef_ws);
    NUSHT_LAUNCH_CHECK();
  }
  SynthArgs a;
  a.lmax = L; a.s = t.spin; a.nslot = t.nslot; a.sstride = t.sstride; a.nsb = 0; a.nmaps = nmaps;
  a.slot_theta = t.slot_theta.as<double>();
  a.slot_north = t.slot_north.as<int>();
  a.slot_south = t.slot_south.as<int>();
  a.mofs = t.mofs.as<long>();
  a.rec = t.rec.as<double2>();
  a.st_mu = t.st_mu.as<double2>();
  a.st_p = t.st_p.as<double2>();
  a.st_k = t.st_k.as<double2>();
  a.start = t.start.as<int>();
  a.coef = coef_ws;
  a.G = G; a.sm = sm; a.sr = sr; a.smap = smap;
  const bool sp = t.spin > 0;
  const bool gr = sp && ncomp == 1;  // E only: the gradient-only contraction
  const int tv = gr ? t.variant_grad : t.variant;
  const int var = (tv >= 0 && tv < kNumSynVariants) ? tv : 0;
  StageScope sc(t.n_synth.c_str(), st);
  switch (nbpad) {  // the kernel's NB must equal the coefficient stride written by k_prep
    case 1:
      if (gr) launch_synth<true, 1, true>(a, var, st);
      else if (sp) launch_synth<true, 1>(a, var, st);
      else launch_synth<false, 1>(a, var, st);
      break;
    case 2:
      if (gr) launch_synth<true, 2, true>(a, var, st);
      else if (sp) launch_synth<true, 2>(a, var, st);
      else launch_synth<false, 2>(a, var, st);
      break;
    default: throw Error{kErrUnsupported, "legendre_synth: at most 2 maps per launch"};
  }
}

int legendre_variant_count() { return kNumSynVariants; }

void legendre_autotune(LegendreTables& t, int ncomp, double2* coef_ws, double2* G, long sm, long sr, hipStream_t st) {
  // NUSHT_LEGENDRE_VARIANT=k fixes the variant (tests run every variant this way)
  if (const char* fv = getenv("NUSHT_LEGENDRE_VARIANT"); fv && fv[0]) {
    const int v = atoi(fv);
    NUSHT_REQUIRE(v >= 0 && v < kNumSynVariants, kErrArg, "NUSHT_LEGENDRE_VARIANT out of range");
    t.variant = t.variant_grad = v;
    t.tuned_ms = 0.f;
    return;
  }
  const bool gr = t.spin > 0 && ncomp == 1;
  int& slot = gr ? t.variant_grad : t.variant;
  // one measurement per (lmax, spin, ring pairs) and process
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, std::pair<int, float>> cache;
  const auto key = std::make_tuple(t.lmax, t.spin, t.nslot, ncomp);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      slot = it->second.first;
      if (!gr) t.tuned_ms = it->second.second;
      return;
    }
  }
  DevBuf dummy(sizeof(double2) * (size_t)(t.lmax + 1) * (t.lmax + 2) / 2 * ncomp);
  NUSHT_HIP(hipMemsetAsync(dummy.get(), 0, dummy.bytes(), st));
  const double2* alm = dummy.as<double2>();
  hipEvent_t e0, e1;
  NUSHT_HIP(hipEventCreate(&e0));
  NUSHT_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bestv = 0;
  const char* env = getenv("NUSHT_VERBOSE");
  const bool verbose = env && env[0] && env[0] != '0';
  for (int v = 0; v < kNumSynVariants; ++v) {
    slot = v;
    legendre_synth(t, 1, ncomp, alm, 0, nullptr, coef_ws, G, sm, sr, 0, st);  // warm-up
    NUSHT_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < 2; ++k) legendre_synth(t, 1, ncomp, alm, 0, nullptr, coef_ws, G, sm, sr, 0, st);
    NUSHT_HIP(hipEventRecord(e1, st));
    NUSHT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NUSHT_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (verbose)
      fprintf(stderr, "[nusht] legendre autotune %s lmax %d spin %d pairs %d: variant %d (CH %d, %d waves) %.3f ms\n",
              t.tag.c_str(), t.lmax, t.spin, t.nslot, v, kSynVariants[v].CH, kSynVariants[v].WPB, ms / 2);
    if (ms < best) { best = ms; bestv = v; }
  }
  slot = bestv;
  if (!gr) t.tuned_ms = best / 2;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> g(mu);
  cache[key] = {bestv, best / 2};
}

void legendre_adjoint(const LegendreTables& t, int ncomp, const double2* G, long sm, long sr, double2* alm,
                      hipStream_t st) {
  const int L = t.lmax;
  AdjArgs a;
  a.lmax = L; a.s = t.spin; a.nslot = t.nslot; a.sstride = t.sstride; a.ngroups = t.ngroups; a.ncomp = ncomp;
  a.slot_theta = t.slot_theta.as<double>();
  a.slot_north = t.slot_north.as<int>();
  a.slot_south = t.slot_south.as<int>();
  a.mofs = t.mofs.as<long>();
  a.rec = t.rec.as<double2>();
  a.cnorm = t.cnorm.as<double>();
  a.st_mu = t.st_mu.as<double2>();
  a.st_p = t.st_p.as<double2>();
  a.st_k = t.st_k.as<double2>();
  a.start = t.start.as<int>();
  a.G = G; a.sm = sm; a.sr = sr; a.alm = alm;
  const int R = t.spin ? AdjCfg<true>::R : AdjCfg<false>::R;
  const int maxw = (t.spin ? AdjCfg<true>::MaxThreads : AdjCfg<false>::MaxThreads) / 64;
  int nw = std::min(maxw, ceil_div(t.nslot, 64 * R));  // more rings run as further passes
  if (nw < 1) nw = 1;
  const int na = t.spin ? 2 : 1;
  const size_t lds = sizeof(double2) * (size_t)(L + 1) * na;
  NUSHT_REQUIRE(lds <= 160 * 1024, kErrUnsupported, "legendre_adjoint: lmax too large for the LDS accumulator");
  StageScope sc(t.n_adj.c_str(), st);
  if (t.spin) {
    NUSHT_HIP(hipFuncSetAttribute((const void*)k_adjoint<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_adjoint<true>, dim3(L + 1), dim3(64 * nw), lds, st, a);
  } else {
    NUSHT_HIP(hipFuncSetAttribute((const void*)k_adjoint<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_adjoint<false>, dim3(L + 1), dim3(64 * nw), lds, st, a);
  }
  NUSHT_LAUNCH_CHECK();
}

}  // namespace nusht
