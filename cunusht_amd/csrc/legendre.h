// Spin-weighted Legendre ring transforms on gfx950 (FP64).
//
// Field conventions (orthonormal, Condon-Shortley, healpy/libsharp spin convention):
//   sY_lm(theta, 0) = (-1)^s sqrt((2l+1)/4pi) d^l_{m,-s}(theta)
//   spin 0 : f = sum a_lm Y_lm                (real field, m >= 0 stored)
//   spin s : Q + iU = - sum (E_lm + i B_lm) sY_lm   (E, B real-field alm)
// The ring coefficients G_m(theta) satisfy f(theta, phi) = sum_m G_m(theta) e^{i m phi}.
//
// Per m the Wigner-d functions u_l = N_l d^l_{m,-s}, v_l = N_l d^l_{m,+s} obey
//   u_{l+1} = (alpha_l x + b_l) u_l - gamma_l u_{l-1},   v_{l+1} = (alpha_l x - b_l) v_l - gamma_l v_{l-1}.
// Writing u_l = c_l mu_l with c_{l+1} = gamma_l c_{l-1} turns this into
//   mu_{l+1} = (A_l x +- B_l) mu_l - mu_{l-1},   A_l = alpha_l c_l / c_{l+1},  B_l = b_l c_l / c_{l+1},
// and c_l is folded into the (pre-scaled) alm, so each step costs one multiply and two FMAs.
// North/south ring pairs share one recursion: u(pi - theta) = (-1)^{l+m} v(theta).
#pragma once

#include "common.h"

namespace nusht {

// Per-plan device tables for one (lmax, spin, ring set).
struct LegendreTables {
  int lmax = 0;
  int spin = 0;
  int nslot = 0;          // ring pairs + unpaired rings
  int sstride = 0;        // row stride of the state tables (odd, >= nslot)
  int ngroups = 0;        // groups of 64 consecutive slots
  int nring = 0;
  DevBuf slot_theta;      // double[nslot]   theta of the slot's northern ring
  DevBuf slot_north;      // int[nslot]      ring index
  DevBuf slot_south;      // int[nslot]      mirrored ring index or -1
  DevBuf mofs;            // int64[lmax+2]   offset of m's l-range in the (l, m) tables
  DevBuf rec;             // double2[nlm]    (A_l, B_l)
  DevBuf cnorm;           // double[nlm]     c_l
  // Recursion state per (m, slot) at the first l its 64-slot group needs: every l below it has
  // all values of the group under 2^-300 (they would add only zeros).  start - l0 is a multiple of 8.
  DevBuf st_mu;           // double2[(lmax+1) * sstride]  (mu_u, mu_v) at l = start
  DevBuf st_p;            // double2[(lmax+1) * sstride]  (mu_u, mu_v) at l = start - 1 (0 at l0)
  DevBuf st_k;            // double2[(lmax+1) * sstride]  scale exponents (k_u, k_v) at l = start
  DevBuf start;           // int[(lmax+1) * ngroups]
  DevBuf mblk;            // int64[lmax+1]  first block of 4 l of each m in the matrix-core coefficient table
  long nblk = 0;          // sum over m of ceil((lmax - l0(m) + 1) / 4)
  long nlm = 0;           // sum over m of (lmax - l0(m) + 1)
  long pair_steps = 0;    // (l, ring pair) recursion steps one synthesis runs (from `start`)
  std::string tag = "rings";  // stage-timer label ("cc" for the nuSHT grid, "ring" for the pointing grid)
  std::string n_prep, n_synth, n_adj;
  int chunk = 2;          // maps per synthesis launch (VALU kernels)
  bool mfma = true;       // batches of >= 3 maps on the matrix cores (NUSHT_LEGENDRE_MFMA=0 turns it off)
  int variant = -1;       // synthesis kernel variant (-1: default); set by legendre_autotune
  int variant_grad = -1;  // the same for the E-only (gradient) contraction of spin-s tables
  float tuned_ms = 0.f;

  void build(int lmax, int spin, const std::vector<double>& ring_theta, hipStream_t st);
  size_t bytes() const {
    return slot_theta.bytes() + slot_north.bytes() + slot_south.bytes() + mofs.bytes() + rec.bytes() +
           cnorm.bytes() + st_mu.bytes() + st_p.bytes() + st_k.bytes() + start.bytes() + mblk.bytes();
  }
};

// Synthesis: alm -> G.  alm: nmaps x ncomp x nalm complex (healpy layout, stride alm_stride between maps).
// Maps are processed `chunk` at a time; one recursion per (m, ring pair) serves all maps of a chunk.
// coef_ws: workspace of legendre_coef_bytes() bytes.  G row of m: spin 0 -> m, spin>0 -> lmax + m
// (negative m -> lmax - |m|); element (row, ring) at G + row * g_stride_m + ring * g_stride_r (+ map * g_stride_map).
// fl (optional, device double[lmax+1]) multiplies E by fl[l] before synthesis.
size_t legendre_coef_bytes(const LegendreTables& t, int nmaps);
// Maps per launch of the matrix-core synthesis path (used for batches of at least 3 maps); batches of at least
// kLegendreMfma8Min maps take up to 8 per launch (two 16-column groups sharing one recursion).
constexpr int kLegendreMfmaMaps = 4;
constexpr int kLegendreMfma8Min = 7;
void legendre_synth(const LegendreTables& t, int nmaps, int ncomp, const double2* alm, long alm_stride_map,
                    const double* fl, double2* coef_ws, double2* G, long g_stride_m, long g_stride_r,
                    long g_stride_map, hipStream_t st);

// Per-map alm of a matrix-core launch (up to 8 maps at their own addresses).
struct AlmPtrs8 {
  const double2* p[8];
};
// The same synthesis for nmaps maps whose alm lie at separate addresses (one pointer per map, host array);
// G as in legendre_synth.
void legendre_synth_maps(const LegendreTables& t, int nmaps, int ncomp, const double2* const* alm, double2* coef_ws,
                         double2* G, long g_stride_m, long g_stride_r, long g_stride_map, hipStream_t st);

// Number of synthesis kernel variants.
int legendre_variant_count();

// Times every synthesis variant on the given buffers (contents are overwritten) and keeps the fastest for
// this ncomp (spin-s tables keep one variant for E + B and one for E alone)
// (or takes the one named by the environment variable NUSHT_LEGENDRE_VARIANT).
// (Plan creation only: allocates a scratch alm set and synchronizes.)
void legendre_autotune(LegendreTables& t, int ncomp, double2* coef_ws, double2* G, long sm, long sr, hipStream_t st);

// Adjoint: G -> alm (exact transpose of legendre_synth under the real-field alm inner product).
// For spin 0 only rows m = 0..lmax are read (the field is real); m = 0 imaginary parts are dropped.
void legendre_adjoint(const LegendreTables& t, int ncomp, const double2* G, long g_stride_m, long g_stride_r,
                      double2* alm, hipStream_t st);

}  // namespace nusht
