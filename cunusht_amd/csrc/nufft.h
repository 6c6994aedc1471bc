// Double-Fourier-sphere transform + exponential-of-semicircle nuFFT on gfx950.
//
// Replaces the reference's doubling -> cuFFT fft2 -> cuFINUFFT type 2 chain
// (cunusht/deflection/GPU_nufft_transformer.py:421-435) and its adjoint
// (cuFINUFFT type 1 -> ifft2 -> adjoint doubling, :475-482) with three passes over
// per-m ring coefficients G_m(theta_j) on an equiangular grid including the poles:
//
//   theta-pass  (one workgroup per m): double G_m onto the torus with sign (-1)^{s+m},
//               forward FFT of length ntd, keep |k| <= lmax, divide by the kernel's Fourier
//               transform in theta and phi, zero-pad and inverse-FFT onto nf theta cells;
//   phi-pass    (one workgroup per fine theta row): place m = -lmax..lmax on nf phi cells,
//               inverse FFT -> fine grid row;
//   interp      type-2 ES interpolation at the (sorted) points, with the spin-s rotation.
//
// The adjoint runs the exact transposes in reverse (spread, forward phi FFT, theta FFT /
// deconvolution / inverse FFT / fold).  Only the fine-grid rows that the kernel footprint of
// points with theta in [0, pi] can touch (about nf/2 + w of nf) are ever formed.
#pragma once

#include <memory>

#include "common.h"
#include "fft_engine.h"

namespace nusht {

struct NufftParams {
  int lmax = 0;
  int spin = 0;
  double eps = 1e-7;
  bool single = false;   // FP32 FFTs and nuFFT (coordinates stay FP64)
  int ntheta = 0;        // CC rings, theta_j = j pi / (ntheta - 1)
  int ntd = 0;           // 2 ntheta - 2 (torus length in theta)
  int nf = 0;            // fine grid size (both axes)
  double sigma = 0;      // nf / (2 lmax + 1)
  int w = 0;             // kernel width in cells
  double beta = 0;
  int row_lo = 0;        // first fine theta row kept (may be negative: wraps)
  int nrow = 0;          // number of fine theta rows kept
  int nm = 0;            // 2 lmax + 1 phi modes
  int poly_deg = 0;      // degree of the piecewise-polynomial kernel (interior cells)
  double poly_err = 0;   // its measured max abs error against the exact kernel

  static NufftParams choose(int lmax, int spin, double eps, double sigma_hint = 0.0);
};

// Precomputed tables and per-plan buffers.
class NufftPlan {
 public:
  // tile_rows: rows of the register row-window spreading tiles (0: 8, or NUSHT_SPREAD_TR)
  NufftPlan(const NufftParams& p, long npts_max, hipStream_t st, int tile_rows = 0);
  int tile_rows() const { return tile_r_; }
  int tile_cols() const { return tile_c_; }
  bool rows_spread() const { return rows_spread_; }
  const NufftParams& params() const { return p_; }

  // Sort points into fine-grid tiles (theta, phi in radians; theta in [0, pi]).
  // Bins the points by fine-grid tile.  With `check`, synchronizes once and rejects points with
  // theta outside the fine grid's rows (or NaN); without it (points produced on the device by
  // this library), such points are binned to a tile whose footprint does not contain them, so
  // interpolation returns NaN for them and spreading skips them.
  void set_pts(const double* theta, const double* phi, long npts, hipStream_t st, bool check = true);
  long npts() const { return npts_; }

  // type 2: G (rows m, ntheta rings; row of m = (spin ? lmax + m : m)) -> values.
  //   spin 0: out_re[j] = Re c_j ;  spin s: (out_re, out_im)[j] = c_j * e^{i s gamma_j} (gamma optional)
  void type2(const double2* G, long g_stride_m, const double* gamma, double* out_re, double* out_im,
             hipStream_t st);
  // type 1 (exact adjoint of type2): in_re (and in_im for spin s) -> G rows.
  void type1(const double* in_re, const double* in_im, const double* gamma, double2* G, long g_stride_m,
             hipStream_t st);

  size_t bytes() const;

 private:
  template <typename T> void type2_impl(const double2* G, long sm, const double* gamma, double* ore, double* oim,
                                        hipStream_t st);
  template <typename T> void type1_impl(const double* ire, const double* iim, const double* gamma, double2* G,
                                        long sm, hipStream_t st);

  NufftParams p_;
  long npts_max_ = 0;
  long npts_ = 0;
  int tile_r_ = 32, tile_c_ = 32;  // spreading tile (fine cells)
  bool rows_spread_ = false;       // register row-window spreading (8 <= w <= 16)
  int rows_cap_ = 0;               // its LDS point records per pass
  int rows_seg_ = 1;               // tiles one workgroup walks down a column strip
  int ntile_r_ = 0, ntile_c_ = 0;
  // tables
  DevBuf deconv_;                  // double[lmax+1] : 1 / psi_hat(k)
  DevBuf kpoly_;                   // T[(poly_deg+1) * (w-2)] kernel polynomials, degree-major
  FftLength<float> ftd_f_, fnf_f_;
  FftLength<double> ftd_d_, fnf_d_;
  // buffers
  DevBuf T_;                       // [nrow][nm] complex (theta-pass output / phi-pass input)
  DevBuf grid_;                    // [nrow][nf] complex fine grid
  DevBuf pts_;                     // double2[npts_max] sorted (theta, phi)
  DevBuf idx_;                     // int[npts_max] original index of sorted point
  DevBuf tile_start_;              // int[ntiles + 1]
  DevBuf tile_cnt_;                // int[ntiles + 1]
  DevBuf key_;                     // int[npts_max] tile of each point
  DevBuf cursor_;                  // int[ntiles]
  DevBuf scan_tmp_;                // hipcub scan temp storage
  size_t scan_tmp_bytes_ = 0;
};

// Plan-time choice of the fine-grid oversampling sigma and the spreading tile height by measurement: every
// candidate (sigma in {1.25, 1.5, 2} and the default rule, as far as the fine grid fits the in-LDS
// transforms; tile heights 8, 16, 32 where the register row-window spreading applies) is built and timed on
// npts_max synthetic points (theta = acos U(-1, 1), phi = U(0, 2 pi)) for one type-2 and one type-1
// transform from / into G (rows m with stride g_stride_m; contents are overwritten); the plan with the
// least type-2 + type-1 time is returned.  Results are cached per (device, lmax, spin, eps, npts_max).
struct NufftTune {
  double sigma = 0;
  int tile_rows = 0;
  float type2_ms = 0, type1_ms = 0;
  int candidates = 0;
};
std::unique_ptr<NufftPlan> nufft_autotune(int lmax, int spin, double eps, long npts_max, double2* G, long g_stride_m,
                                          hipStream_t st, NufftTune& tune);

}  // namespace nusht
