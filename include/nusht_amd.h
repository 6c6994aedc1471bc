/*
 * nusht_amd.h -- C ABI of the MI355X (gfx950) nuSHT library  (libnusht_amd.so)
 *
 * Drop-in boundary for the GPU backend of cunuSHT.  Every pointer argument that names an array
 * is a DEVICE pointer unless the comment says "host".  Every call takes an explicit stream
 * (hipStream_t, NULL = default stream), enqueues its work and returns without synchronizing
 * unless stated.  Every call returns NUSHT_OK (0) or an error code; nusht_last_error() gives
 * the message of the calling thread's last failure.  No call allocates device memory except
 * the *_create functions.
 *
 * Complex arrays are interleaved (re, im) doubles.  alm arrays use the healpy m-major layout
 * idx(l, m) = m (2 lmax + 1 - m) / 2 + l with nalm = (lmax + 1)(lmax + 2) / 2 (mmax = lmax),
 * one block of nalm per component (E then B for spin > 0).
 *
 * Conventions (orthonormal, Condon-Shortley): spin 0 f = sum a_lm Y_lm (real);
 * spin s > 0: Q + iU = - sum (E_lm + i B_lm) sY_lm, returned as the two real arrays (Q, U).
 */
#ifndef NUSHT_AMD_H
#define NUSHT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* nusht_stream_t;                 /* hipStream_t */
typedef struct nusht_plan_s* nusht_plan_t;     /* nuSHT plan: one (lmax, spin, epsilon) */
typedef struct nusht_ringplan_s* nusht_ringplan_t; /* SHT on the iso-latitude pointing grid */
typedef struct nusht_pointplan_s* nusht_pointplan_t; /* exact synthesis at a few arbitrary points */

#define NUSHT_OK 0
#define NUSHT_ERR_ARG 1
#define NUSHT_ERR_HIP 2
#define NUSHT_ERR_UNSUPPORTED 3
#define NUSHT_ERR_STATE 4
#define NUSHT_ERR_NOMEM 5

const char* nusht_last_error(void);
int nusht_version(void);

/* ---------------------------------------------------------------- stage kernels
 * Theta doubling onto the double Fourier sphere.
 * Replaces cunusht/c/doubling.cu:203-220 (CUdoubling_contig_1D, exported at :237-250):
 *   map  (nphi, ntheta) theta-contiguous  ->  dmap (nphi, 2 ntheta - 2),
 *   dmap[p][r] = map[p][r] (r < ntheta), (-1)^spin map[(p + nphi/2) % nphi][2 ntheta - 2 - r] otherwise
 * (spin sign from cunusht/deflection/CPU_nufft_transformer.py:284-296).  nphi must be even. */
int nusht_doubling_f64(const double* map, int ntheta, int nphi, int spin, double* dmap, nusht_stream_t s);
int nusht_doubling_f32(const float* map, int ntheta, int nphi, int spin, float* dmap, nusht_stream_t s);
int nusht_doubling_c128(const double* map, int ntheta, int nphi, int spin, double* dmap, nusht_stream_t s);

/* Adjoint doubling.  Replaces cunusht/c/doubling.cu:75-92 (CUadjoint_doubling_1D, exported at :251-264):
 *   dmap (2 ntheta - 2, nphi) ring-major -> map (ntheta, nphi),
 *   map[r][p] = dmap[r][p] + (-1)^spin dmap[2 ntheta - 2 - r][(p + nphi/2) % nphi] for 0 < r < ntheta - 1
 * (CPU_nufft_transformer.py:298-318). */
int nusht_adjoint_doubling_f64(const double* dmap, int ntheta, int nphi, int spin, double* map, nusht_stream_t s);
int nusht_adjoint_doubling_f32(const float* dmap, int ntheta, int nphi, int spin, float* map, nusht_stream_t s);
int nusht_adjoint_doubling_c128(const double* dmap, int ntheta, int nphi, int spin, double* map, nusht_stream_t s);

/* Deflected pointing.  Replaces cunusht/c/pointing.cu:228-255 (CUpointing_1Dto1D_lowmem, exported at
 * :510-520) and adds the rotation angle gamma of CPU_nufft_transformer.py:158-160.
 *   theta, phi0 (double[nring]), nph, ofs (int64[nring]): ring table; d_theta, d_phi (double[npix]):
 *   spin-1 deflection; out_theta, out_phi (double[npix]); out_gamma may be NULL.
 *   max_nph: host value, the largest entry of nph; npix: host value, the length of the pixel arrays.
 *   Pixels ofs[r] + k outside [0, npix) (or k >= max_nph) are skipped, so a ring table that does not
 *   match the arrays leaves pixels unwritten rather than writing out of bounds. */
int nusht_pointing_f64(const double* theta, const double* phi0, const int64_t* nph, const int64_t* ofs,
                       int nring, int64_t max_nph, int64_t npix, const double* d_theta, const double* d_phi,
                       double* out_theta, double* out_phi, double* out_gamma, nusht_stream_t s);

/* Determinant of the lensing magnification matrix per pixel.  Replaces the pointwise part of
 * cunusht/helper_GPU.py:113-161 (dlm2A); kappa, g1, g2, d1, d2, omega (omega may be NULL) are the
 * syntheses on the pointing grid of 0.5 sqrt(l(l+1)) dlm (spin 0), -0.5 sqrt((l-1)(l+2)) [dlm, dclm]
 * (spin 2), [dlm, dclm] (spin 1) and 0.5 sqrt(l(l+1)) dclm (spin 0).  exact != 0: the solid-angle
 * ratio of the deflection map; exact = 0: the reference's formula (shear sign of the |d|-dependent term
 * as at helper_GPU.py:126; equal through second order in |d|).  |d| -> 0 is handled by series. */
int nusht_magnification_f64(const double* kappa, const double* g1, const double* g2, const double* d1,
                            const double* d2, const double* omega, int64_t npix, int exact, double* A,
                            nusht_stream_t s);
/* out = fl[l] alm (ncomp blocks of nalm complex in the layout above; out may alias alm).  Replaces the
 * almxfl calls of helper_GPU.py:126-130 on the device. */
int nusht_almxfl_c128(const double* alm, int lmax, int ncomp, const double* fl, double* out, nusht_stream_t s);

/* ---------------------------------------------------------------- ring plan (pointing grid SHT)
 * Replaces the GL-grid SHTns transformer of cunusht/sht/GPU_sht_transformer.py:28-34 and its
 * synthesis_cupy (:56-62), adjoint_synthesis_cupy (:65-73: SHTns analys, the quadrature-weighted
 * adjoint; with iw = 1 / weight the exact transpose) and synthesis_der1_cupy (:88-96) together with
 * deflection.dlm2pointing (cunusht/deflection/GPU_nufft_transformer.py:64-104).
 * theta, phi0, nph, ofs, weight are HOST arrays of length nring.  Rings may differ in length
 * (1 <= nph <= 8320; HEALPix, thinned GL and other geometries of cunusht/geometry.py:194-297);
 * rings shorter than 2 lmax + 1 are aliased as exact sampling requires.  weight (per pixel of the
 * ring, as Geom.weight of cunusht/geometry.py:29-57) may be NULL; nusht_ringplan_analysis then fails. */
typedef struct {
  int lmax;
  int spin;             /* 0: scalar maps, 1: deflection / pointing, 2: polarisation */
  int nring;
  const double* theta;  /* host */
  const double* phi0;   /* host */
  const int64_t* nph;   /* host */
  const int64_t* ofs;   /* host */
  const double* weight; /* host, may be NULL */
} nusht_ring_desc;

int nusht_ringplan_create(const nusht_ring_desc* desc, nusht_stream_t s, nusht_ringplan_t* out);
int nusht_ringplan_destroy(nusht_ringplan_t p);
int64_t nusht_ringplan_npix(nusht_ringplan_t p);

typedef struct {
  int64_t npix;
  int nring;
  int ring_pairs;                /* rings after north/south pairing */
  int ring_dft;                  /* DFT of the longest ring: 0 direct mixed-radix, 1 Rader, 2 Bluestein, 3 split */
  int64_t legendre_pair_steps;   /* (l, ring pair) recursion steps one synthesis runs */
  int legendre_variant;          /* synthesis kernel variant chosen by the plan-time autotuner */
  float legendre_ms;             /* its measured time for one map */
  int ring_lengths;              /* distinct ring lengths (one DFT description each) */
  int64_t workspace_bytes;       /* device bytes the plan holds */
} nusht_ringplan_info;
/* Plan properties (no counterpart in the reference; reported for measurement). */
int nusht_ringplan_get_info(nusht_ringplan_t p, nusht_ringplan_info* info);
/* Number of Legendre synthesis kernel variants; plan creation times them and keeps the fastest unless the
 * environment variable NUSHT_LEGENDRE_VARIANT names one (no counterpart in the reference). */
int nusht_legendre_variants(void);
/* alm (ncomp x nalm) -> map0 (spin 0 map, or Q) and map1 (U; NULL for spin 0) */
int nusht_ringplan_synthesis(nusht_ringplan_t p, const double* alm, int ncomp, double* map0, double* map1,
                             nusht_stream_t s);
/* quadrature analysis: the adjoint of the synthesis applied to map * weight
 * (GPU_nufft_transformer.py:372-383 -> GPU_sht_transformer.py:65-73 SHTns analys;
 * CPU_sht_transformer.py:37-48 ducc adjoint_synthesis with apply_weights) */
int nusht_ringplan_analysis(nusht_ringplan_t p, const double* map0, const double* map1, double* alm, int ncomp,
                            nusht_stream_t s);
/* exact adjoint of the synthesis (no weights) */
int nusht_ringplan_adjoint(nusht_ringplan_t p, const double* map0, const double* map1, double* alm, int ncomp,
                           nusht_stream_t s);
/* spin-1 plans: gradient of a scalar field, (d/dtheta, d/dphi / sin theta) of f = sum alm Y_lm on the ring
 * grid.  Replaces synthesis_der1_cupy / SHTns synth_grad (cunusht/sht/GPU_sht_transformer.py:88-96). */
int nusht_sht_synth_grad_gl(nusht_ringplan_t p, const double* alm, double* dtheta, double* dphi, nusht_stream_t s);
/* spin-1 plans: dlm_scaled = dlm / sqrt(l(l+1)) -> deflected (theta, phi) and gamma (may be NULL) */
int nusht_ringplan_dlm2pointing(nusht_ringplan_t p, const double* dlm_scaled, double* ptg_theta, double* ptg_phi,
                                double* gamma, nusht_stream_t s);

/* ---------------------------------------------------------------- point plan
 * Exact synthesis (no nuFFT approximation) at arbitrary points, for few points.  Replaces the brute-force
 * evaluation of gclm2lenpixs (cunusht/deflection/CPU_nufft_transformer.py:531-558).  theta: HOST array of
 * npts colatitudes in [0, pi]; phi (and gamma, may be NULL) are device arrays given per call.  spin 0:
 * out0 = f; spin s: (out0, out1) = (Q, U), rotated by exp(i s gamma) when gamma != NULL. */
int nusht_pointplan_create(int lmax, int spin, int64_t npts, const double* theta, nusht_stream_t s,
                           nusht_pointplan_t* out);
int nusht_pointplan_destroy(nusht_pointplan_t p);
int nusht_pointplan_synthesis(nusht_pointplan_t p, const double* alm, int ncomp, const double* phi,
                              const double* gamma, double* out0, double* out1, nusht_stream_t s);
/* Allocation-free variant of nusht_pointplan_synthesis: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): no plan, no hipMalloc / hipFree, nothing allocated by any library routine.  theta, phi,
 * gamma (may be NULL), alm (ncomp x nalm complex, healpy layout), out0 / out1 are device arrays; theta in
 * [0, pi].  The ring coefficients live in one 32 MiB array per GPU inside the library (2^21 complex
 * values): (lmax + 1) x npts for spin 0, (2 lmax + 1) x npts for spin > 0 must not exceed 2^21 (e.g. 256
 * points at lmax 4095 spin 2, 1024 at lmax 1023), or NUSHT_ERR_UNSUPPORTED.  Every call on one GPU uses
 * that GPU's array: calls that overlap on one GPU (different streams, threads) race and give wrong
 * results; calls ordered on one stream are safe.  Same results as nusht_pointplan_synthesis (FP64,
 * different summation order). */
int nusht_pointplan_synthesis_noalloc(int lmax, int spin, int64_t npts, const double* theta, const double* alm,
                                      int ncomp, const double* phi, const double* gamma, double* out0, double* out1,
                                      nusht_stream_t s);

/* ---------------------------------------------------------------- nuSHT plan
 * Replaces GPU_cufinufft_transformer.__init__ / plan (cunusht/deflection/GPU_nufft_transformer.py:108-173,
 * :227-246): geometry, recursion seeds, FFT twiddles, ES kernel and deconvolution tables and all
 * work buffers are created here and stay resident until nusht_plan_destroy. */
typedef struct {
  int lmax;             /* mmax must equal lmax */
  int mmax;
  int spin;             /* 0, 1, 2 */
  int ncomp;            /* 1 (spin 0 or GRAD_ONLY) or 2 (E, B) */
  double epsilon;       /* target accuracy; FFT + nuFFT in FP32 when epsilon >= 1e-6 */
  int64_t npts_max;     /* largest point set passed to set_pts */
  double sigma;         /* fine-grid oversampling: > 0 given; 0 automatic (1.25, or 2 below eps 1e-8, as far
                           as the grid fits the in-LDS transforms); < 0 measured at plan creation (candidates
                           sigma 1.25, 1.5, 2 x spreading tile heights 8, 16, 32; the fastest type 2 + type 1
                           on npts_max synthetic points); 0 is measured too when NUSHT_NUFFT_AUTOTUNE=1 */
  int max_maps;         /* maps per nusht_nusht2d2_batch call the plan holds buffers for (1..8; 0 = 8) */
} nusht_plan_desc;

typedef struct {
  int ntheta;           /* equiangular rings incl. poles used internally */
  int ntheta_dfs;       /* 2 ntheta - 2 */
  int nfine;            /* fine grid size per axis */
  int nrow_fine;        /* fine theta rows actually formed */
  int kernel_width;
  int single_precision;
  double sigma;
  double beta;
  int64_t workspace_bytes;
  int64_t legendre_pair_steps;  /* (l, ring pair) recursion steps one synthesis runs: l-ranges in
                                   which a whole group of 64 ring pairs is below 2^-300 are skipped */
  int legendre_variant;  /* synthesis kernel variant chosen by the plan-time autotuner */
  float legendre_ms;     /* its measured time for one map */
  int nufft_tile_rows;   /* point tile (fine cells of first footprint cell): rows x cols */
  int nufft_tile_cols;
  int nufft_candidates;  /* nuFFT configurations timed at plan creation (0: not tuned) */
  float nufft_type2_ms;  /* the chosen one's type-2 and type-1 times on the synthetic points */
  float nufft_type1_ms;
  int nufft_spread_rows; /* 1: register row-window spreading (8 <= w <= 16), 0: LDS-atomic spreading */
} nusht_plan_info;

int nusht_plan_create(const nusht_plan_desc* desc, nusht_stream_t s, nusht_plan_t* out);
int nusht_plan_destroy(nusht_plan_t p);
int nusht_plan_get_info(nusht_plan_t p, nusht_plan_info* info);
/* Device bytes the plan holds (tables and work buffers; the same as workspace_bytes of the info), or -1 for
 * a null plan.  For HBM-budget admission; no counterpart in the reference, which allocates per call. */
int64_t nusht_plan_workspace_bytes(nusht_plan_t p);

/* Points (theta in [0, pi], any phi).  Replaces set_pts / cuFINUFFT setpts
 * (GPU_nufft_transformer.py:249-251).  Synchronizes the stream once (bounds check). */
int nusht_plan_set_pts(nusht_plan_t p, const double* theta, const double* phi, int64_t npts, nusht_stream_t s);
/* Same without the bounds check and without synchronizing (the setpts of cuFINUFFT checks
 * nothing either).  A point with theta outside [0, pi] beyond the kernel half-width, or NaN,
 * gives NaN in nusht2d2 and is ignored by nusht2d1.  For points computed on the device, e.g. by
 * nusht_ringplan_dlm2pointing. */
int nusht_plan_set_pts_async(nusht_plan_t p, const double* theta, const double* phi, int64_t npts,
                             nusht_stream_t s);

/* alm -> values at the points.  Replaces nusht2d2 (GPU_nufft_transformer.py:386-435).
 * spin 0: out0[npts] real; spin s: out0 = Q, out1 = U, rotated by exp(i s gamma) if gamma != NULL. */
int nusht_nusht2d2(nusht_plan_t p, const double* alm, double* out0, double* out1, const double* gamma,
                   nusht_stream_t s);
/* exact adjoint.  Replaces nusht2d1 (GPU_nufft_transformer.py:437-484). */
int nusht_nusht2d1(nusht_plan_t p, const double* in0, const double* in1, const double* gamma, double* alm,
                   nusht_stream_t s);

/* Lensing.  Replace gclm2lenmap (GPU_nufft_transformer.py:486-578) and lenmap2gclm (:580-673), with the
 * spin-s rotation of the CPU backend (CPU_nufft_transformer.py:356-368, 429-529) when polrot != 0.
 * ptg must be a spin-1 ring plan with the same lmax; its npix must not exceed npts_max. */
int nusht_gclm2lenmap(nusht_plan_t p, nusht_ringplan_t ptg, const double* gclm, const double* dlm_scaled,
                      double* lenmap0, double* lenmap1, int polrot, nusht_stream_t s);
int nusht_lenmap2gclm(nusht_plan_t p, nusht_ringplan_t ptg, const double* lenmap0, const double* lenmap1,
                      const double* dlm_scaled, double* gclm, int polrot, nusht_stream_t s);

/* Batched in-place 2D complex FFT of `batch` row-major n0 x n1 arrays (interleaved complex), forward
 * (direction -1, sign -1) or backward (direction +1, sign +1, unnormalised) as cuFFT C2C.  Replaces the
 * reference's _C2C / _iC2C (GPU_nufft_transformer.py:271-287).  Sides that factor into 2, 3, 5, 7, 13 may
 * reach 10240 (f64) / 20480 (f32); any other side (Rader, Bluestein or split transforms) up to 8320, except a
 * prime above 4160 whose n - 1 is not smooth.  The first call for a length makes its tables (device
 * allocations kept for the life of the process, one stream synchronization); later calls allocate nothing. */
int nusht_fft2d_c2c_f64(double* data, int n0, int n1, int batch, int direction, nusht_stream_t s);
int nusht_fft2d_c2c_f32(float* data, int n0, int n1, int batch, int direction, nusht_stream_t s);

/* Stages of the plan, for callers that compose them (the reference's _synthesis / _nuFFT2d2 / _nuFFT2d1 /
 * _adjoint_synthesis, GPU_nufft_transformer.py:254-310).  The intermediate G holds the Fourier coefficients
 * in phi of the synthesis on the plan's equiangular rings theta_j = pi j / (ntheta - 1), j = 0..ntheta-1
 * (ntheta from nusht_plan_info; the rings include both poles): G[r * ntheta + j] complex, rows r = m for
 * spin 0 (m = 0..lmax) and r = m + lmax for spin s (m = -lmax..lmax), so that the field on ring j is
 * f(phi) = sum_m G_m(theta_j) e^{i m phi} (real part 2 Re sum_{m>0} + G_0 for spin 0; Q + iU for spin s).
 *   nusht_sht_synth_cc:   alm -> G                  (SHT synthesis on the rings)
 *   nusht_sht_adjoint_cc: G -> alm                  (its exact adjoint)
 *   nusht_nufft_interp:   G -> values at the plan's points (doubling onto the torus, 2D FFT, ES
 *                         interpolation; rotation by exp(i s gamma) if gamma != NULL)
 *   nusht_nufft_spread:   values -> G               (exact adjoint of nusht_nufft_interp)
 * so nusht_nusht2d2 = synth_cc + nufft_interp and nusht_nusht2d1 = nufft_spread + adjoint_cc.
 * nusht_nufft_setpts is nusht_plan_set_pts (cuFINUFFT setpts, GPU_nufft_transformer.py:249-251). */
int nusht_sht_synth_cc(nusht_plan_t p, const double* alm, double* G, nusht_stream_t s);
int nusht_sht_adjoint_cc(nusht_plan_t p, const double* G, double* alm, nusht_stream_t s);
int nusht_nufft_setpts(nusht_plan_t p, const double* theta, const double* phi, int64_t npts, nusht_stream_t s);
int nusht_nufft_interp(nusht_plan_t p, const double* G, const double* gamma, double* out0, double* out1,
                       nusht_stream_t s);
int nusht_nufft_spread(nusht_plan_t p, const double* in0, const double* in1, const double* gamma, double* G,
                       nusht_stream_t s);

/* Batched synthesis: nmaps alm sets (stride alm_stride doubles between sets) on the plan's current
 * points, sharing one Legendre recursion per (m, ring); outputs at out0 + k * out_stride (and out1). */
int nusht_nusht2d2_batch(nusht_plan_t p, int nmaps, const double* alm, int64_t alm_stride, double* out0,
                         double* out1, int64_t out_stride, const double* gamma, nusht_stream_t s);

/* Batched lensing: nmaps gclm sets (stride alm_stride doubles) lensed by one deflection, i.e. nmaps
 * gclm2lenmap calls sharing dlm_scaled (GPU_nufft_transformer.py:486-578 per map), with the pointing and binning
 * done once and the synthesis of 3 or more maps on the matrix cores.  Outputs at lenmap0 + k * out_stride (and
 * lenmap1).  Everything runs in order on `s`.  Calls on one plan must not overlap. */
int nusht_gclm2lenmap_batch(nusht_plan_t p, nusht_ringplan_t ptg, int nmaps, const double* gclm,
                            int64_t alm_stride, const double* dlm_scaled, double* lenmap0, double* lenmap1,
                            int64_t out_stride, int polrot, nusht_stream_t s);

/* ---------------------------------------------------------------- batches of independent transforms
 * No counterpart in the reference (one call at a time on the legacy stream).  A batch runs jobs of
 * any lmax, spin and epsilon, each on its own plan(s): jobs that share a plan or ring plan run in
 * submission order on one internal stream, the other groups run concurrently on up to four internal
 * streams (default priority), forked from and joined back into `s`.  The call only enqueues work, so a
 * caller may capture it into a hipGraph on `s` and replay it.  All jobs are validated before any work is
 * enqueued.  Points of NUSHT_JOB_NUSHT2D2 / NUSHT_JOB_NUSHT2D1 jobs are the plan's current points. */
typedef enum {
  NUSHT_JOB_GCLM2LENMAP = 0,  /* gclm2lenmap: alm_in, dlm_scaled -> map_out0 (, map_out1) */
  NUSHT_JOB_LENMAP2GCLM = 1,  /* lenmap2gclm: map_in0 (, map_in1), dlm_scaled -> alm_out */
  NUSHT_JOB_NUSHT2D2 = 2,     /* nusht2d2: alm_in -> map_out0 (, map_out1) at the plan's points */
  NUSHT_JOB_NUSHT2D1 = 3      /* nusht2d1: map_in0 (, map_in1) at the plan's points -> alm_out */
} nusht_job_kind;

typedef struct {
  int kind;                   /* nusht_job_kind */
  nusht_plan_t plan;
  nusht_ringplan_t ring;      /* lensing jobs: spin-1 ring plan of the pointing grid (else NULL) */
  const double* dlm_scaled;   /* lensing jobs */
  int polrot;                 /* lensing jobs: rotate spin-s fields by exp(i s gamma) */
  const double* gamma;        /* nusht2d2 / nusht2d1 jobs: optional per-point rotation angle */
  const double* alm_in;
  double* alm_out;
  const double* map_in0;
  const double* map_in1;
  double* map_out0;
  double* map_out1;
} nusht_job;

int nusht_batch_execute(const nusht_job* jobs, int njobs, nusht_stream_t s);

/* ---------------------------------------------------------------- stage timing (no counterpart in the
 * reference, whose timer decorators synchronize the device around every stage: helper_GPU.py:10-51).
 * When enabled, instrumented launches are bracketed by hipEvents recorded on their own stream;
 * nusht_timing_count() synchronizes on those events and returns the number of stage names. */
int nusht_timing_enable(int on);
int nusht_timing_reset(void);
int nusht_timing_count(void);
int nusht_timing_get(int i, const char** name, double* total_ms, int64_t* calls);

#ifdef __cplusplus
}
#endif

#endif /* NUSHT_AMD_H */
