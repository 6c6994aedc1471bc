/* Host CPU backend of the nuSHT operators: C ABI of cunusht_amd/libnusht_cpu.so (C++17 + OpenMP).
 *
 * These entry points carry the numerical core of get_transformer(backend='CPU'), the counterpart of the
 * reference's CPU_DUCCnufft_transformer (cunusht/deflection/CPU_nufft_transformer.py:181-558).  The
 * reference calls ducc0 for each of them; the ducc0 call each one replaces is named below.  FFTs stay in
 * Python (scipy.fft), as the reference's ducc0.fft.c2c calls do (CPU_nufft_transformer.py:266-279).
 *
 * Conventions (shared with include/nusht_amd.h):
 *   alm     complex128, (ncomp, (lmax+1)(lmax+2)/2), mmax = lmax, index m (2 lmax + 1 - m) / 2 + l.
 *   G       complex128 ring coefficients, row-major (rows, nring): spin 0 rows m = 0..lmax, spin s > 0 rows
 *           m = -lmax..lmax (row m + lmax).  The field on ring r is sum_m G[m][r] e^{i m phi} (spin 0: the
 *           real field with G_{-m} = conj(G_m)); spin s > 0 gives Q + iU = -sum (E + iB) sY_lm.
 *   theta   radians in [0, pi]; all arrays are plain host pointers; nthreads <= 0 uses OpenMP's default.
 * Every function returns 0 on success and 1 on an invalid argument (nothing is written then).
 */
#ifndef NUSHT_CPU_H
#define NUSHT_CPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int nusht_cpu_version(void);

/* Ring coefficients of the spin-s synthesis on nring rings of colatitude theta.
 * Replaces ducc0.sht.experimental.synthesis_2d(..., geometry="CC") up to its phi transform
 * (CPU_nufft_transformer.py:323-325) and the spin-1 gradient synthesis of the pointing
 * (CPU_nufft_transformer.py:86-116, mode GRAD_ONLY when ncomp == 1 and spin > 0). */
int nusht_cpu_synth_rings(int lmax, int spin, int ncomp, const double* alm, int nring, const double* theta,
                          double* G, int nthreads);

/* Exact adjoint of nusht_cpu_synth_rings under the real-field alm inner product
 * (sum_l a_l0 b_l0 + 2 sum_{m>0} Re conj(a_lm) b_lm per component; rows m > 0 of a spin-0 G count twice).
 * Replaces ducc0.sht.experimental.adjoint_synthesis_2d after its phi transform (CPU_nufft_transformer.py:327-331). */
int nusht_cpu_adjoint_rings(int lmax, int spin, int ncomp, const double* G, int nring, const double* theta,
                            double* alm, int nthreads);

/* A window of rows of an nf1 x nf2 periodic fine grid and its ES kernel.  The window holds nrow rows from
 * torus row row_lo (complex128, row-major (nrow, nf2)); nrow == nf1 (with row_lo == 0) is the whole torus and
 * wraps, a smaller window drops rows outside it.  Kernel psi(z) = exp(beta (sqrt(1 - z^2) - 1)) over w cells,
 * z = 2 (distance in cells) / w, evaluated from per-cell polynomials fitted to max error ktol (a cell that
 * no polynomial of degree <= 24 fits that well, at most the two outer ones in practice, is evaluated exactly). */
typedef struct nusht_cpu_grid {
    int nf1, nf2;      /* torus size (theta, phi) */
    int row_lo, nrow;  /* the window of rows held */
    int w;             /* kernel width in cells, 2..32 */
    double beta;       /* kernel shape */
    double ktol;       /* max error of the piecewise-polynomial kernel */
} nusht_cpu_grid;

/* Degree and max error of the kernel polynomials for (w, beta, ktol) (no counterpart: diagnostics); cells
 * evaluated exactly are not counted in max_err. */
int nusht_cpu_es_kernel_info(int w, double beta, double ktol, int* degree, double* max_err);

/* ES-kernel interpolation out[p] = e^{i (mshift phi_p + spin gamma_p)} sum_{i,j} grid[i][j]
 * psi(theta_p - 2 pi (row_lo + i) / nf1) psi(phi_p - 2 pi j / nf2) (gamma may be NULL: no spin term).
 * Replaces ducc0.nufft.u2nu's interpolation stage and the rotation ducc0.misc.lensing_rotate
 * (CPU_nufft_transformer.py:346-368). */
int nusht_cpu_interp(const nusht_cpu_grid* g, const double* grid, int64_t npts, const double* theta,
                     const double* phi, int mshift, const double* gamma, int spin, double* out, int nthreads);

/* Exact adjoint of nusht_cpu_interp: grid += sum_p e^{-i (mshift phi_p + spin gamma_p)} in[p] psi psi (grid
 * zeroed by the caller).  Replaces ducc0.nufft.nu2u's spreading stage (CPU_nufft_transformer.py:336-344). */
int nusht_cpu_spread(const nusht_cpu_grid* g, const double* in, int64_t npts, const double* theta,
                     const double* phi, int mshift, const double* gamma, int spin, double* grid, int nthreads);

/* Deflected angles (tp, pp) of points (theta, phi) moved along the great circle by the deflection
 * (dt, dp) = (d_theta, d_phi / sin theta components), and the rotation angle gamma when gamma != NULL.
 * Replaces ducc0.misc.get_deflected_angles (CPU_nufft_transformer.py:118-166). */
int nusht_cpu_pointing(int64_t n, const double* theta, const double* phi, const double* dt, const double* dp,
                       double* tp, double* pp, double* gamma, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* NUSHT_CPU_H */
