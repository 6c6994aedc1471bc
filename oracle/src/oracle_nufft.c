/* FP64 exponential-of-semicircle (ES) gridding for the oracle's nuFFT (test infrastructure only).
 *
 * The reference delegates its nuFFT to cuFINUFFT (cunusht/deflection/GPU_nufft_transformer.py:242-246,
 * 292-310) and ducc0.nufft (CPU_nufft_transformer.py:336-351).  Both implement the ES-kernel
 * method of Barnett, Magland & af Klinteberg (SIAM J. Sci. Comput. 2019):
 *   psi(z) = exp(beta (sqrt(1 - z^2) - 1)),  |z| < 1,  support w grid cells.
 * Only the gridding steps live here; the FFTs and deconvolution are done with numpy in
 * oracle/nufft.py.  Grid: n1 x n2 complex row-major, cell (i1,i2) at (2 pi i1/n1, 2 pi i2/n2).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double es(double z, double beta) {
    double t = 1.0 - z * z;
    return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

static inline void kernel_row(double x, int n, int w, double beta, int *i0, double *k) {
    double u = x * n / (2.0 * M_PI);
    int start = (int)ceil(u - 0.5 * w);
    *i0 = start;
    for (int a = 0; a < w; ++a) k[a] = es((start + a - u) * 2.0 / w, beta);
}

static inline int wrap(long i, int n) { long r = i % n; return (int)(r < 0 ? r + n : r); }

/* type-2 gridding: out_j = sum_{a,b} grid[i1+a][i2+b] psi psi */
int oracle_es_interp(int n1, int n2, const double *grid, long npts, const double *x1, const double *x2,
                     int w, double beta, double *out) {
    if (w > 64) return 1;
#pragma omp parallel for schedule(static)
    for (long j = 0; j < npts; ++j) {
        double k1[64], k2[64];
        int i1, i2;
        kernel_row(x1[j], n1, w, beta, &i1, k1);
        kernel_row(x2[j], n2, w, beta, &i2, k2);
        double sr = 0, si = 0;
        for (int a = 0; a < w; ++a) {
            const double *row = grid + 2 * (size_t)wrap(i1 + a, n1) * n2;
            double rr = 0, ri = 0;
            for (int b = 0; b < w; ++b) {
                int c = wrap(i2 + b, n2);
                rr += row[2 * c] * k2[b];
                ri += row[2 * c + 1] * k2[b];
            }
            sr += rr * k1[a];
            si += ri * k1[a];
        }
        out[2 * j] = sr;
        out[2 * j + 1] = si;
    }
    return 0;
}

/* type-1 gridding: grid[i1+a][i2+b] += c_j psi psi (grid zeroed by the caller).
 * Deterministic: points are bucketed by their first row; buckets of >= w rows are
 * processed in two colour phases so that no two threads touch the same row. */
int oracle_es_spread(int n1, int n2, double *grid, long npts, const double *x1, const double *x2,
                     const double *c, int w, double beta) {
    if (w > 64) return 1;
    /* an even number of buckets of >= w rows: a window spans at most two adjacent
     * buckets (the last bucket also takes the remainder rows and wraps to bucket 0) */
    int bs = w;
    int nb = n1 / bs;
    if (nb >= 2 && (nb & 1)) nb -= 1;
    if (nb < 2) nb = 1;
    long *cnt = calloc((size_t)nb + 1, sizeof(long));
    long *order = malloc(sizeof(long) * (size_t)(npts > 0 ? npts : 1));
    int *bucket = malloc(sizeof(int) * (size_t)(npts > 0 ? npts : 1));
    for (long j = 0; j < npts; ++j) {
        double u = x1[j] * n1 / (2.0 * M_PI);
        int r = wrap((long)ceil(u - 0.5 * w), n1);
        int b = r / bs;
        if (b >= nb) b = nb - 1;
        bucket[j] = b;
        cnt[b + 1]++;
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    long *pos = malloc(sizeof(long) * (size_t)(nb + 1));
    memcpy(pos, cnt, sizeof(long) * (size_t)(nb + 1));
    for (long j = 0; j < npts; ++j) order[pos[bucket[j]]++] = j;
    int ncol = (nb == 1) ? 1 : 2;
    for (int colour = 0; colour < ncol; ++colour) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int b = 0; b < nb; ++b) {
            if ((b % ncol) != colour) continue;
            double k1[64], k2[64];
            for (long q = cnt[b]; q < cnt[b + 1]; ++q) {
                long j = order[q];
                int i1, i2;
                kernel_row(x1[j], n1, w, beta, &i1, k1);
                kernel_row(x2[j], n2, w, beta, &i2, k2);
                double cr = c[2 * j], ci = c[2 * j + 1];
                for (int a = 0; a < w; ++a) {
                    double *row = grid + 2 * (size_t)wrap(i1 + a, n1) * n2;
                    double fr = cr * k1[a], fi = ci * k1[a];
                    for (int bb = 0; bb < w; ++bb) {
                        int cc = wrap(i2 + bb, n2);
                        row[2 * cc] += fr * k2[bb];
                        row[2 * cc + 1] += fi * k2[bb];
                    }
                }
            }
        }
    }
    free(cnt); free(order); free(bucket); free(pos);
    return 0;
}
