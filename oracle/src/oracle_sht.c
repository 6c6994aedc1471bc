/* FP64 ring Legendre transforms for the oracle (test infrastructure only).
 *
 * Restates, in plain C/OpenMP, the spin-s ring synthesis that the reference CPU
 * backend delegates to ducc0 synthesis_2d / adjoint_synthesis_2d
 * (cunusht/deflection/CPU_nufft_transformer.py:320-331) and that the GPU backend
 * delegates to SHTns synth / analys / synth_grad
 * (cunusht/sht/GPU_sht_transformer.py:56-96).
 *
 * Convention: sY_lm(theta,0) = (-1)^s sqrt((2l+1)/4pi) d^l_{m,-s}(theta);
 * spin 0 field f = sum a_lm Y_lm; spin s>0 field Q+iU = -sum (E+iB) sY_lm.
 * Output ring coefficients G_m(theta) so that f(theta,phi) = sum_m G_m e^{i m phi}.
 *
 * No north/south symmetry is used here on purpose: every ring runs its own
 * recursion, so this code shares no structure with the GPU kernels it checks.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SCALE_BITS 600

static long nalm_(int lmax) { return (long)(lmax + 1) * (lmax + 2) / 2; }
static long almidx_(int lmax, int l, int m) { return (long)m * (2 * lmax + 1 - m) / 2 + l; }

/* log|d^{l0}_{m,mp}(theta)| and sign, l0 = max(m,|mp|), m >= 0 */
static void log_start(int l0, int m, int mp, double theta, double *logv, double *sgn) {
    double c = cos(0.5 * theta), s = sin(0.5 * theta);
    double lc = log(fabs(c)), ls = log(fabs(s));
    double lg, sign;
    int pc, ps;
    if (l0 == m) {
        lg = 0.5 * (lgamma(2.0 * l0 + 1) - lgamma(l0 + mp + 1.0) - lgamma(l0 - mp + 1.0));
        pc = l0 + mp; ps = l0 - mp;
        sign = ((l0 - mp) & 1) ? -1.0 : 1.0;
    } else if (mp == l0) {
        lg = 0.5 * (lgamma(2.0 * l0 + 1) - lgamma(l0 + m + 1.0) - lgamma(l0 - m + 1.0));
        pc = l0 + m; ps = l0 - m;
        sign = 1.0;
    } else {
        lg = 0.5 * (lgamma(2.0 * l0 + 1) - lgamma(l0 + m + 1.0) - lgamma(l0 - m + 1.0));
        pc = l0 - m; ps = l0 + m;
        sign = ((m + l0) & 1) ? -1.0 : 1.0;
    }
    double lv = lg + 0.5 * log((2.0 * l0 + 1) / (4.0 * M_PI));
    if (pc > 0) lv += pc * lc;
    if (ps > 0) lv += ps * ls;
    if ((pc & 1) && c < 0) sign = -sign;
    if ((ps & 1) && s < 0) sign = -sign;
    *logv = lv;
    *sgn = sign;
}

static void split_scaled(double logv, double sgn, double *v, int *k) {
    if (!isfinite(logv)) { *v = 0.0; *k = -1000000; return; }
    double kk = ceil((logv / log(2.0) - SCALE_BITS / 2) / SCALE_BITS);
    if (kk > 0) kk = 0;
    *k = (int)kk;
    *v = sgn * exp(logv - kk * SCALE_BITS * log(2.0));
}

typedef struct { double *alpha, *b, *gamma; } coefs_t;

/* recursion lambda_{l+1} = (alpha_l x - mp/|s| * b_l) lambda_l - gamma_l lambda_{l-1}
 * for m' = -s uses +b_l, for m' = +s uses -b_l. */
static void fill_coefs(int lmax, int m, int s, coefs_t *c) {
    int l0 = m > s ? m : s;
    for (int l = l0; l < lmax; ++l) {
        if (l == 0) { c->alpha[l] = sqrt(3.0); c->b[l] = 0.0; c->gamma[l] = 0.0; continue; }
        double d1 = sqrt(((double)(l + 1) * (l + 1) - (double)m * m) * ((double)(l + 1) * (l + 1) - (double)s * s));
        double d0 = sqrt(((double)l * l - (double)m * m) * ((double)l * l - (double)s * s));
        double f = sqrt((2.0 * l + 3) * (2.0 * l + 1));
        c->alpha[l] = f * (l + 1) / d1;
        c->b[l] = f * (double)m * s / (l * d1);
        c->gamma[l] = sqrt((2.0 * l + 3) / (2.0 * l - 1)) * (l + 1) * d0 / (l * d1);
    }
}

/* Per (m, ring): values N_l d^l_{m,-s} (u) and N_l d^l_{m,+s} (v), l = l0..lmax, into ubuf/vbuf
 * (zero where below the representable range). */
static void ring_lambdas(int lmax, int m, int s, double theta, const coefs_t *c, double *ubuf, double *vbuf) {
    int l0 = m > s ? m : s;
    double x = cos(theta);
    double lvu, sgu, lvv, sgv;
    double u, v, up = 0.0, vp = 0.0;
    int ku, kv;
    log_start(l0, m, -s, theta, &lvu, &sgu);
    split_scaled(lvu, sgu, &u, &ku);
    if (s > 0) {
        log_start(l0, m, s, theta, &lvv, &sgv);
        split_scaled(lvv, sgv, &v, &kv);
    } else { v = 0; kv = -1000000; }
    const double thresh = ldexp(1.0, SCALE_BITS / 2), small = ldexp(1.0, -SCALE_BITS);
    for (int l = l0; l <= lmax; ++l) {
        ubuf[l] = (ku == 0) ? u : 0.0;
        if (s > 0) vbuf[l] = (kv == 0) ? v : 0.0;
        if (l == lmax) break;
        double ax = c->alpha[l] * x;
        double un = (ax + c->b[l]) * u - c->gamma[l] * up;
        up = u; u = un;
        if (ku < 0 && fabs(u) > thresh) { u *= small; up *= small; ku++; }
        if (s > 0) {
            double vn = (ax - c->b[l]) * v - c->gamma[l] * vp;
            vp = v; v = vn;
            if (kv < 0 && fabs(v) > thresh) { v *= small; vp *= small; kv++; }
        }
    }
}

/* Synthesis.  alm: ncomp x nalm complex (interleaved re,im).  spin 0: ncomp=1,
 * G shape (lmax+1) x nring complex (m = 0..lmax).  spin>0: ncomp in {1,2} (E[,B]),
 * G shape (2 lmax + 1) x nring complex, row = m + lmax. */
int oracle_sht_synth_rings(int lmax, int spin, int ncomp, const double *alm, int nring,
                           const double *theta, double *G) {
    if (spin < 0 || spin > 3 || lmax < 0 || nring <= 0) return 1;
    const long na = nalm_(lmax);
    const int nm = spin == 0 ? lmax + 1 : 2 * lmax + 1;
    memset(G, 0, sizeof(double) * 2 * (size_t)nm * nring);
    const double sgs = (spin & 1) ? -1.0 : 1.0;
#pragma omp parallel
    {
        coefs_t c;
        c.alpha = malloc(sizeof(double) * (lmax + 1));
        c.b = malloc(sizeof(double) * (lmax + 1));
        c.gamma = malloc(sizeof(double) * (lmax + 1));
        double *ub = malloc(sizeof(double) * (lmax + 1));
        double *vb = malloc(sizeof(double) * (lmax + 1));
#pragma omp for schedule(dynamic, 1)
        for (int m = 0; m <= lmax; ++m) {
            int l0 = m > spin ? m : spin;
            if (l0 > lmax) continue;
            fill_coefs(lmax, m, spin, &c);
            for (int i = 0; i < nring; ++i) {
                ring_lambdas(lmax, m, spin, theta[i], &c, ub, vb);
                if (spin == 0) {
                    double gr = 0, gi = 0;
                    for (int l = l0; l <= lmax; ++l) {
                        long id = almidx_(lmax, l, m);
                        gr += alm[2 * id] * ub[l];
                        gi += alm[2 * id + 1] * ub[l];
                    }
                    G[2 * ((size_t)m * nring + i)] = gr;
                    G[2 * ((size_t)m * nring + i) + 1] = gi;
                } else {
                    double pr = 0, pi = 0, nr = 0, ni = 0;
                    for (int l = l0; l <= lmax; ++l) {
                        long id = almidx_(lmax, l, m);
                        double er = alm[2 * id], ei = alm[2 * id + 1];
                        double br = ncomp > 1 ? alm[2 * (na + id)] : 0.0;
                        double bi = ncomp > 1 ? alm[2 * (na + id) + 1] : 0.0;
                        /* W = E + iB ; Wt = conj(E) + i conj(B) */
                        double wr = er - bi, wi = ei + br;
                        double tr = er + bi, ti = -ei + br;
                        pr += wr * ub[l]; pi += wi * ub[l];
                        nr += tr * vb[l]; ni += ti * vb[l];
                    }
                    /* G_{+m} = -(-1)^s sum W u ; G_{-m} = - sum Wt v */
                    size_t ip = (size_t)(m + lmax) * nring + i;
                    G[2 * ip] = -sgs * pr;
                    G[2 * ip + 1] = -sgs * pi;
                    if (m > 0) {
                        size_t in = (size_t)(lmax - m) * nring + i;
                        G[2 * in] = -nr;
                        G[2 * in + 1] = -ni;
                    }
                }
            }
        }
        free(c.alpha); free(c.b); free(c.gamma); free(ub); free(vb);
    }
    return 0;
}

/* Exact adjoint of oracle_sht_synth_rings under the real-field alm inner product. */
int oracle_sht_adjoint_rings(int lmax, int spin, int ncomp, const double *G, int nring,
                             const double *theta, double *alm) {
    if (spin < 0 || spin > 3 || lmax < 0 || nring <= 0) return 1;
    const long na = nalm_(lmax);
    memset(alm, 0, sizeof(double) * 2 * (size_t)na * ncomp);
    const double sgs = (spin & 1) ? -1.0 : 1.0;
#pragma omp parallel
    {
        coefs_t c;
        c.alpha = malloc(sizeof(double) * (lmax + 1));
        c.b = malloc(sizeof(double) * (lmax + 1));
        c.gamma = malloc(sizeof(double) * (lmax + 1));
        double *ub = malloc(sizeof(double) * (lmax + 1));
        double *vb = malloc(sizeof(double) * (lmax + 1));
        double *acc = malloc(sizeof(double) * 4 * (lmax + 1));
#pragma omp for schedule(dynamic, 1)
        for (int m = 0; m <= lmax; ++m) {
            int l0 = m > spin ? m : spin;
            if (l0 > lmax) continue;
            fill_coefs(lmax, m, spin, &c);
            memset(acc, 0, sizeof(double) * 4 * (lmax + 1));
            for (int i = 0; i < nring; ++i) {
                ring_lambdas(lmax, m, spin, theta[i], &c, ub, vb);
                if (spin == 0) {
                    double gr = G[2 * ((size_t)m * nring + i)], gi = G[2 * ((size_t)m * nring + i) + 1];
                    for (int l = l0; l <= lmax; ++l) { acc[4 * l] += ub[l] * gr; acc[4 * l + 1] += ub[l] * gi; }
                } else {
                    size_t ip = (size_t)(m + lmax) * nring + i, in = (size_t)(lmax - m) * nring + i;
                    double pr = G[2 * ip], pi = G[2 * ip + 1];
                    double qr = G[2 * in], qi = -G[2 * in + 1]; /* conj(G'_{-m}) */
                    for (int l = l0; l <= lmax; ++l) {
                        acc[4 * l] += ub[l] * pr; acc[4 * l + 1] += ub[l] * pi;
                        acc[4 * l + 2] += vb[l] * qr; acc[4 * l + 3] += vb[l] * qi;
                    }
                }
            }
            for (int l = l0; l <= lmax; ++l) {
                long id = almidx_(lmax, l, m);
                if (spin == 0) {
                    alm[2 * id] = acc[4 * l];
                    alm[2 * id + 1] = (m == 0) ? 0.0 : acc[4 * l + 1];
                } else {
                    /* X = (-1)^s sum G'_m u ; Y = sum conj(G'_{-m}) v
                     * E' = -(X + Y)/2 ; B' = i (X - Y)/2 */
                    double xr = sgs * acc[4 * l], xi = sgs * acc[4 * l + 1];
                    double yr = acc[4 * l + 2], yi = acc[4 * l + 3];
                    alm[2 * id] = -0.5 * (xr + yr);
                    alm[2 * id + 1] = -0.5 * (xi + yi);
                    if (ncomp > 1) {
                        alm[2 * (na + id)] = -0.5 * (xi - yi);
                        alm[2 * (na + id) + 1] = 0.5 * (xr - yr);
                    }
                }
            }
        }
        free(c.alpha); free(c.b); free(c.gamma); free(ub); free(vb); free(acc);
    }
    return 0;
}
