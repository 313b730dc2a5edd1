/*
 * CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 *
 * Complex FFT of arbitrary length for the oracle's restatement of the
 * reference CPU Poisson path (projection.f90:210-288), which calls FFTW-MPI
 * 2-D plans built at initialization.f90:305-308.  FFTW is absent here, so
 * the published algorithms are restated:
 *   - mixed-radix decimation-in-time Cooley-Tukey (radix 4, 2, 3, 5, and a
 *     generic odd radix for small primes), recursive, out of place;
 *   - Bluestein's chirp-z for prime factors larger than 64, embedding the
 *     length-p DFT in a power-of-two cyclic convolution.
 * Sign convention follows FFTW: forward uses exp(-2*pi*i*j*k/n), backward
 * uses exp(+2*pi*i*j*k/n), neither is normalised.
 * Twiddles are computed in long double and rounded once.
 */
#include "orc_fft.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI_L 3.141592653589793238462643383279502884L


struct orc_fft_plan {
    int n;
    int sign;
    int nfac;
    int fac[64];
    /* per-stage twiddles, one table of n entries is enough (w^k, k<n) */
    orc_cplx *w;
    /* Bluestein sub-plans per large prime factor (at most a few) */
    int nblue;
    int blue_p[8];
    int blue_m[8];
    orc_cplx *blue_chirp[8];   /* exp(sign*i*pi*k^2/p), k < p          */
    orc_cplx *blue_kern[8];    /* FFT of conj chirp, length m          */
    orc_fft_plan *blue_fwd[8]; /* length m forward                     */
    orc_fft_plan *blue_bwd[8]; /* length m backward                    */
    orc_cplx *scratch;         /* n */
    orc_cplx *bscratch;        /* max m */
};

static orc_cplx cmul(orc_cplx a, orc_cplx b) {
    orc_cplx r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

static orc_cplx twiddle(long num, long den, int sign) {
    /* exp(sign*2*pi*i*num/den) with exact integer argument reduction */
    long r = num % den;
    if (r < 0) r += den;
    long double a = 2.0L * ORC_PI_L * (long double)r / (long double)den;
    orc_cplx c;
    c.re = (double)cosl(a);
    c.im = (double)((long double)sign * sinl(a));
    return c;
}

orc_fft_plan *orc_fft_plan_create(int n, int sign) {
    orc_fft_plan *p = (orc_fft_plan *)calloc(1, sizeof(*p));
    if (!p) return NULL;
    p->n = n;
    p->sign = sign;
    int m = n;
    while (m > 1) {
        int f;
        if (m % 4 == 0) f = 4;
        else if (m % 2 == 0) f = 2;
        else if (m % 3 == 0) f = 3;
        else if (m % 5 == 0) f = 5;
        else {
            f = 7;
            while (m % f) f += 2;
        }
        p->fac[p->nfac++] = f;
        m /= f;
    }
    p->w = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) p->w[k] = twiddle(k, n, sign);
    p->scratch = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)(n > 0 ? n : 1));
    int maxm = 0;
    for (int i = 0; i < p->nfac; i++) {
        int f = p->fac[i];
        if (f > 64) {
            int known = 0;
            for (int b = 0; b < p->nblue; b++)
                if (p->blue_p[b] == f) known = 1;
            if (known) continue;
            int b = p->nblue++;
            int mm = 1;
            while (mm < 2 * f - 1) mm <<= 1;
            p->blue_p[b] = f;
            p->blue_m[b] = mm;
            p->blue_chirp[b] = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)f);
            for (long k = 0; k < f; k++)
                p->blue_chirp[b][k] = twiddle(k * k, 2L * f, sign);
            p->blue_fwd[b] = orc_fft_plan_create(mm, -1);
            p->blue_bwd[b] = orc_fft_plan_create(mm, +1);
            orc_cplx *kern = (orc_cplx *)calloc((size_t)mm, sizeof(orc_cplx));
            for (int k = 0; k < f; k++) {
                orc_cplx c = p->blue_chirp[b][k];
                c.im = -c.im;
                kern[k] = c;
                if (k) kern[mm - k] = c;
            }
            orc_cplx *kk = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)mm);
            orc_fft_exec(p->blue_fwd[b], kern, kk, 1, 1);
            p->blue_kern[b] = kk;
            free(kern);
            if (mm > maxm) maxm = mm;
        }
    }
    if (maxm) p->bscratch = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)(3 * maxm));
    return p;
}

void orc_fft_plan_destroy(orc_fft_plan *p) {
    if (!p) return;
    for (int b = 0; b < p->nblue; b++) {
        free(p->blue_chirp[b]);
        free(p->blue_kern[b]);
        orc_fft_plan_destroy(p->blue_fwd[b]);
        orc_fft_plan_destroy(p->blue_bwd[b]);
    }
    free(p->w);
    free(p->scratch);
    free(p->bscratch);
    free(p);
}

/* Length-f DFT of f points x[j*is] -> y[k*os] by Bluestein. */
static void blue_dft(orc_fft_plan *p, int f, const orc_cplx *x, long is, orc_cplx *y, long os) {
    int b = 0;
    while (p->blue_p[b] != f) b++;
    int mm = p->blue_m[b];
    orc_cplx *a = p->bscratch, *A = p->bscratch + mm, *c = p->bscratch + 2 * mm;
    memset(a, 0, sizeof(orc_cplx) * (size_t)mm);
    for (int k = 0; k < f; k++) a[k] = cmul(x[k * is], p->blue_chirp[b][k]);
    orc_fft_exec(p->blue_fwd[b], a, A, 1, 1);
    for (int k = 0; k < mm; k++) A[k] = cmul(A[k], p->blue_kern[b][k]);
    orc_fft_exec(p->blue_bwd[b], A, c, 1, 1);
    double inv = 1.0 / (double)mm;
    for (int k = 0; k < f; k++) {
        orc_cplx v = cmul(c[k], p->blue_chirp[b][k]);
        y[k * os].re = v.re * inv;
        y[k * os].im = v.im * inv;
    }
}

/*
 * Recursive DIT: out[k] for k<n of the length-n DFT of in[j*is], j<n.
 * n = fac[stage]*m.  First transform the f interleaved sub-sequences of
 * length m into out, then combine with butterflies of radix f.
 */
static void dit(orc_fft_plan *p, int stage, int n, const orc_cplx *in, long is, orc_cplx *out) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int f = p->fac[stage];
    int m = n / f;
    for (int r = 0; r < f; r++) dit(p, stage + 1, m, in + r * is, is * f, out + r * m);
    /* twiddle step: root of unity of order n is w[k*(N/n)] */
    long wstride = p->n / n;
    orc_cplx tmp[128];
    orc_cplx *t = tmp;
    orc_cplx *big = NULL;
    if (f > 64) big = (orc_cplx *)malloc(sizeof(orc_cplx) * (size_t)(2 * f));
    if (big) t = big;
    for (int k = 0; k < m; k++) {
        /* gather the f inputs of this butterfly, apply twiddles */
        for (int r = 0; r < f; r++) {
            orc_cplx v = out[r * m + k];
            if (r && k) v = cmul(v, p->w[((long)r * k * wstride) % p->n]);
            t[r] = v;
        }
        if (f == 2) {
            orc_cplx a = t[0], b = t[1];
            out[k].re = a.re + b.re;
            out[k].im = a.im + b.im;
            out[m + k].re = a.re - b.re;
            out[m + k].im = a.im - b.im;
        } else if (f == 4) {
            orc_cplx a0 = t[0], a1 = t[1], a2 = t[2], a3 = t[3];
            orc_cplx s02 = {a0.re + a2.re, a0.im + a2.im};
            orc_cplx d02 = {a0.re - a2.re, a0.im - a2.im};
            orc_cplx s13 = {a1.re + a3.re, a1.im + a3.im};
            orc_cplx d13 = {a1.re - a3.re, a1.im - a3.im};
            /* multiply d13 by sign*i */
            orc_cplx jd13;
            if (p->sign < 0) {
                jd13.re = d13.im;
                jd13.im = -d13.re;
            } else {
                jd13.re = -d13.im;
                jd13.im = d13.re;
            }
            out[k].re = s02.re + s13.re;
            out[k].im = s02.im + s13.im;
            out[m + k].re = d02.re + jd13.re;
            out[m + k].im = d02.im + jd13.im;
            out[2 * m + k].re = s02.re - s13.re;
            out[2 * m + k].im = s02.im - s13.im;
            out[3 * m + k].re = d02.re - jd13.re;
            out[3 * m + k].im = d02.im - jd13.im;
        } else if (f > 64) {
            blue_dft(p, f, t, 1, t + f, 1);
            for (int q = 0; q < f; q++) out[q * m + k] = t[f + q];
        } else {
            /* generic radix: direct DFT of length f with root w^(N/f) */
            long fstride = p->n / f;
            for (int q = 0; q < f; q++) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < f; r++) {
                    orc_cplx w = p->w[((long)q * r % f) * fstride];
                    sr += t[r].re * w.re - t[r].im * w.im;
                    si += t[r].re * w.im + t[r].im * w.re;
                }
                /* keep results in tmp2 to avoid overwriting inputs */
                t[f + q].re = sr;
                t[f + q].im = si;
            }
            for (int q = 0; q < f; q++) out[q * m + k] = t[f + q];
        }
    }
    if (big) free(big);
}

void orc_fft_exec(orc_fft_plan *p, const orc_cplx *in, orc_cplx *out, long istride, long ostride) {
    int n = p->n;
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    if (ostride == 1 && in != out) {
        dit(p, 0, n, in, istride, out);
    } else {
        dit(p, 0, n, in, istride, p->scratch);
        for (int k = 0; k < n; k++) out[k * ostride] = p->scratch[k];
    }
}

/* one-shot transform for tests: n complex values interleaved re,im */
void orc_fft_once(int n, int sign, const double *in, double *out) {
    orc_fft_plan *p = orc_fft_plan_create(n, sign);
    orc_fft_exec(p, (const orc_cplx *)in, (orc_cplx *)out, 1, 1);
    orc_fft_plan_destroy(p);
}
