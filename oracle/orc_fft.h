/* CPU-ORACLE TEST INFRASTRUCTURE — not part of the product.
 * Arbitrary-length complex FFT used by the oracle's Poisson restatement. */
#ifndef ORC_FFT_H
#define ORC_FFT_H

typedef struct {
    double re, im;
} orc_cplx;

typedef struct orc_fft_plan orc_fft_plan;

/* sign = -1 forward (exp(-2 pi i jk/n)), +1 backward; unnormalised.
 * A plan owns scratch memory: use one plan per thread. */
orc_fft_plan *orc_fft_plan_create(int n, int sign);
void orc_fft_plan_destroy(orc_fft_plan *p);
/* out[k*ostride] = sum_j in[j*istride] w^(jk); in and out may alias. */
void orc_fft_exec(orc_fft_plan *p, const orc_cplx *in, orc_cplx *out, long istride, long ostride);
void orc_fft_once(int n, int sign, const double *in, double *out);

#endif
