/*
 * corr_field.c -- CPU restatement of make_corr_rand_3d (sim/corr_rand_3d.f90:
 * 191-386) for the oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference transforms with Ooura's rdft3d/rdft3dsort (vendor/fft2d/
 * fftsg3d.f:416-526), whose forward+inverse pair scales by n1*n2*n3/2, undone
 * by the factor 2/(n1 n2 n3) (C.f90:319).  The filter multiplies every
 * Hermitian-half coefficient by sqrt(|Re C_k|) (C.f90:303-310); C is real and
 * even, so the result is the normalised inverse DFT of DFT(x)*sqrt|C_k| and
 * does not depend on the transform's sign convention.  This file computes it
 * with a plain radix-2 complex FFT, so it agrees with the reference up to
 * floating-point rounding of the two transforms.
 */
#include "j1j2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int32_t nextpow2(int32_t n) { int32_t p = 1; while (p < n) p *= 2; return p; }  /* C.f90:350-357 */

static void fft1(double *re, double *im, int32_t n, int32_t stride, int sign)
{
    /* in-place iterative radix-2 DIT on a strided line */
    for (int32_t i = 1, j = 0; i < n; ++i) {
        int32_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i * stride]; re[i * stride] = re[j * stride]; re[j * stride] = t;
            t = im[i * stride]; im[i * stride] = im[j * stride]; im[j * stride] = t;
        }
    }
    for (int32_t len = 2; len <= n; len <<= 1) {
        double ang = sign * 2.0 * M_PI / (double)len;
        for (int32_t i = 0; i < n; i += len)
            for (int32_t k = 0; k < len / 2; ++k) {
                double wr = cos(ang * k), wi = sin(ang * k);
                int32_t a = (i + k) * stride, b = (i + k + len / 2) * stride;
                double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
    }
}

static void fft3(double *re, double *im, int32_t n1, int32_t n2, int32_t n3, int sign)
{
    for (int32_t k = 0; k < n3; ++k)
        for (int32_t j = 0; j < n2; ++j) fft1(re + (size_t)(k * n2 + j) * n1, im + (size_t)(k * n2 + j) * n1, n1, 1, sign);
    for (int32_t k = 0; k < n3; ++k)
        for (int32_t i = 0; i < n1; ++i) fft1(re + (size_t)k * n2 * n1 + i, im + (size_t)k * n2 * n1 + i, n2, n1, sign);
    for (int32_t j = 0; j < n2; ++j)
        for (int32_t i = 0; i < n1; ++i) fft1(re + (size_t)j * n1 + i, im + (size_t)j * n1 + i, n3, n1 * n2, sign);
}

double ora_erfcc(double x)
{
    /* Numerical Recipes erfcc, C.f90:31-43 */
    double z = fabs(x);
    double t = 1.0 / (1.0 + 0.5 * z);
    double r = t * exp(-z * z - 1.26551223 + t * (1.00002368 + t * (.37409196 + t *
               (.09678418 + t * (-.18628806 + t * (.27886807 + t * (-1.13520398 + t *
               (1.48851587 + t * (-.82215223 + t * .17087277)))))))));
    if (x < 0.0) r = 2.0 - r;
    return r;
}

int ora_corr_field(int32_t n1, int32_t n2, int32_t n3, double xi, int32_t seed,
                   int32_t do_uniform, double max_value, double *out)
{
    int32_t f1 = nextpow2(n1), f2 = nextpow2(n2), f3 = nextpow2(n3);
    size_t nf = (size_t)f1 * f2 * f3;
    double *cr = (double *)calloc(nf, sizeof(double)), *ci = (double *)calloc(nf, sizeof(double));
    double *xr = (double *)calloc(nf, sizeof(double)), *xim = (double *)calloc(nf, sizeof(double));
    if (!cr || !ci || !xr || !xim) { free(cr); free(ci); free(xr); free(xim); return -1; }
    ora_kiss k;
    ora_kiss_init(&k, seed);                             /* gkissinit, C.f90:256 */
    for (int32_t kk = 0; kk < f3; ++kk)                  /* C(r), C.f90:262-268, 362-384 */
        for (int32_t j = 0; j < f2; ++j)
            for (int32_t i = 0; i < f1; ++i) {
                int32_t dx = i < f1 - i ? i : f1 - i, dy = j < f2 - j ? j : f2 - j,
                        dz = kk < f3 - kk ? kk : f3 - kk;
                double r = sqrt((double)(dx * dx + dy * dy + dz * dz));
                double c;
                if (xi <= 0.0) c = (r == 0.0) ? 1.0 : 0.0;
                else c = exp(-r / xi);
                cr[((size_t)kk * f2 + j) * f1 + i] = c;
            }
    fft3(cr, ci, f1, f2, f3, -1);
    for (int32_t kk = 0; kk < f3; ++kk)                  /* white noise, C.f90:281-287 */
        for (int32_t j = 0; j < f2; ++j)
            for (int32_t i = 0; i < f1; ++i)
                xr[((size_t)kk * f2 + j) * f1 + i] = ora_kiss_gauss(&k);
    fft3(xr, xim, f1, f2, f3, -1);
    for (size_t t = 0; t < nf; ++t) {                    /* filter, C.f90:303-310 */
        double w = sqrt(fabs(cr[t]));
        xr[t] *= w; xim[t] *= w;
    }
    fft3(xr, xim, f1, f2, f3, +1);
    double scale = 1.0 / ((double)f1 * (double)f2 * (double)f3);
    for (int32_t kk = 0; kk < n3; ++kk)                  /* crop, C.f90:322-328 */
        for (int32_t j = 0; j < n2; ++j)
            for (int32_t i = 0; i < n1; ++i)
                out[((size_t)kk * n2 + j) * n1 + i] = xr[((size_t)kk * f2 + j) * f1 + i] * scale;
    if (do_uniform) {                                    /* C.f90:333-341 */
        size_t n = (size_t)n1 * n2 * n3;
        for (size_t t = 0; t < n; ++t)
            out[t] = (2.0 * (1.0 - 0.5 * ora_erfcc(out[t] / sqrt(2.0))) - 1.0) * max_value;
    }
    free(cr); free(ci); free(xr); free(xim);
    return 0;
}
