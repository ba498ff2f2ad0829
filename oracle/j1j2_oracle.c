/*
 * j1j2_oracle.c -- CPU restatement of the J1-J2 Metropolis/PT core.
 *
 * TEST INFRASTRUCTURE ONLY (see j1j2_oracle.h).  Compiled with
 * -ffp-contract=off so every floating-point operation below is exactly the
 * one written; fma/fmaf are called explicitly where the engine rule uses them.
 */
#include "j1j2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * KISS05 -- J.f90:3614-3695 (rkiss05, kissinit); C.f90:48-85 (gkiss05,
 * gkissinit).  Fortran INTEGER(4) arithmetic wraps modulo 2^32 in practice;
 * ishft(v,-k) is a logical shift.  Done here in uint32_t.
 * ====================================================================== */
double ora_kiss_uniform(ora_kiss *k)
{
    uint32_t x = (uint32_t)k->x, y = (uint32_t)k->y, z = (uint32_t)k->z, w = (uint32_t)k->w;
    x = 69069u * x + 1327217885u;                       /* J.f90:3627 */
    y ^= y << 13; y ^= y >> 17; y ^= y << 5;            /* J.f90:3628 */
    z = 18000u * (z & 65535u) + (z >> 16);              /* J.f90:3629 */
    w = 30903u * (w & 65535u) + (w >> 16);              /* J.f90:3630 */
    uint32_t kiss = (x + y + (z << 16) + w) >> 1;        /* J.f90:3631 */
    k->x = (int32_t)x; k->y = (int32_t)y; k->z = (int32_t)z; k->w = (int32_t)w;
    return (double)(int32_t)kiss * 4.656612873077392578e-10;   /* J.f90:3632 */
}

static int32_t pm_step(int32_t idum)
{
    /* Park-Miller minimal standard via Schrage, J.f90:3662-3664 */
    const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
    int32_t kq = idum / IQ;
    int64_t v = (int64_t)IA * (int64_t)(idum - kq * IQ) - (int64_t)IR * kq;
    int32_t r = (int32_t)(uint32_t)(uint64_t)v;
    if (r < 0) r = (int32_t)((uint32_t)r + (uint32_t)IM);
    return r;
}

void ora_kiss_init(ora_kiss *k, int32_t iinit)
{
    const int32_t IM = 2147483647;
    /* J.f90:3658-3661: idum = abs(1099087573*idum), 32-bit wrap */
    int32_t idum = (int32_t)((uint32_t)1099087573u * (uint32_t)iinit);
    if (idum < 0) idum = (int32_t)(0u - (uint32_t)idum);
    if (idum == 0) idum = 1;
    if (idum >= IM) idum = IM - 1;
    int32_t v[4];
    for (int i = 0; i < 4; ++i) {             /* J.f90:3662-3691 */
        idum = pm_step(idum);
        v[i] = (idum < 1) ? idum + 1 : idum;
    }
    k->x = v[0]; k->y = v[1]; k->z = v[2]; k->w = v[3];
    k->iset = 0; k->gset = 0.0;               /* gkissinit, C.f90:76-85 */
    (void)ora_kiss_uniform(k);                /* J.f90:3693: one discarded draw */
}

double ora_kiss_gauss(ora_kiss *k)
{
    /* Marsaglia polar method, C.f90:48-74 */
    if (k->iset != 1) {
        double v1, v2, s;
        for (;;) {
            v1 = 2.0 * ora_kiss_uniform(k) - 1.0;
            v2 = 2.0 * ora_kiss_uniform(k) - 1.0;
            s = v1 * v1 + v2 * v2;
            if (s < 1.0 && s > 0.0) break;
        }
        double fac = sqrt(-2.0 * log(s) / s);
        k->gset = v1 * fac;
        k->iset = 1;
        return v2 * fac;
    }
    k->iset = 0;
    return k->gset;
}

void ora_kiss_fill(int32_t seed, int64_t n, double *out)
{
    ora_kiss k;
    ora_kiss_init(&k, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = ora_kiss_uniform(&k);
}

/* ======================================================================
 * Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11.  Round i uses key + i*(W0, W1).
 * ====================================================================== */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ======================================================================
 * Deterministic exponentials shared (as a rule, not as code) with the
 * engine.  Both sides use only +,-,*,fma,rint,ldexp, which IEEE-754 fixes.
 * ====================================================================== */
float ora_accept_threshold_f32(float y)
{
    /* 2^32 * 2^y.  y = n + f, n = rint(y), |f| <= 1/2; 2^f by a degree-6
     * polynomial (near-minimax on [-1/2, 1/2], relative error < 1.6e-8),
     * coefficients pre-scaled by 2^32 (exact power-of-two scaling).  No
     * clamp: for y >= 0 the result is >= 2^32, for y <= -183 it is 0 (the
     * engine's v_ldexp_f32 and ldexpf agree; n is bounded here only to keep
     * the int conversion defined, which leaves ldexp's 0 / inf unchanged). */
    const float c0 = 0x1.000000p+32f, c1 = 0x1.62e430p+31f, c2 = 0x1.ebfbdap+29f,
                c3 = 0x1.c6aed4p+27f, c4 = 0x1.3b2dbcp+25f, c5 = 0x1.5f456ap+22f,
                c6 = 0x1.41d332p+19f;
    float n = rintf(y);
    float f = y - n;
    float p = c6;
    p = fmaf(p, f, c5);
    p = fmaf(p, f, c4);
    p = fmaf(p, f, c3);
    p = fmaf(p, f, c2);
    p = fmaf(p, f, c1);
    p = fmaf(p, f, c0);
    if (n < -1000.0f) n = -1000.0f;
    if (n > 1000.0f) n = 1000.0f;
    return ldexpf(p, (int)n);
}

double ora_exp64(double x)
{
    /* Cody & Waite argument reduction x = k ln2 + r, |r| <= ln2/2, with
     * ln2 split hi/lo (hi has 21 trailing zero bits so k*hi is exact), then
     * e^r by its Taylor series to degree 13 in Horner form. */
    if (x < -700.0) return 0.0;
    if (x > 709.0) return INFINITY;
    const double LN2_HI = 0x1.62e42feep-1, LN2_LO = 0x1.a39ef35793c76p-33,
                 INV_LN2 = 0x1.71547652b82fep+0;
    double k = rint(x * INV_LN2);
    double r = fma(-k, LN2_HI, x);
    r = fma(-k, LN2_LO, r);
    double p = 0x1.6124613a86d09p-33;          /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);      /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);      /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);      /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);      /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);      /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);      /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);      /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);       /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);       /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);       /* 1/3!  */
    p = fma(p, r, 0x1.0000000000000p-1);       /* 1/2!  */
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

/* ======================================================================
 * Lattice geometry.  site = x + L*y + L*L*z (J.f90:539).  Neighbour tables
 * of Setup_neighbor_table (J.f90:532-586) computed arithmetically:
 * m1=+y m2=-y m3=+x m4=-x m5=+x+y m6=-x+y m7=+x-y m8=-x-y m9=+z m10=-z,
 * all periodic.
 * ====================================================================== */
static inline int32_t nb(int32_t L, int32_t H, int32_t site, int dx, int dy, int dz)
{
    int32_t L2 = L * L;
    int32_t x = site % L, y = (site / L) % L, z = site / L2;
    x = (x + dx + L) % L; y = (y + dy + L) % L; z = (z + dz + H) % H;
    return x + L * y + L2 * z;
}

int32_t ora_colour(int32_t L, int32_t site)
{
    int32_t L2 = L * L;
    int32_t x = site % L, y = (site / L) % L, z = site / L2;
    return (x & 1) | (((y + z) & 1) << 1);
}

int32_t ora_colour_index(int32_t L, int32_t site)
{
    int32_t L2 = L * L, half = L / 2;
    int32_t x = site % L, y = (site / L) % L, z = site / L2;
    return (z * half + (y >> 1)) * half + (x >> 1);
}

int ora_lattice_init(ora_lattice *lat, int32_t L, int32_t H, double J1, double deltaJ1,
                     double J2, double J3, const double *p)
{
    memset(lat, 0, sizeof(*lat));
    if (L < 2 || H < 2 || (L & 1) || (H & 1)) return -1;
    lat->L = L; lat->H = H; lat->N = L * L * H;
    lat->J1h = J1 + deltaJ1;   /* J.f90:47 */
    lat->J1v = J1 - deltaJ1;   /* J.f90:48 */
    lat->J2 = J2; lat->J3 = J3;
    int32_t N = lat->N;
    lat->p = (double *)malloc(sizeof(double) * N);
    lat->kc = (float *)malloc(sizeof(float) * N);
    lat->k2xp = (float *)malloc(sizeof(float) * N);
    lat->k2xm = (float *)malloc(sizeof(float) * N);
    lat->k2yp = (float *)malloc(sizeof(float) * N);
    lat->k2ym = (float *)malloc(sizeof(float) * N);
    lat->q = (int64_t *)malloc(sizeof(int64_t) * N);
    memcpy(lat->p, p, sizeof(double) * N);
    /* Forward bond couplings of the FP32 rule, rounded once from FP64:
     * Kx[i] = J1h + p_i + p_{i+x},  Ky[i] = J1v - p_i - p_{i+y}.  These are
     * the coefficients of s_i s_j in E (derived from J.f90:2512-2531). */
    float *Kx = (float *)malloc(sizeof(float) * N), *Ky = (float *)malloc(sizeof(float) * N);
    for (int32_t i = 0; i < N; ++i) {
        Kx[i] = (float)((lat->J1h + p[i]) + p[nb(L, H, i, 1, 0, 0)]);
        Ky[i] = (float)((lat->J1v - p[i]) - p[nb(L, H, i, 0, 1, 0)]);
    }
    for (int32_t i = 0; i < N; ++i) {
        float kxp = Kx[i], kxm = Kx[nb(L, H, i, -1, 0, 0)];
        float kyp = Ky[i], kym = Ky[nb(L, H, i, 0, -1, 0)];
        double sK = (((double)kxp + (double)kxm) + (double)kyp) + (double)kym;
        lat->kc[i] = (float)((-sK - 4.0 * J2) - 2.0 * J3);
        lat->k2xp[i] = 2.0f * kxp; lat->k2xm[i] = 2.0f * kxm;
        lat->k2yp[i] = 2.0f * kyp; lat->k2ym[i] = 2.0f * kym;
        lat->q[i] = llrint(p[i] * 0x1p32);
    }
    lat->j2x2 = (float)(2.0 * J2);
    lat->j3x2 = (float)(2.0 * J3);
    free(Kx); free(Ky);
    return 0;
}

void ora_lattice_free(ora_lattice *lat)
{
    free(lat->p); free(lat->kc); free(lat->k2xp); free(lat->k2xm);
    free(lat->k2yp); free(lat->k2ym); free(lat->q);
    memset(lat, 0, sizeof(*lat));
}

double ora_field_f64(const ora_lattice *lat, const int8_t *s, int32_t site, double *n_out)
{
    /* J.f90:2510-2540 with occu = jh = jv = 1, same operation order. */
    int32_t L = lat->L, H = lat->H;
    const double *p = lat->p;
    int32_t m1 = nb(L, H, site, 0, 1, 0), m2 = nb(L, H, site, 0, -1, 0);
    int32_t m3 = nb(L, H, site, 1, 0, 0), m4 = nb(L, H, site, -1, 0, 0);
    int32_t m5 = nb(L, H, site, 1, 1, 0), m6 = nb(L, H, site, -1, 1, 0);
    int32_t m7 = nb(L, H, site, 1, -1, 0), m8 = nb(L, H, site, -1, -1, 0);
    int32_t m9 = nb(L, H, site, 0, 0, 1), m10 = nb(L, H, site, 0, 0, -1);
    double field = 0.0, n = 0.0;
    field = field + (double)s[m1] * lat->J1v; n = n - (double)s[m1]; field = field - p[m1] * (double)s[m1];
    field = field + (double)s[m2] * lat->J1v; n = n - (double)s[m2]; field = field - p[m2] * (double)s[m2];
    field = field + (double)s[m3] * lat->J1h; n = n + (double)s[m3]; field = field + p[m3] * (double)s[m3];
    field = field + (double)s[m4] * lat->J1h; n = n + (double)s[m4]; field = field + p[m4] * (double)s[m4];
    field = field + (double)s[m5] * lat->J2;
    field = field + (double)s[m6] * lat->J2;
    field = field + (double)s[m7] * lat->J2;
    field = field + (double)s[m8] * lat->J2;
    field = field + (double)s[m9] * lat->J3;
    field = field + (double)s[m10] * lat->J3;
    field = field + p[site] * n;
    if (n_out) *n_out = n;
    return field;
}

float ora_field_f32(const ora_lattice *lat, const int8_t *s, int32_t site)
{
    /* Engine FP32 rule.  With b = (s+1)/2 in {0,1}:
     *   h = C_i + 2Kx[i] b(+x) + 2Kx[i-x] b(-x) + 2Ky[i] b(+y) + 2Ky[i-y] b(-y)
     *       + 2 J2 D + 2 J3 Z,
     * D = sum of the four diagonal b's, Z = sum of the two z b's,
     * C_i = -(sum of the four K) - 4 J2 - 2 J3, evaluated in this order. */
    int32_t L = lat->L, H = lat->H;
    float bpx = s[nb(L, H, site, 1, 0, 0)] > 0 ? 1.0f : 0.0f;
    float bmx = s[nb(L, H, site, -1, 0, 0)] > 0 ? 1.0f : 0.0f;
    float bpy = s[nb(L, H, site, 0, 1, 0)] > 0 ? 1.0f : 0.0f;
    float bmy = s[nb(L, H, site, 0, -1, 0)] > 0 ? 1.0f : 0.0f;
    int D = (s[nb(L, H, site, 1, 1, 0)] > 0) + (s[nb(L, H, site, -1, 1, 0)] > 0) +
            (s[nb(L, H, site, 1, -1, 0)] > 0) + (s[nb(L, H, site, -1, -1, 0)] > 0);
    int Z = (s[nb(L, H, site, 0, 0, 1)] > 0) + (s[nb(L, H, site, 0, 0, -1)] > 0);
    float h = lat->kc[site];
    h = fmaf(bpx, lat->k2xp[site], h);
    h = fmaf(bmx, lat->k2xm[site], h);
    h = fmaf(bpy, lat->k2yp[site], h);
    h = fmaf(bmy, lat->k2ym[site], h);
    h = fmaf((float)D, lat->j2x2, h);
    h = fmaf((float)Z, lat->j3x2, h);
    return h;
}

/* Decision rules.  r = the site's 32-bit Philox word.
 * FP32: y = -s (b2 h) with b2 = (float)(2 beta log2 e); accept iff
 * (float)r <= 2^32 2^y.  Every dE <= 0 move is taken (y >= 0 gives a
 * threshold >= 2^32, the largest value (float)r can round to). */
static int decide_f32(const ora_lattice *lat, const int8_t *s, int32_t site, float b2, uint32_t r)
{
    float h = ora_field_f32(lat, s, site);
    float g = b2 * h;
    float y = s[site] > 0 ? -g : g;
    return (float)r <= ora_accept_threshold_f32(y);
}

static int decide_f64(const ora_lattice *lat, const int8_t *s, int32_t site, double beta, uint32_t r)
{
    /* Metro_sweep acceptance, J.f90:2572-2604 */
    double field = ora_field_f64(lat, s, site, NULL);
    double u = (double)r * 0x1p-32;
    if (s[site] == 1) {
        if (field <= 0.0) return 1;
        return u < ora_exp64(-((2.0 * field) * beta));
    } else {
        if (field >= 0.0) return 1;
        return u < ora_exp64((2.0 * field) * beta);
    }
}

/* The engine's random words (no counterpart in the reference, whose
 * Metro_sweep draws KISS05 uniforms in site order).  Colour c of replica
 * `replica` in sweep `sweep`: the z-layers 2zp and 2zp+1 share the Philox
 * calls m = zp*GL + g (g = the site's 4-site group within its layer, GL =
 * ceil((L/2)^2 / 4) groups per layer); the site's word is r = v*2^16 + l,
 * v = one 16-bit half of word j of call (m, c | replica<<2, sweep lo, sweep
 * hi), l = the same half of word j of the call with ORA_LO_CTR added to the
 * second counter word; the even layer takes the high halves, the odd layer
 * the low halves.  out[] receives r for the colour indices qi = z*(L/2)^2 +
 * 4g + j of the call's sites (up to 8: j = 0..3 of the even layer, then of
 * the odd layer; n_out of them exist). */
#define ORA_LO_CTR 0x40000000u

static uint32_t half16(uint32_t w, int odd) { return odd ? (w & 0xFFFFu) : (w >> 16); }

int ora_call_words(int32_t L, uint32_t seed, uint32_t iconf, uint32_t replica, uint64_t sweep, int32_t c,
                   int32_t m, int32_t qi_out[8], uint32_t r_out[8])
{
    int32_t half = L / 2, lay = half * half, GL = (lay + 3) / 4;
    int32_t zp = m / GL, g = m % GL, n = 0;
    uint32_t key[2] = { seed, iconf };
    uint32_t c1 = (uint32_t)c | (replica << 2);
    uint32_t ch[4] = { (uint32_t)m, c1, (uint32_t)sweep, (uint32_t)(sweep >> 32) };
    uint32_t cl[4] = { (uint32_t)m, c1 | ORA_LO_CTR, (uint32_t)sweep, (uint32_t)(sweep >> 32) };
    uint32_t hi[4], lo[4];
    ora_philox4x32_10(ch, key, hi);
    ora_philox4x32_10(cl, key, lo);
    for (int odd = 0; odd < 2; ++odd)
        for (int j = 0; j < 4; ++j) {
            int32_t qz = 4 * g + j;
            if (qz >= lay) break;
            qi_out[n] = (2 * zp + odd) * lay + qz;
            r_out[n] = (half16(hi[j], odd) << 16) | half16(lo[j], odd);
            ++n;
        }
    return n;
}

void ora_sweep(const ora_lattice *lat, int8_t *s, int precision, double beta,
               uint32_t seed, uint32_t iconf, uint32_t replica, uint64_t sweep,
               int32_t *flipcounts, int64_t *shadow)
{
    int32_t L = lat->L, H = lat->H, half = L / 2, L2 = L * L;
    int32_t ncalls = (H / 2) * ((half * half + 3) / 4);
    float b2 = (float)(2.0 * beta * 0x1.71547652b82fep+0);
    for (int32_t c = 0; c < 4; ++c) {
        for (int32_t m = 0; m < ncalls; ++m) {
            int32_t qis[8];
            uint32_t rs[8];
            int n = ora_call_words(L, seed, iconf, replica, sweep, c, m, qis, rs);
            for (int k = 0; k < n; ++k) {
                int32_t qi = qis[k];
                int32_t xh = qi % half, rr = qi / half;
                int32_t z = rr / half, yh = rr % half;
                int32_t x = 2 * xh + (c & 1);
                int32_t y = 2 * yh + (((c >> 1) ^ z) & 1);
                int32_t site = x + L * y + L2 * z;
                int acc;
                if (precision == 0) acc = decide_f64(lat, s, site, beta, rs[k]);
                else acc = decide_f32(lat, s, site, b2, rs[k]);
                if (shadow) {
                    int other = (precision == 0) ? decide_f32(lat, s, site, b2, rs[k])
                                                 : decide_f64(lat, s, site, beta, rs[k]);
                    shadow[0] += 1;
                    shadow[1] += (other != acc);
                }
                if (acc) {
                    s[site] = (int8_t)(-s[site]);
                    if (flipcounts) flipcounts[site] += 1;
                }
            }
        }
    }
}

void ora_observables(const ora_lattice *lat, const int8_t *s, int64_t out[8])
{
    int32_t L = lat->L, H = lat->H, N = lat->N;
    int64_t M = 0, SX = 0, SY = 0, Bx = 0, By = 0, Bd = 0, Bz = 0, Eq = 0;
    for (int32_t i = 0; i < N; ++i) {
        int32_t x = i % L, y = (i / L) % L;
        int64_t si = s[i];
        M += si;
        SX += (x & 1) ? -si : si;
        SY += (y & 1) ? -si : si;
        int64_t spx = s[nb(L, H, i, 1, 0, 0)], smx = s[nb(L, H, i, -1, 0, 0)];
        int64_t spy = s[nb(L, H, i, 0, 1, 0)], smy = s[nb(L, H, i, 0, -1, 0)];
        Bx += si * spx;
        By += si * spy;
        Bd += si * s[nb(L, H, i, 1, 1, 0)] + si * s[nb(L, H, i, -1, 1, 0)];
        Bz += si * s[nb(L, H, i, 0, 0, 1)];
        int64_t l = si * (spx + smx - spy - smy);     /* J.f90:1017-1020 */
        Eq += lat->q[i] * l;
    }
    out[0] = M; out[1] = SX; out[2] = SY; out[3] = Bx; out[4] = By;
    out[5] = Bd; out[6] = Bz; out[7] = Eq;
}

double ora_energy_total(const ora_lattice *lat, const int64_t obs[8])
{
    /* E = -J1h Bx - J1v By - J2 Bd - J3 Bz - sum_i p_i l_i  (equals
     * 1/2 sum_i (-s_i h_i) of J.f90:2828-2838) */
    double ej = ((lat->J1h * (double)obs[3]) + (lat->J1v * (double)obs[4])) +
                ((lat->J2 * (double)obs[5]) + (lat->J3 * (double)obs[6]));
    return -ej - (double)obs[7] * 0x1p-32;
}

void ora_measurement_ref(const ora_lattice *lat, const int8_t *s, double out[4])
{
    /* J.f90:2807-2842, literal */
    int32_t L = lat->L, L2 = L * L, H = lat->H;
    double imag = 0, ien = 0, stripx = 0, stripy = 0;
    for (int32_t iz = 0; iz < H; ++iz)
        for (int32_t is = 0; is < L2; ++is) {
            int32_t site = is + L2 * iz;
            imag = imag + s[site];
            double field = ora_field_f64(lat, s, site, NULL);
            ien = ien - s[site] * field;
            stripx = stripx + s[site] * ((is % L) % 2 ? -1.0 : 1.0);
            stripy = stripy + s[site] * ((is / L) % 2 ? -1.0 : 1.0);
        }
    double N = lat->N;
    out[0] = imag / N; out[1] = (0.5 * ien) / N; out[2] = stripx / N; out[3] = stripy / N;
}

double ora_nbar_ref(const ora_lattice *lat, const int8_t *s)
{
    /* J.f90:1028-1048 */
    int32_t L = lat->L, H = lat->H;
    double sum = 0.0;
    for (int32_t i = 0; i < lat->N; ++i) {
        double l = (double)(s[i] * s[nb(L, H, i, 1, 0, 0)]) - (double)(s[i] * s[nb(L, H, i, 0, 1, 0)]) +
                   (double)(s[i] * s[nb(L, H, i, -1, 0, 0)]) - (double)(s[i] * s[nb(L, H, i, 0, -1, 0)]);
        sum = sum + l;
    }
    return sum / (double)lat->N;
}

void ora_initial_spins(int32_t L, int32_t H, int32_t start, int32_t straxis,
                       int32_t seed, int8_t *s)
{
    /* Initialize_spin_array, J.f90:1179-1209 */
    int32_t N = L * L * H;
    if (start == 0) {
        ora_kiss k;
        ora_kiss_init(&k, seed);
        for (int32_t i = 0; i < N; ++i) s[i] = ora_kiss_uniform(&k) < 0.5 ? 1 : -1;
    } else if (start == 1) {
        for (int32_t i = 0; i < N; ++i) s[i] = 1;
    } else {
        for (int32_t i = 0; i < N; ++i) {
            int32_t x = i % L, y = (i / L) % L;
            s[i] = (straxis == 0) ? ((y & 1) ? -1 : 1) : ((x & 1) ? -1 : 1);
        }
    }
}

int ora_ladder(int32_t mode, int32_t ntemp, double tmin, double tmax, double dt,
               double tmid, double tfrac, double *t)
{
    /* Setup_temperature_ladder, J.f90:928-962 */
    int32_t ninterval = ntemp - 1;
    if (mode == 1 && ninterval > 0) {
        int32_t nlow = (int32_t)lround(ninterval * tfrac);
        if (nlow > ninterval - 1) nlow = ninterval - 1;
        if (nlow < 1) nlow = 1;
        int32_t nhigh = ninterval - nlow;
        t[0] = tmin;
        for (int32_t it = 1; it <= nlow; ++it)
            t[it] = tmin + (tmid - tmin) * (double)it / (double)nlow;
        for (int32_t it = 1; it <= nhigh; ++it)
            t[nlow + it] = tmid + (tmax - tmid) * (double)it / (double)nhigh;
    } else if (mode == 0 || mode == 1) {
        if (ntemp <= 1) t[0] = tmin;
        else
            for (int32_t it = 1; it <= ntemp; ++it)
                t[it - 1] = tmin + (tmax - tmin) * (double)(it - 1) / (double)(ntemp - 1);
    } else {
        for (int32_t it = 1; it <= ntemp; ++it) t[it - 1] = tmin + (double)(it - 1) * dt;
    }
    return 0;
}

/* ======================================================================
 * PT engine restatement.  Cycle structure of Pt_cycle / Pt_attempt_swaps /
 * Pt_measurement_loop (J.f90:1282-1338, 1437-1524), engine RNG mapping:
 *   sweep draw: Philox key (seed, iconf), counter (group, colour|rep<<2,
 *               sweep lo, sweep hi), word = colour index & 3
 *   swap draw : counter (lower slot, 0x80000000, cycle lo, cycle hi), word 0
 * ====================================================================== */
int ora_pt_init(ora_pt *pt, int32_t L, int32_t H, double J1, double deltaJ1, double J2, double J3,
                const double *p, int32_t nrep, const double *temps, int32_t ptswapever,
                int32_t ptmeasindex, int32_t bins, int32_t precision, int32_t local,
                uint32_t seed, uint32_t iconf, const int8_t *spins0, int64_t series_cap,
                int32_t count_flips, int32_t shadow)
{
    memset(pt, 0, sizeof(*pt));
    if (ora_lattice_init(&pt->lat, L, H, J1, deltaJ1, J2, J3, p)) return -1;
    int32_t N = pt->lat.N;
    pt->nrep = nrep; pt->ptswapever = ptswapever; pt->ptmeasindex = ptmeasindex;
    pt->bins = bins; pt->precision = precision; pt->local = local;
    pt->seed = seed; pt->iconf = iconf;
    pt->spins = (int8_t *)malloc((size_t)nrep * N);
    memcpy(pt->spins, spins0, (size_t)nrep * N);
    pt->rep_at_temp = (int32_t *)malloc(sizeof(int32_t) * nrep);
    pt->temp_of_rep = (int32_t *)malloc(sizeof(int32_t) * nrep);
    pt->beta = (double *)malloc(sizeof(double) * nrep);
    for (int32_t i = 0; i < nrep; ++i) {
        pt->rep_at_temp[i] = i; pt->temp_of_rep[i] = i;
        pt->beta[i] = 1.0 / temps[i];
    }
    int32_t np = nrep > 1 ? nrep - 1 : 1;
    pt->att = (int64_t *)calloc(np, sizeof(int64_t));
    pt->acc = (int64_t *)calloc(np, sizeof(int64_t));
    pt->sums = (double *)calloc((size_t)nrep * ORA_NSUMS, sizeof(double));
    pt->best = (double *)malloc(sizeof(double) * nrep);
    pt->final_en = (double *)calloc(nrep, sizeof(double));
    for (int32_t i = 0; i < nrep; ++i) pt->best[i] = 1.7976931348623157e308;
    pt->hist = (int64_t *)calloc((size_t)nrep * bins, sizeof(int64_t));
    pt->hmin = (double *)calloc(nrep, sizeof(double));
    pt->hmax = (double *)calloc(nrep, sizeof(double));
    pt->hbin = (double *)calloc(nrep, sizeof(double));
    pt->vcold = (uint8_t *)calloc(nrep, 1);
    pt->vhot = (uint8_t *)calloc(nrep, 1);
    pt->locspin = (int64_t *)calloc(N, sizeof(int64_t));
    pt->locnem = (int64_t *)calloc(N, sizeof(int64_t));
    if (count_flips) pt->flips = (int32_t *)calloc((size_t)nrep * N, sizeof(int32_t));
    pt->do_shadow = shadow;
    pt->obs = (int64_t *)calloc((size_t)nrep * 8, sizeof(int64_t));
    pt->etot = (double *)calloc(nrep, sizeof(double));
    pt->series_cap = series_cap;
    if (series_cap > 0) pt->series = (double *)calloc((size_t)nrep * 5 * series_cap, sizeof(double));
    pt->swap_phase = 1;
    return 0;
}

void ora_pt_free(ora_pt *pt)
{
    ora_lattice_free(&pt->lat);
    free(pt->spins); free(pt->rep_at_temp); free(pt->temp_of_rep); free(pt->beta);
    free(pt->att); free(pt->acc); free(pt->sums); free(pt->best); free(pt->final_en);
    free(pt->hist); free(pt->hmin); free(pt->hmax); free(pt->hbin); free(pt->vcold);
    free(pt->vhot); free(pt->locspin); free(pt->locnem); free(pt->flips); free(pt->obs);
    free(pt->etot); free(pt->series);
    memset(pt, 0, sizeof(*pt));
}

void ora_pt_begin_phase(ora_pt *pt) { pt->swap_phase = 1; }

void ora_pt_reset_measurement(ora_pt *pt)
{
    /* the sums, extremes, histogram and visited flags zeroed at the start of
     * Pt_measurement_loop (J.f90:1375-1399, 1402-1403), Reset_local_nematic_
     * average (J.f90:1000-1006), and the LOCAL series of Measurement_loop */
    int32_t nrep = pt->nrep, N = pt->lat.N;
    memset(pt->sums, 0, sizeof(double) * (size_t)nrep * ORA_NSUMS);
    for (int32_t i = 0; i < nrep; ++i) pt->best[i] = 1.7976931348623157e308;
    memset(pt->final_en, 0, sizeof(double) * (size_t)nrep);
    memset(pt->hist, 0, sizeof(int64_t) * (size_t)nrep * pt->bins);
    memset(pt->hmin, 0, sizeof(double) * (size_t)nrep);
    memset(pt->hmax, 0, sizeof(double) * (size_t)nrep);
    memset(pt->hbin, 0, sizeof(double) * (size_t)nrep);
    memset(pt->vcold, 0, (size_t)nrep);
    memset(pt->vhot, 0, (size_t)nrep);
    memset(pt->locspin, 0, sizeof(int64_t) * (size_t)N);
    memset(pt->locnem, 0, sizeof(int64_t) * (size_t)N);
    pt->nmeasure = 0;
    pt->locsamples = 0;
    pt->series_len = 0;
}

static void swap_pass(ora_pt *pt)
{
    /* Pt_attempt_swaps, J.f90:1303-1338, with the engine's draw */
    int32_t start = (pt->swap_phase == 1) ? 0 : 1;
    uint32_t key[2] = { pt->seed, pt->iconf };
    for (int32_t sl = start; sl < pt->nrep - 1; sl += 2) {
        int32_t rlo = pt->rep_at_temp[sl], rhi = pt->rep_at_temp[sl + 1];
        pt->att[sl] += 1;
        double delta = (pt->beta[sl] - pt->beta[sl + 1]) * (pt->etot[rlo] - pt->etot[rhi]);
        uint32_t ctr[4] = { (uint32_t)sl, 0x80000000u, (uint32_t)pt->cycle,
                            (uint32_t)(pt->cycle >> 32) };
        uint32_t r4[4];
        ora_philox4x32_10(ctr, key, r4);
        double u = (double)r4[0] * 0x1p-32;
        if (delta >= 0.0 || u < ora_exp64(delta)) {
            pt->acc[sl] += 1;
            pt->rep_at_temp[sl] = rhi;
            pt->rep_at_temp[sl + 1] = rlo;
            pt->temp_of_rep[rlo] = sl + 1;
            pt->temp_of_rep[rhi] = sl;
        }
    }
}

static void replica_values(const ora_pt *pt, int32_t r, double v[5])
{
    const int64_t *o = pt->obs + 8 * r;
    double N = (double)pt->lat.N;
    v[0] = (double)o[0] / N;                 /* m    */
    v[1] = pt->etot[r] / N;                  /* e    */
    v[2] = (double)o[1] / N;                 /* strx */
    v[3] = (double)o[2] / N;                 /* stry */
    v[4] = (double)(2 * (o[3] - o[4])) / N;  /* nbar = sum l / N */
}

void ora_pt_cycle(ora_pt *pt, int measure)
{
    int32_t N = pt->lat.N;
    for (int32_t sl = 0; sl < pt->nrep; ++sl) {
        int32_t r = pt->rep_at_temp[sl];
        for (int32_t k = 0; k < pt->ptswapever; ++k)
            ora_sweep(&pt->lat, pt->spins + (size_t)r * N, pt->precision, pt->beta[sl],
                      pt->seed, pt->iconf, (uint32_t)r, pt->sweep + (uint64_t)k,
                      pt->flips ? pt->flips + (size_t)sl * N : NULL,
                      pt->do_shadow ? pt->shadow : NULL);
    }
    pt->sweep += (uint64_t)pt->ptswapever;
    for (int32_t r = 0; r < pt->nrep; ++r) {
        ora_observables(&pt->lat, pt->spins + (size_t)r * N, pt->obs + 8 * r);
        pt->etot[r] = ora_energy_total(&pt->lat, pt->obs + 8 * r);
    }
    if (!pt->local) swap_pass(pt);
    if (measure) {
        if (!pt->local)
            for (int32_t r = 0; r < pt->nrep; ++r) {      /* J.f90:1441-1442 */
                if (pt->temp_of_rep[r] == 0) pt->vcold[r] = 1;
                if (pt->temp_of_rep[r] == pt->nrep - 1) pt->vhot[r] = 1;
            }
        /* Accumulate_local_nematic_average, J.f90:1009-1025, for the replica
         * at slot PTMEASINDEX (J.f90:1477) */
        {
            int32_t r = pt->rep_at_temp[pt->ptmeasindex - 1];
            const int8_t *s = pt->spins + (size_t)r * N;
            int32_t L = pt->lat.L, H = pt->lat.H;
            for (int32_t i = 0; i < N; ++i) {
                pt->locspin[i] += s[i];
                pt->locnem[i] += s[i] * (s[nb(L, H, i, 1, 0, 0)] - s[nb(L, H, i, 0, 1, 0)] +
                                         s[nb(L, H, i, -1, 0, 0)] - s[nb(L, H, i, 0, -1, 0)]);
            }
            pt->locsamples += 1;
        }
        for (int32_t sl = 0; sl < pt->nrep; ++sl) {       /* J.f90:1480-1524 */
            int32_t r = pt->rep_at_temp[sl];
            double v[5];
            replica_values(pt, r, v);
            if (pt->local) {
                if (pt->series_len < pt->series_cap) {
                    double *sr = pt->series + (size_t)sl * 5 * pt->series_cap;
                    for (int k = 0; k < 5; ++k) sr[(size_t)k * pt->series_cap + pt->series_len] = v[k];
                }
                continue;
            }
            double am = v[0], ae = v[1], ax = v[2], ay = v[3], an = v[4];
            double ax2 = ax * ax, ay2 = ay * ay, m2 = am * am, e2 = ae * ae, n2 = an * an;
            double str2 = ax2 + ay2;
            double *S = pt->sums + (size_t)sl * ORA_NSUMS;
            S[0] += am;            S[1] += fabs(am);      S[2] += m2;          S[3] += m2 * m2;
            S[4] += ae;            S[5] += e2;            S[6] += fabs(ax);    S[7] += fabs(ay);
            S[8] += ax2;           S[9] += ax2 * ax2;     S[10] += ay2;        S[11] += ay2 * ay2;
            S[12] += str2;         S[13] += str2 * str2;  S[14] += sqrt(str2); S[15] += fabs(ay2 - ax2);
            S[16] += an;           S[17] += fabs(an);     S[18] += n2;         S[19] += n2 * n2;
            if (ae < pt->best[sl]) pt->best[sl] = ae;
            pt->final_en[sl] = ae;
            if (pt->hbin[sl] <= 0.0) {
                pt->hmin[sl] = ae - (fabs(ae) + 1.0);
                pt->hmax[sl] = ae + (fabs(ae) + 1.0);
                pt->hbin[sl] = (pt->hmax[sl] - pt->hmin[sl]) / (double)pt->bins;
            }
            double tq = (ae - pt->hmin[sl]) / pt->hbin[sl];
            int32_t idx;
            if (tq < 0.0) idx = 1;
            else if (tq >= (double)pt->bins) idx = pt->bins;
            else idx = 1 + (int32_t)tq;
            if (idx < 1) idx = 1;
            if (idx > pt->bins) idx = pt->bins;
            pt->hist[(size_t)sl * pt->bins + idx - 1] += 1;
        }
        if (pt->local) pt->series_len += 1;
        pt->nmeasure += 1;
    }
    pt->cycle += 1;
    pt->swap_phase = (pt->swap_phase == 1) ? 2 : 1;
}

void ora_derive(const double *S, int64_t nm_i, double beta, int32_t N_i,
                double best, double final_en, double o[49])
{
    /* J.f90:1581-1684 */
    double nm = (double)nm_i, N = (double)N_i;
    double mag2 = S[2] / nm, mag4 = S[3] / nm, absmag = S[1] / nm;
    double en = S[4] / nm, en2 = S[5] / nm;
    double tsx = S[6] / nm, tsy = S[7] / nm;
    double strx2 = S[8] / nm, strx4 = S[9] / nm, stry2 = S[10] / nm, stry4 = S[11] / nm;
    double str2 = S[12] / nm, str4 = S[13] / nm, absstr = S[14] / nm, nem = S[15] / nm;
    double nbar = S[16] / nm, nbarabs = S[17] / nm, nbar2 = S[18] / nm, nbar4 = S[19] / nm;
    double stripediff = tsx * tsx - tsy * tsy;
    double suscabs = ((mag2 - absmag * absmag) * N) * beta;
    double susstrip = ((str2 - absstr * absstr) * N) * beta;
    double UL = mag2 > 0.0 ? 1.0 - mag4 / (3.0 * (mag2 * mag2)) : 0.0;
    double US = str2 > 0.0 ? 2.0 - str4 / (str2 * str2) : 0.0;
    double USX = strx2 > 0.0 ? 1.0 - strx4 / (3.0 * (strx2 * strx2)) : 0.0;
    double USY = stry2 > 0.0 ? 1.0 - stry4 / (3.0 * (stry2 * stry2)) : 0.0;
    double UBN = nbar2 > 0.0 ? 1.0 - nbar4 / (3.0 * (nbar2 * nbar2)) : 0.0;
    double sph = ((en2 - en * en) * N) * (beta * beta);
    o[0] = absmag;   o[1] = mag2;    o[2] = mag4;    o[3] = suscabs;  o[4] = suscabs * suscabs;
    o[5] = UL;       o[6] = UL * UL; o[7] = absmag * absmag;
    o[8] = absmag > 0.0 ? log(absmag) : log(2.2250738585072014e-308);
    o[9] = en;       o[10] = en * en; o[11] = sph;   o[12] = sph * sph; o[13] = absstr;
    o[14] = tsx;     o[15] = tsy;    o[16] = str2;   o[17] = str4;    o[18] = susstrip;
    o[19] = susstrip * susstrip;     o[20] = US;     o[21] = US * US; o[22] = nem;
    o[23] = nem * nem; o[24] = absstr * absstr;      o[25] = strx2;   o[26] = strx4;
    o[27] = USX;     o[28] = USX * USX; o[29] = stry2; o[30] = stry4; o[31] = USY;
    o[32] = USY * USY; o[33] = absstr; o[34] = stripediff; o[35] = best; o[36] = final_en;
    o[37] = nbar;    o[38] = nbar * nbar; o[39] = nbar2; o[40] = nbar4; o[41] = UBN;
    o[42] = UBN * UBN; o[43] = nbar; o[44] = nbar2; o[45] = nbar4; o[46] = UBN;
    o[47] = nbarabs; o[48] = nbarabs;
}

void ora_pt_basefields(const ora_pt *pt, double *out)
{
    double o[49];
    for (int32_t sl = 0; sl < pt->nrep; ++sl) {
        ora_derive(pt->sums + (size_t)sl * ORA_NSUMS, pt->nmeasure, pt->beta[sl], pt->lat.N,
                   pt->best[sl], pt->final_en[sl], o);
        for (int f = 0; f < 49; ++f) out[(size_t)f * pt->nrep + sl] = o[f];
    }
}

void ora_pt_hist(const ora_pt *pt, double *out)
{
    for (int32_t sl = 0; sl < pt->nrep; ++sl)
        for (int32_t b = 0; b < pt->bins; ++b)
            out[(size_t)sl * pt->bins + b] =
                (double)pt->hist[(size_t)sl * pt->bins + b] / (double)pt->nmeasure;
}

void ora_process_local(const double *mag, const double *en, const double *strx,
                       const double *stry, const double *nbar, int64_t nmess,
                       double beta, int32_t N, int32_t bins, double out49[49],
                       double *hist, double *hmin_o, double *hbin_o)
{
    /* Process_data, J.f90:2219-2474 (NMESS > 0 branch) */
    double S[ORA_NSUMS] = { 0 };
    double hmin = en[0], hmax = en[0];
    for (int64_t i = 1; i < nmess; ++i) {
        if (en[i] < hmin) hmin = en[i];
        if (en[i] > hmax) hmax = en[i];
    }
    double besten = hmin, finalen = en[nmess - 1];
    if (hmax > hmin) {
        double hr = hmax - hmin;
        hmin = hmin - 0.50 * hr;
        hmax = hmax + 0.50 * hr;
    }
    if (hmax <= hmin) { hmin -= 1.0e-6; hmax += 1.0e-6; }
    double hbin = (hmax - hmin) / (double)bins;
    for (int32_t b = 0; b < bins; ++b) hist[b] = 0.0;
    for (int64_t i = 0; i < nmess; ++i) {
        double am = mag[i], ax = strx[i], ay = stry[i], an = nbar[i], ae = en[i];
        double ax2 = ax * ax, ay2 = ay * ay, m2 = am * am, n2 = an * an;
        S[0] += am;
        S[6] += fabs(ax); S[8] += ax2; S[9] += ax2 * ax2;
        S[7] += fabs(ay); S[10] += ay2; S[11] += ay2 * ay2;
        S[12] += ax2 + ay2; S[13] += (ax2 + ay2) * (ax2 + ay2);
        S[14] += sqrt(ax2 + ay2);
        S[1] += fabs(am); S[2] += m2; S[3] += m2 * m2;
        S[15] += fabs(ay2 - ax2);
        S[16] += an; S[17] += fabs(an); S[18] += n2; S[19] += n2 * n2;
        S[4] += ae; S[5] += ae * ae;
        long idx = lround((ae - hmin) / hbin);
        if (idx >= 1 && idx <= bins) hist[idx - 1] += 1.0;
    }
    for (int32_t b = 0; b < bins; ++b) hist[b] = hist[b] / (double)nmess;
    ora_derive(S, nmess, beta, N, besten, finalen, out49);
    *hmin_o = hmin; *hbin_o = hbin;
}

int ora_enumerate(const ora_lattice *lat, double beta, double out[8])
{
    int32_t N = lat->N, L = lat->L, H = lat->H;
    if (N > 24) return -1;
    int8_t *s = (int8_t *)malloc(N);
    int64_t ns = (int64_t)1 << N;
    double *E = (double *)malloc(sizeof(double) * ns);
    double emin = 1e300;
    for (int64_t st = 0; st < ns; ++st) {
        for (int32_t i = 0; i < N; ++i) s[i] = ((st >> i) & 1) ? 1 : -1;
        double pl = 0.0;
        int64_t o[8];
        ora_observables(lat, s, o);
        for (int32_t i = 0; i < N; ++i)
            pl += lat->p[i] * (double)(s[i] * (s[nb(L, H, i, 1, 0, 0)] + s[nb(L, H, i, -1, 0, 0)] -
                                               s[nb(L, H, i, 0, 1, 0)] - s[nb(L, H, i, 0, -1, 0)]));
        E[st] = -((lat->J1h * o[3] + lat->J1v * o[4]) + (lat->J2 * o[5] + lat->J3 * o[6])) - pl;
        if (E[st] < emin) emin = E[st];
    }
    double Z = 0, se = 0, se2 = 0, sam = 0, sm2 = 0, snb = 0, sx2 = 0, sy2 = 0;
    for (int64_t st = 0; st < ns; ++st) {
        for (int32_t i = 0; i < N; ++i) s[i] = ((st >> i) & 1) ? 1 : -1;
        int64_t o[8];
        ora_observables(lat, s, o);
        double w = exp(-beta * (E[st] - emin));
        double e = E[st] / N, m = (double)o[0] / N, nbv = (double)(2 * (o[3] - o[4])) / N;
        double sx = (double)o[1] / N, sy = (double)o[2] / N;
        Z += w; se += w * e; se2 += w * e * e; sam += w * fabs(m); sm2 += w * m * m;
        snb += w * nbv; sx2 += w * sx * sx; sy2 += w * sy * sy;
    }
    out[0] = se / Z; out[1] = se2 / Z; out[2] = sam / Z; out[3] = sm2 / Z;
    out[4] = snb / Z; out[5] = sx2 / Z; out[6] = sy2 / Z; out[7] = log(Z) - beta * emin;
    free(s); free(E);
    return 0;
}
