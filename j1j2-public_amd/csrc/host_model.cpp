// host_model.cpp -- host-side model pieces (see host_model.h).
#include "host_model.h"

#include <cmath>
#include <complex>
#include <vector>

namespace j1j2 {

// ---------------------------------------------------------------- KISS05
Kiss05::Kiss05(int32_t seed) : iset(0), gset(0.0)
{
    // kissinit (J1J2_6tv.f90:3635-3695): idum = |1099087573*seed| (32-bit
    // wrap), four Park-Miller steps (Schrage) give x, y, z, w, then one draw.
    const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
    int32_t idum = (int32_t)((uint32_t)1099087573u * (uint32_t)seed);
    if (idum < 0) idum = (int32_t)(0u - (uint32_t)idum);
    if (idum == 0) idum = 1;
    if (idum >= IM) idum = IM - 1;
    uint32_t v[4];
    for (int i = 0; i < 4; ++i) {
        const int32_t kq = idum / IQ;
        const int64_t t = (int64_t)IA * (idum - kq * IQ) - (int64_t)IR * kq;
        idum = (int32_t)(uint32_t)(uint64_t)t;
        if (idum < 0) idum = (int32_t)((uint32_t)idum + (uint32_t)IM);
        v[i] = (uint32_t)(idum < 1 ? idum + 1 : idum);
    }
    x = v[0]; y = v[1]; z = v[2]; w = v[3];
    (void)uniform();
}

double Kiss05::uniform()
{
    // rkiss05 (J1J2_6tv.f90:3614-3632): LCG + xorshift + two MWC, >> 1, * 2^-31
    x = 69069u * x + 1327217885u;
    y ^= y << 13;
    y ^= y >> 17;
    y ^= y << 5;
    z = 18000u * (z & 65535u) + (z >> 16);
    w = 30903u * (w & 65535u) + (w >> 16);
    const uint32_t kiss = (x + y + (z << 16) + w) >> 1;
    return (double)kiss * 4.656612873077392578e-10;
}

double Kiss05::gauss()
{
    // Marsaglia polar method with one cached deviate (corr_rand_3d.f90:48-74)
    if (iset == 1) {
        iset = 0;
        return gset;
    }
    double v1, v2, s;
    do {
        v1 = 2.0 * uniform() - 1.0;
        v2 = 2.0 * uniform() - 1.0;
        s = v1 * v1 + v2 * v2;
    } while (!(s < 1.0 && s > 0.0));
    const double fac = std::sqrt(-2.0 * std::log(s) / s);
    gset = v1 * fac;
    iset = 1;
    return v2 * fac;
}

// ---------------------------------------------------------------- FFT
namespace {

using cplx = std::complex<double>;

// Radix-2 complex FFT along one axis of a 3D array, with a twiddle table.
struct Axis {
    int n;
    std::vector<cplx> tw;     // e^{sign 2 pi i k / n}, k < n/2
    std::vector<int> rev;
    Axis(int n_, int sign) : n(n_), tw(n_ / 2 > 0 ? n_ / 2 : 1), rev(n_)
    {
        for (int k = 0; k < n / 2; ++k) {
            const double a = sign * 2.0 * M_PI * (double)k / (double)n;
            tw[k] = cplx(std::cos(a), std::sin(a));
        }
        int lg = 0;
        while ((1 << lg) < n) ++lg;
        for (int i = 0; i < n; ++i) {
            int r = 0;
            for (int b = 0; b < lg; ++b)
                if (i & (1 << b)) r |= 1 << (lg - 1 - b);
            rev[i] = r;
        }
    }
    void run(cplx *a, long stride, std::vector<cplx> &buf) const
    {
        for (int i = 0; i < n; ++i) buf[rev[i]] = a[i * stride];
        for (int len = 2; len <= n; len <<= 1) {
            const int step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; ++k) {
                    const cplx t = buf[i + k + len / 2] * tw[k * step];
                    buf[i + k + len / 2] = buf[i + k] - t;
                    buf[i + k] += t;
                }
        }
        for (int i = 0; i < n; ++i) a[i * stride] = buf[i];
    }
};

void fft3d(std::vector<cplx> &a, int n1, int n2, int n3, int sign)
{
    const Axis a1(n1, sign), a2(n2, sign), a3(n3, sign);
    std::vector<cplx> buf(std::max(n1, std::max(n2, n3)));
    for (int k = 0; k < n3; ++k)
        for (int j = 0; j < n2; ++j) a1.run(&a[((size_t)k * n2 + j) * n1], 1, buf);
    for (int k = 0; k < n3; ++k)
        for (int i = 0; i < n1; ++i) a2.run(&a[(size_t)k * n2 * n1 + i], n1, buf);
    for (int j = 0; j < n2; ++j)
        for (int i = 0; i < n1; ++i) a3.run(&a[(size_t)j * n1 + i], (long)n1 * n2, buf);
}

int next_pow2(int n)
{
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

double erfcc(double x)
{
    // Numerical Recipes erfc approximation used by the reference (corr_rand_3d.f90:31-43)
    const double z = std::fabs(x);
    const double t = 1.0 / (1.0 + 0.5 * z);
    double r = t * std::exp(-z * z - 1.26551223 +
                            t * (1.00002368 + t * (.37409196 + t * (.09678418 + t * (-.18628806 +
                            t * (.27886807 + t * (-1.13520398 + t * (1.48851587 +
                            t * (-.82215223 + t * .17087277)))))))));
    if (x < 0.0) r = 2.0 - r;
    return r;
}

}  // namespace

int corr_field(int32_t n1, int32_t n2, int32_t n3, double xi, int32_t seed, bool map_uniform,
               double max_value, double *out)
{
    // Fourier filtering on the grid padded to powers of two: field =
    // IDFT(DFT(noise) * sqrt|DFT(C)|) / (f1 f2 f3), C(r) = exp(-|r|/xi) with
    // minimum-image distances; cropped to n1 x n2 x n3; optional erf map to
    // U(-max, max).  The reference's Ooura transform pair (scale 2/(n1 n2 n3))
    // is the same normalised pair.
    const int f1 = next_pow2(n1), f2 = next_pow2(n2), f3 = next_pow2(n3);
    const size_t nf = (size_t)f1 * f2 * f3;
    std::vector<cplx> c(nf), x(nf);
    for (int k = 0; k < f3; ++k)
        for (int j = 0; j < f2; ++j)
            for (int i = 0; i < f1; ++i) {
                const int dx = std::min(i, f1 - i), dy = std::min(j, f2 - j), dz = std::min(k, f3 - k);
                const double r = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
                const double v = xi <= 0.0 ? (r == 0.0 ? 1.0 : 0.0) : std::exp(-r / xi);
                c[((size_t)k * f2 + j) * f1 + i] = cplx(v, 0.0);
            }
    Kiss05 rng(seed);                       // gkissinit(seed)
    for (size_t t = 0; t < nf; ++t) x[t] = cplx(rng.gauss(), 0.0);   // i fastest
    fft3d(c, f1, f2, f3, -1);
    fft3d(x, f1, f2, f3, -1);
    for (size_t t = 0; t < nf; ++t) x[t] *= std::sqrt(std::fabs(c[t].real()));
    fft3d(x, f1, f2, f3, +1);
    const double scale = 1.0 / ((double)f1 * (double)f2 * (double)f3);
    for (int k = 0; k < n3; ++k)
        for (int j = 0; j < n2; ++j)
            for (int i = 0; i < n1; ++i) {
                double v = x[((size_t)k * f2 + j) * f1 + i].real() * scale;
                if (map_uniform) v = (2.0 * (1.0 - 0.5 * erfcc(v / std::sqrt(2.0))) - 1.0) * max_value;
                out[((size_t)k * n2 + j) * n1 + i] = v;
            }
    return 0;
}

void initial_spins(int32_t L, int32_t H, int32_t start, int32_t straxis, int32_t seed, int64_t skip, int8_t *s)
{
    const int64_t N = (int64_t)L * L * H;
    if (start == 0) {
        Kiss05 rng(seed);
        for (int64_t i = 0; i < skip; ++i) (void)rng.uniform();
        for (int64_t i = 0; i < N; ++i) s[i] = rng.uniform() < 0.5 ? 1 : -1;
    } else if (start == 1) {
        for (int64_t i = 0; i < N; ++i) s[i] = 1;
    } else {
        for (int64_t i = 0; i < N; ++i) {
            const int64_t x = i % L, y = (i / L) % L;
            const int64_t par = straxis == 0 ? y : x;    // STRAXIS X alternates along y
            s[i] = (par & 1) ? -1 : 1;
        }
    }
}

void temperature_ladder(int32_t mode, int32_t ntemp, double tmin, double tmax, double dt,
                        double tmid, double tfrac, double *t)
{
    const int32_t nint = ntemp - 1;
    if (mode == 1 && nint > 0) {
        int32_t nlow = (int32_t)std::lround(nint * tfrac);
        nlow = std::max(1, std::min(nint - 1, nlow));
        const int32_t nhigh = nint - nlow;
        t[0] = tmin;
        for (int32_t i = 1; i <= nlow; ++i) t[i] = tmin + (tmid - tmin) * (double)i / (double)nlow;
        for (int32_t i = 1; i <= nhigh; ++i) t[nlow + i] = tmid + (tmax - tmid) * (double)i / (double)nhigh;
    } else if (mode == 0 || mode == 1) {
        if (ntemp <= 1) t[0] = tmin;
        else
            for (int32_t i = 0; i < ntemp; ++i) t[i] = tmin + (tmax - tmin) * (double)i / (double)(ntemp - 1);
    } else {
        for (int32_t i = 0; i < ntemp; ++i) t[i] = tmin + (double)i * dt;
    }
}

void derive_basefields(const double *S, int64_t nmeasure, double beta, int32_t N_, double best,
                       double final_en, double o[49])
{
    const double nm = (double)nmeasure, N = (double)N_;
    const double absmag = S[1] / nm, mag2 = S[2] / nm, mag4 = S[3] / nm;
    const double en = S[4] / nm, en2 = S[5] / nm;
    const double tsx = S[6] / nm, tsy = S[7] / nm;
    const double strx2 = S[8] / nm, strx4 = S[9] / nm, stry2 = S[10] / nm, stry4 = S[11] / nm;
    const double str2 = S[12] / nm, str4 = S[13] / nm, absstr = S[14] / nm, nem = S[15] / nm;
    const double nbar = S[16] / nm, nbarabs = S[17] / nm, nbar2 = S[18] / nm, nbar4 = S[19] / nm;
    const double stripediff = tsx * tsx - tsy * tsy;
    const double suscabs = ((mag2 - absmag * absmag) * N) * beta;
    const double susstrip = ((str2 - absstr * absstr) * N) * beta;
    const double UL = mag2 > 0.0 ? 1.0 - mag4 / (3.0 * (mag2 * mag2)) : 0.0;
    const double US = str2 > 0.0 ? 2.0 - str4 / (str2 * str2) : 0.0;
    const double USX = strx2 > 0.0 ? 1.0 - strx4 / (3.0 * (strx2 * strx2)) : 0.0;
    const double USY = stry2 > 0.0 ? 1.0 - stry4 / (3.0 * (stry2 * stry2)) : 0.0;
    const double UBN = nbar2 > 0.0 ? 1.0 - nbar4 / (3.0 * (nbar2 * nbar2)) : 0.0;
    const double sph = ((en2 - en * en) * N) * (beta * beta);
    const double logm = absmag > 0.0 ? std::log(absmag) : std::log(2.2250738585072014e-308);
    const double v[49] = {
        absmag, mag2, mag4, suscabs, suscabs * suscabs, UL, UL * UL, absmag * absmag, logm,
        en, en * en, sph, sph * sph, absstr, tsx, tsy, str2, str4, susstrip, susstrip * susstrip,
        US, US * US, nem, nem * nem, absstr * absstr, strx2, strx4, USX, USX * USX, stry2, stry4,
        USY, USY * USY, absstr, stripediff, best, final_en, nbar, nbar * nbar, nbar2, nbar4, UBN,
        UBN * UBN, nbar, nbar2, nbar4, UBN, nbarabs, nbarabs};
    for (int f = 0; f < 49; ++f) o[f] = v[f];
}

void process_local_series(const double *mag, const double *en, const double *strx,
                          const double *stry, const double *nbar, int64_t nmess, double beta,
                          int32_t N, int32_t bins, double out49[49], double *hist, double *hmin_o,
                          double *hbin_o)
{
    double S[20] = {0};
    double hmin = en[0], hmax = en[0];
    for (int64_t i = 1; i < nmess; ++i) {
        hmin = std::min(hmin, en[i]);
        hmax = std::max(hmax, en[i]);
    }
    const double besten = hmin, finalen = en[nmess - 1];
    if (hmax > hmin) {
        const double hr = hmax - hmin;
        hmin = hmin - 0.50 * hr;
        hmax = hmax + 0.50 * hr;
    }
    if (hmax <= hmin) {
        hmin -= 1.0e-6;
        hmax += 1.0e-6;
    }
    const double hbin = (hmax - hmin) / (double)bins;
    for (int32_t b = 0; b < bins; ++b) hist[b] = 0.0;
    for (int64_t i = 0; i < nmess; ++i) {
        const double am = mag[i], ax = strx[i], ay = stry[i], an = nbar[i], ae = en[i];
        const double ax2 = ax * ax, ay2 = ay * ay, m2 = am * am, n2 = an * an;
        S[0] += am;
        S[6] += std::fabs(ax); S[8] += ax2; S[9] += ax2 * ax2;
        S[7] += std::fabs(ay); S[10] += ay2; S[11] += ay2 * ay2;
        S[12] += ax2 + ay2; S[13] += (ax2 + ay2) * (ax2 + ay2);
        S[14] += std::sqrt(ax2 + ay2);
        S[1] += std::fabs(am); S[2] += m2; S[3] += m2 * m2;
        S[15] += std::fabs(ay2 - ax2);
        S[16] += an; S[17] += std::fabs(an); S[18] += n2; S[19] += n2 * n2;
        S[4] += ae; S[5] += ae * ae;
        const long idx = std::lround((ae - hmin) / hbin);     // NINT
        if (idx >= 1 && idx <= bins) hist[idx - 1] += 1.0;
    }
    for (int32_t b = 0; b < bins; ++b) hist[b] = hist[b] / (double)nmess;
    derive_basefields(S, nmess, beta, N, besten, finalen, out49);
    *hmin_o = hmin;
    *hbin_o = hbin;
}

}  // namespace j1j2
