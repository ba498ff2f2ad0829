// device_common.h -- device-side building blocks of the J1-J2 engine (gfx950).
//
// The arithmetic here defines the engine's update rule; a CPU restatement of
// the same rule is the test oracle.  Every floating-point expression that feeds
// a flip decision is written as explicit operations (the file is compiled with
// -ffp-contract=off), so CPU and GPU agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace j1j2 {

// ---------------------------------------------------------------------------
// Descriptors shared by host and device
// ---------------------------------------------------------------------------
struct LatDesc {          // one replica lattice
    int32_t L, H, N;      // N = L*L*H
    int32_t conf;         // global config index
    int32_t rep;          // replica index within its config
    int32_t slot_base;    // index of the config's slot 0 in per-slot arrays
    int32_t packed;       // spins in HBM: 0 one byte per site, 1 one bit per site, 2 unit words
    int64_t spin_off;     // byte offset of this lattice's spins
    int64_t site_off;     // offset (in sites) of the config's per-site planes
};

struct ConfDesc {         // one disorder config
    int32_t L, H, N, nrep;
    int32_t lat0;         // first lattice index of the config
    int32_t slot_base;    // per-slot arrays offset (slot-major blocks of nrep)
    int32_t pair_base;    // per-pair arrays offset (nrep-1 entries)
    int32_t iconf;        // Philox key word 1
    int64_t site_off;     // per-site planes offset
    int64_t hist_off;     // histogram offset (nrep*bins)
    int64_t series_off;   // LOCAL series offset (nrep*5*cap)
    int64_t flip_off;     // flip-count offset (nrep*N)
};

struct DeviceClock {      // counted by the PT step of every cycle
    uint64_t ticks;       // PT-step blocks finished since the engine was created (nconf per cycle)
    uint64_t phase_start; // ticks when the current phase began (a multiple of nconf)
};

// Per-lattice integer observables written after every cycle.
enum { OBS_M = 0, OBS_SX, OBS_SY, OBS_BX, OBS_BY, OBS_BD, OBS_BZ, OBS_EQ, OBS_N };

// Per-slot moment sums, in the accumulation order of Pt_measurement_loop
// (J1J2_6tv.f90:1493-1512).
enum { NSUMS = 20 };

// Internal site layout ("colour planes"): the N/4 sites of colour
// c = (x&1) | ((y+z)&1)<<1 form plane c, ordered by the colour index
// q = (z*L/2 + y/2)*L/2 + x/2, so a colour pass touches one contiguous plane.
// Within a plane, unit u = z*L/2 + y/2 is one half-row (L/2 sites with the
// same x parity).  Spins are stored as one byte b = (s+1)/2 per site.
__host__ __device__ inline int64_t internal_index(int32_t L, int32_t H, int32_t x, int32_t y, int32_t z)
{
    const int64_t half = L >> 1, nq = (int64_t)L * L * H / 4;
    const int c = (x & 1) | (((y + z) & 1) << 1);
    return c * nq + ((int64_t)z * half + (y >> 1)) * half + (x >> 1);
}

// Spins in HBM of the bit-layout sweep's lattices (LatDesc::packed == 2): its
// LDS unit words as they are.  Half-row unit u (internal sites u HW .. u HW +
// HW - 1, HW = L/2) is two dwords at dwords 2u, 2u + 1: sites m < N0 at bits
// 1..N0 of dword 0, sites N0..HW-1 at bits 1..HW-N0 of dword 1, where N0 =
// 4 ceil(G/2) for G = HW/4 four-site groups; bit 0 of each dword repeats the
// site before its first one and the bit after its last site the site after it
// (dword 0: sites HW-1 and N0, dword 1: sites N0-1 and 0).  Dword and bit of
// internal site i:
__host__ __device__ inline void unit_word_bit(int32_t L, int64_t i, int64_t &dw, int &bit)
{
    const int HW = L >> 1, N0 = 4 * ((HW / 4 + 1) / 2);
    const int64_t u = i / HW;
    const int m = (int)(i - u * HW);
    dw = 2 * u + (m >= N0 ? 1 : 0);
    bit = (m >= N0 ? m - N0 : m) + 1;
}

// Inverse of internal_index.
__host__ __device__ inline void site_of_internal(int32_t L, int32_t H, int64_t i, int &x, int &y, int &z)
{
    const int64_t half = L >> 1, nq = (int64_t)L * L * H / 4;
    const int c = (int)(i / nq);
    const int64_t q = i - c * nq;
    const int64_t u = q / half;
    z = (int)(u / half);
    const int yh = (int)(u - (int64_t)z * half);
    x = 2 * (int)(q - u * half) + (c & 1);
    y = 2 * yh + (((c >> 1) ^ z) & 1);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Round i uses key + i*(W0, W1).
// The 32x32->64 products map onto v_mad_u64_u32.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // three-input XOR in one v_bitop3 (truth table 0x96 is symmetric in its inputs)
        uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// The same generator when only c0 varies across the lanes (c1, c2, c3 and the
// key are wave-uniform, as in the sweeps): the first three rounds are written
// so that every uniform sub-expression is formed once on the scalar unit and
// each vector xor takes a single scalar operand.
__device__ __forceinline__ U4 philox4x32_10_u(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1)
{
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u, M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    // round 1: p0 vector, p1 uniform
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t a0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;                  // uniform
    uint32_t u3 = c3 ^ k1;
    asm volatile("" : "+s"(u3));                                    // kept as one scalar (no reassociation)
    uint32_t a2 = (uint32_t)(p0 >> 32) ^ u3;                       // vector
    uint32_t a1 = (uint32_t)p1, a3 = (uint32_t)p0;                 // uniform, vector
    k0 += W0; k1 += W1;
    // round 2: p0 uniform, p1 vector
    p0 = (uint64_t)M0 * a0;
    p1 = (uint64_t)M1 * a2;
    uint32_t b0 = (uint32_t)(p1 >> 32) ^ (a1 ^ k0);                // vector
    uint32_t b2 = a3 ^ ((uint32_t)(p0 >> 32) ^ k1);                // vector
    uint32_t b1 = (uint32_t)p1, b3 = (uint32_t)p0;                 // vector, uniform
    k0 += W0; k1 += W1;
    // round 3: c3 uniform
    p0 = (uint64_t)M0 * b0;
    p1 = (uint64_t)M1 * b2;
    c0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), b1, k0, 0x96);
    c2 = (uint32_t)(p0 >> 32) ^ (b3 ^ k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    k0 += W0; k1 += W1;
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        p0 = (uint64_t)M0 * c0;
        p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += W0;
        k1 += W1;
    }
    return U4{c0, c1, c2, c3};
}

// ---------------------------------------------------------------------------
// Acceptance threshold of the FP32 rule: 2^32 * 2^y.  y = n + f, n = rint(y),
// |f| <= 1/2, 2^f by a degree-6 near-minimax polynomial (relative error
// < 1.6e-8) with coefficients pre-scaled by 2^32.  No clamp: y >= 0 gives
// >= 2^32 and very negative y gives 0 (v_ldexp_f32 saturates n); the host
// keeps |y| far below 2^30 (see the temperature check of the FP32 form).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float threshold32(float y)
{
    float n = rintf(y);
    float f = y - n;
    float p = 0x1.41d332p+19f;
    p = __builtin_fmaf(p, f, 0x1.5f456ap+22f);
    p = __builtin_fmaf(p, f, 0x1.3b2dbcp+25f);
    p = __builtin_fmaf(p, f, 0x1.c6aed4p+27f);
    p = __builtin_fmaf(p, f, 0x1.ebfbdap+29f);
    p = __builtin_fmaf(p, f, 0x1.62e430p+31f);
    p = __builtin_fmaf(p, f, 0x1.000000p+32f);
    return ldexpf(p, (int)n);
}

// Exponent of the FP32 rule for the site whose spin is bit `bit` of the word
// `sw` (set for s = +1), local field h, b2 = 2 beta log2(e): y = -s b2 h.
__device__ __forceinline__ float y_f32_word(float h, uint32_t sw, int bit, float b2)
{
    const float g = b2 * h;
    const uint32_t sgn = (sw << (31 - bit)) & 0x80000000u;
    return __uint_as_float(__float_as_uint(g) ^ sgn);
}

// Metropolis decision of the FP32 rule with the site's Philox word r: accept
// iff (float)r <= 2^32 2^y.  A move with y >= 0 (dE <= 0) is always taken:
// its threshold is >= 2^32, the largest value (float)r rounds to.
__device__ __forceinline__ bool decide_f32_word(float h, uint32_t sw, int bit, float b2, uint32_t r)
{
    return (float)r <= threshold32(y_f32_word(h, sw, bit, b2));
}

// The same decision, first through the hardware exp2 (v_exp_f32) as a filter.
// Both 2^32 v_exp(y) and threshold32(y) lie within a relative 2^-20 of
// 2^32 2^y wherever the threshold is a normal float (the polynomial within
// 1.6e-8; the hardware exp2 within a few ulp), so (float)r below
// 2^32 (1 - 2^-16) v_exp(y) is accepted by the exact rule and above
// 2^32 (1 + 2^-16) v_exp(y) rejected by it.  Below the normal range both
// bounds are < 1 <= (float)r (rejected either way) unless r = 0, which is
// left open like every (float)r inside the band: those sites (a fraction
// ~2^-15 of the proposals) take the exact rule, decide_f32_open.
struct Filt {
    bool acc, open;
};

__device__ __forceinline__ Filt decide_f32_filter(float y, uint32_t r)
{
    const float ur = (float)r, e = __builtin_amdgcn_exp2f(y);
    const bool acc = ur < e * 0x1.fffep31f;          // 2^32 (1 - 2^-16)
    const bool rej = ur > e * 0x1.0001p32f;          // 2^32 (1 + 2^-16)
    return Filt{acc, !(acc || rej)};
}

__device__ __forceinline__ bool decide_f32_open(float y, uint32_t r)
{
    return (float)r <= threshold32(y);
}

// ---------------------------------------------------------------------------
// The sites' random words.  A site's 32-bit word is r = v 2^16 + l: v is one
// 16-bit half of word j of the Philox call with counter (m, c | rep << 2,
// sweep) and l the same half of word j of the call with LO_CTR added to the
// second counter word.  The z-layers 2 zp and 2 zp + 1 of a colour plane
// share the calls m = zp GL + g, g the site's 4-site group within its layer
// (GL per layer): layer 2 zp takes the high halves, layer 2 zp + 1 the low
// halves.  The kernels draw l only when v cannot decide (decide16_filter), so
// one call serves eight sites.
// ---------------------------------------------------------------------------
constexpr uint32_t LO_CTR = 0x40000000u;

__host__ __device__ inline uint32_t half16(uint32_t w, int odd) { return odd ? (w & 0xFFFFu) : (w >> 16); }
__host__ __device__ inline uint32_t site_word(uint32_t hi, uint32_t lo, int odd)
{
    return (half16(hi, odd) << 16) | half16(lo, odd);
}

// v as the float 2^16 + v 2^-7 (v fills the low 16 bits of the mantissa).
__device__ __forceinline__ float v16_float(uint32_t w, int odd) { return __uint_as_float(half16(w, odd) | 0x47800000u); }

// The FP32 rule's decision from v alone, for every l.  With F = 2^16 + v 2^-7
// (exact) and e = v_exp_f32(y), d = RN(F - 2^9 e) is one fma, rounded once
// (error <= 2^-9 below 2^16, <= 2^-8 above).  The rule accepts every l when
// (v + 1) 2^16 <= T(y) (every (float)r is then <= (v + 1) 2^16, a float), and
// T(y) and 2^32 e agree to 2^-22 relative, so for e <= 1 it suffices that
// t = F - 2^9 e <= 2^16 - 2^-7 - 2^-13; d < 2^16 - 2^-7 means d <= 2^16 -
// 3 2^-8 on the float grid, so t <= 2^16 - 5 2^-9 (for y > 0 the rule accepts
// every move anyway).  It rejects every l when v 2^16 > T(y), for which t >
// 2^16 + 2^-13 suffices; d > 2^16 means d >= 2^16 + 2^-7, so t >= 2^16 + 2^-8
// (below the normal range: v >= 1, r >= 2^16 > the threshold).  Open
// otherwise (v within about one of 2^16 2^y: a fraction ~2^-16 of the
// proposals), decided by the exact rule on the whole word.
__device__ __forceinline__ Filt decide16_filter(float y, float F)
{
    const float e = __builtin_amdgcn_exp2f(y);
    const float d = __builtin_fmaf(e, -0x1p9f, F);
    const bool acc = d < 0x1.fffffcp15f;             // 2^16 - 2^-7
    const bool rej = d > 0x1p16f;                    // 2^16
    return Filt{acc, !(acc || rej)};
}

// ---------------------------------------------------------------------------
// e^x in FP64: Cody & Waite reduction x = k ln2 + r (ln2 split so k*ln2_hi is
// exact), Taylor series of e^r to degree 13 in Horner form, ldexp.
// ---------------------------------------------------------------------------
__host__ __device__ inline double exp64(double x)
{
    if (x < -700.0) return 0.0;
    if (x > 709.0) return __builtin_huge_val();
    const double LN2_HI = 0x1.62e42feep-1, LN2_LO = 0x1.a39ef35793c76p-33,
                 INV_LN2 = 0x1.71547652b82fep+0;
    double k = __builtin_rint(x * INV_LN2);
    double r = __builtin_fma(-k, LN2_HI, x);
    r = __builtin_fma(-k, LN2_LO, r);
    double p = 0x1.6124613a86d09p-33;
    p = __builtin_fma(p, r, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, r, 0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, r, 0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, r, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, r, 0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, r, 0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, r, 0x1.1111111111111p-7);
    p = __builtin_fma(p, r, 0x1.5555555555555p-5);
    p = __builtin_fma(p, r, 0x1.5555555555555p-3);
    p = __builtin_fma(p, r, 0x1.0000000000000p-1);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

// The same for the FP64 rule (u = r 2^-32 < E): v accepts when (v 2^16 +
// 2^16 - 1) 2^-32 < E and rejects when v 2^16 2^-32 >= E (both exact in FP64).
struct Dec64 {
    bool acc, open;
};

__device__ __forceinline__ Dec64 decide16_f64(double field, bool up, double beta, uint32_t v)
{
    double E;
    if (up) {
        if (field <= 0.0) return Dec64{true, false};
        E = exp64(-((2.0 * field) * beta));
    } else {
        if (field >= 0.0) return Dec64{true, false};
        E = exp64((2.0 * field) * beta);
    }
    const double u0 = (double)(v << 16) * 0x1p-32, u1 = (double)((v << 16) | 0xFFFFu) * 0x1p-32;
    if (u1 < E) return Dec64{true, false};
    if (u0 >= E) return Dec64{false, false};
    return Dec64{false, true};
}

}  // namespace j1j2
