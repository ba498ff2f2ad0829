// sweep_common.h -- device helpers shared by the byte- and bit-layout sweep
// kernels (kernels.hip, kernels_bits.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace j1j2 {
namespace {
// The current PT cycle.  Every block of the PT step adds one tick when it is
// done, nconf per cycle; kernels between PT steps see a multiple of nconf,
// and a PT-step block reads the clock before its own tick, so it sees fewer
// than nconf ticks of its own cycle.
__device__ __forceinline__ uint64_t clock_cycle(const DeviceState &st)
{
    return st.clk->ticks / (uint64_t)st.nconf;
}

__device__ __forceinline__ uint32_t lds_bit(uint32_t w, int j) { return (w >> (8 * j + 2)) & 1u; }


// FP64 decision of Metro_sweep (J1J2_6tv.f90:2572-2604) with the engine's
// uniform u = r 2^-32 from the site's Philox word
__device__ __forceinline__ bool decide_f64(double field, bool up, double beta, uint32_t r)
{
    double u = (double)r * 0x1p-32;
    if (up) {
        if (field <= 0.0) return true;
        return u < exp64(-((2.0 * field) * beta));
    }
    if (field >= 0.0) return true;
    return u < exp64((2.0 * field) * beta);
}

// Compute_local_field (J1J2_6tv.f90:2510-2540), occu = jh = jv = 1.
// s[k] = spin of neighbour m_{k+1} (+y,-y,+x,-x,+x+y,-x+y,+x-y,-x-y,+z,-z) as +-1;
// p = (p_i, p(m1), p(m2), p(m3), p(m4)).
__device__ __forceinline__ double field_f64(const double s[10], const double p[5], const ConfConst &cc)
{
    double field = 0.0, n = 0.0;
    field = field + s[0] * cc.J1v; n = n - s[0]; field = field - p[1] * s[0];
    field = field + s[1] * cc.J1v; n = n - s[1]; field = field - p[2] * s[1];
    field = field + s[2] * cc.J1h; n = n + s[2]; field = field + p[3] * s[2];
    field = field + s[3] * cc.J1h; n = n + s[3]; field = field + p[4] * s[3];
    field = field + s[4] * cc.J2;
    field = field + s[5] * cc.J2;
    field = field + s[6] * cc.J2;
    field = field + s[7] * cc.J2;
    field = field + s[8] * cc.J3;
    field = field + s[9] * cc.J3;
    field = field + p[0] * n;
    return field;
}

__device__ __forceinline__ double pm1(uint32_t b) { return b ? 1.0 : -1.0; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of OBS_N int64 values over a group of NT threads (the block, or one of
// the groups a block is split into, NT a multiple of 64 then), tid the thread's
// index in its group; the result is valid in the group's thread 0.  red holds
// a row per wave of the block.
template <int NT>
__device__ void block_reduce_obs(int64_t v[OBS_N], int64_t (*red)[OBS_N], int tid = threadIdx.x)
{
    constexpr int NW = (NT + 63) / 64;      // waves, the last one possibly partial
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < OBS_N; ++k) v[k] = wave_sum<long long>((long long)v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < OBS_N; ++k) red[wave][k] = v[k];
    __syncthreads();
    if (tid == 0)
        for (int w = 1; w < NW; ++w)
#pragma unroll
            for (int k = 0; k < OBS_N; ++k) v[k] += red[wave + w][k];
}

struct SweepCtx {
    const float4 *kc4, *kxp4, *kxm4, *kyp4, *kym4;
    const double *p64;     // config's FP64 field (shadow evaluation)
    float b2, j2d, j3d;        // j2d = 2^24 (2 J2 / 4), j3d = 2^24 (2 J3 / 4) (exact scalings)
    double beta;
    uint32_t ctr1, sweep_lo, sweep_hi, key0, key1;
    int cnt_off;               // resident sweep with flip counts: dwords from a lattice word to its
                               // counter word (one byte per site); 0: no counts
};


// float4 load at a 32-bit byte offset from a block-uniform base, so the load
// can address as SGPR base + VGPR offset (no 64-bit address arithmetic).
__device__ __forceinline__ float4 ld16(const float4 *base, uint32_t byte_off)
{
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + byte_off);
}

// In the resident sweep's LDS copy a site's byte is 4b (spin bit at bit 2),
// which is the high byte of the f16 value 2^-14 b.  With the coupling planes
// stored as 2^14 * 2K, fma(2^-14 b, 2^14 2K, h) has the exact product b 2K, so
// it rounds exactly as fma(b, 2K, h) and the compiler issues it as one
// v_fma_mix_f32 reading the f16 half directly (no byte-to-float conversion).
// f16_odd keeps bytes 1 and 3 in place (sites 1, 3: low and high half);
// f16_even moves bytes 0 and 2 up into the high bytes of the halves.
__device__ __forceinline__ uint32_t f16_odd(uint32_t w) { return w & 0xFF00FF00u; }
__device__ __forceinline__ uint32_t f16_even(uint32_t w) { return __builtin_amdgcn_perm(w, w, 0x020c000cu); }
// The bytewise bond counts D = 4 (0..4) and Z = 4 (0..2) enter the same way,
// as f16 denormals: a byte v as the low byte of a half whose high byte is 0 is
// the f16 value v 2^-24, so against the coefficient 2^24 (2J/4) the product is
// exactly v (2J/4) and the fma rounds as fma(v, 2J/4, h).  lo_even keeps bytes
// 0 and 2 in place, lo_odd moves bytes 1 and 3 down; both zero the high bytes.
__device__ __forceinline__ uint32_t lo_even(uint32_t w) { return w & 0x00FF00FFu; }
__device__ __forceinline__ uint32_t lo_odd(uint32_t w) { return __builtin_amdgcn_perm(w, w, 0x0c030c01u); }
__device__ __forceinline__ float f16_site(uint32_t even, uint32_t odd, int j)
{
    const uint32_t w = (j & 1) ? odd : even;
    return (float)__builtin_bit_cast(_Float16, (uint16_t)((j & 2) ? (w >> 16) : w));
}


// The field term of the energy, sum_i q_i l_i with q_i = llrint(p_i 2^32)
// (|q_i| < 2^38) and l_i = 2 (a(+y) + a(-y) - a(+x) - a(-x)) in -4..4 (a = 1
// for an anti-aligned bond), is summed as 2 sum q_i e_i - 4 sum q_i with
// e_i = l_i / 2 + 2 in 0..4, the last sum a per-config constant.  The
// disorder is held as five byte planes of q_i + 2^38, so for a word of four
// sites (e_i in its bytes) one v_dot4_u32_u8 per plane adds the four products
// e_i * byte_b(q_i + 2^38) exactly into a 32-bit partial sum (at most 1020
// per site), and a sixth adds sum e_i for the offset QOFF (kernels.h).

struct QSum {
    uint32_t a[5], se;
};

// The config's five byte planes (block-uniform pointers).
struct QPlanes {
    const uint8_t *p[5];
};

__device__ __forceinline__ QPlanes qplanes(const uint8_t *qb, int64_t nsites)
{
    QPlanes q;
#pragma unroll
    for (int b = 0; b < 5; ++b) q.p[b] = qb + b * nsites;
    return q;
}

// boff: 4 x the word's dword index within the config's planes (a 32-bit
// offset from the uniform plane bases)
__device__ __forceinline__ void qsum_word(QSum &s, const QPlanes &q, uint32_t boff, uint32_t e)
{
#pragma unroll
    for (int b = 0; b < 5; ++b)
        s.a[b] = __builtin_amdgcn_udot4(e, *reinterpret_cast<const uint32_t *>(q.p[b] + boff), s.a[b], false);
    s.se = __builtin_amdgcn_udot4(e, 0x01010101u, s.se, false);
}

// sum q_i e_i over the words added
__device__ __forceinline__ int64_t qsum_total(const QSum &s)
{
    int64_t t = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) t += (int64_t)s.a[b] << (8 * b);
    return t - QOFF * (int64_t)s.se;
}

// Per-thread partial observables of the fixed-geometry paths: raw popcounts
// and the field term's sum q e (see observe_fast).
struct ObsFast {
    uint32_t nw;                       // words visited
    uint32_t ps, px, py, pz, pd;       // popcounts of S, S^X+, S^Y+, S^Z+, S^D(+x+y) + S^D(-x+y)
    int32_t sxm, sym;                  // signed sums of 2 popc(S) - 4 (x and y stripe)
    QSum q;                            // sum q e
};

}  // namespace
}  // namespace j1j2
