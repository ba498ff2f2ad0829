// resident.h -- device code of the LDS-resident FP32 sweep (byte layout,
// L % 8 == 0, L <= 48) and the helpers it shares with the global-memory
// kernels of kernels.hip.  The kernels are templates; each block size is
// instantiated in its own translation unit (kernels_res<NT>.hip), so the
// instantiations compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "sweep_common.h"

namespace j1j2 {

namespace {

// ---------------------------------------------------------------------------
// Half-row units of the colour-plane layout (L % 8 == 0, WPH = L/8 dwords per
// half-row).  Plane c holds N/4 bytes = NQ dwords; unit u = z*L/2 + y/2 of
// plane c sits at dwords [c*NQ + u*WPH, +WPH).  Its neighbour half-rows:
//   +-x and the diagonals' x part : plane c^1 (same row), plane c^3 (rows y+-1)
//   +-y                           : plane c^2, units uyp / uym
//   +-z                           : plane c^2, units u +- L/2 (periodic in z)
// Consecutive threads take consecutive units, so every neighbour read of a
// wave is one contiguous run of LDS (or global) memory.
// ---------------------------------------------------------------------------
struct UnitGeo {
    int z, p;                 // p = y & 1
    int uyp, uym, uzp, uzm;
};

template <int WPH>
__device__ __forceinline__ UnitGeo unit_geo(int u, int c, int H)
{
    constexpr int HALF = 4 * WPH;
    UnitGeo g;
    g.z = u / HALF;
    const int yh = u - g.z * HALF;
    g.p = ((c >> 1) ^ g.z) & 1;
    if (g.p) {                                  // y odd: y+1 is the next unit, y-1 this one
        g.uyp = (yh == HALF - 1) ? u - (HALF - 1) : u + 1;
        g.uym = u;
    } else {                                    // y even: y+1 is this unit, y-1 the previous one
        g.uyp = u;
        g.uym = (yh == 0) ? u + (HALF - 1) : u - 1;
    }
    g.uzp = (g.z == H - 1) ? u - (H - 1) * HALF : u + HALF;
    g.uzm = (g.z == 0) ? u + (H - 1) * HALF : u - HALF;
    return g;
}

// Spins in HBM: one byte b = (s+1)/2 per site in the internal (colour-plane)
// order; for the lattices of the byte-layout resident sweep (packed = 1) one
// bit per site (bit i & 31 of word i >> 5 for internal index i), which that
// sweep loads and stores once per cycle; for those of the bit-layout sweep
// (packed = 2) its unit words (unit_word_bit).  Dword idx of the byte form
// (sites 4 idx .. 4 idx + 3) from the first two: PK = 1 reads nibble idx of
// the bits.
template <int PK>
__device__ __forceinline__ uint32_t rd_word(const uint32_t *sp32, int idx)
{
    if constexpr (PK) return (((sp32[idx >> 3] >> (4 * (idx & 7))) & 0xFu) * 0x00204081u) & 0x01010101u;
    else return sp32[idx];
}

__device__ __forceinline__ uint32_t spin_at(const uint8_t *sp, int packed, int L, int64_t i)
{
    const uint32_t *sp32 = reinterpret_cast<const uint32_t *>(sp);
    if (packed == 2) {
        int64_t dw;
        int bit;
        unit_word_bit(L, i, dw, bit);
        return (sp32[dw] >> bit) & 1u;
    }
    return packed ? (sp32[i >> 5] >> (i & 31)) & 1u : sp[i];
}

// +x / -x neighbour bytes of dword w of the half-row at dword offset `uoff`,
// taken from the other-parity half-row of the same row (periodic in x).
template <int WPH, int HX, int PK = 0>
__device__ __forceinline__ void x_pair(const uint32_t *base, int uoff, int w, uint32_t &Xp, uint32_t &Xm)
{
    const uint32_t A = rd_word<PK>(base, uoff + w);
    if constexpr (HX == 0) {
        const uint32_t B = rd_word<PK>(base, uoff + (w == 0 ? WPH - 1 : w - 1));
        Xp = A;
        Xm = __builtin_amdgcn_alignbyte(A, B, 3);
    } else {
        const uint32_t B = rd_word<PK>(base, uoff + (w == WPH - 1 ? 0 : w + 1));
        Xm = A;
        Xp = __builtin_amdgcn_alignbyte(B, A, 1);
    }
}

// Accumulate_local_nematic_average (J1J2_6tv.f90:1009-1025) for one word
// (four sites) of plane c in the colour-plane layout: l_i = 2 (a(+y) + a(-y) -
// a(+x) - a(-x)), a = 1 if the bond is anti-aligned.  A site's spin is bit SH
// of its byte: SH = 0 in global memory, 2 in the resident sweep's LDS copy;
// PK = 1: global memory holds one bit per site (rd_word).
template <int WPH, int HX, int SH, int PK = 0>
__device__ __forceinline__ void locnem_word(const uint32_t *sp32, int c, int H, int t, int4 &ls, int4 &lv)
{
    const int nq = 4 * WPH * WPH * H;
    const int u = t / WPH, w = t - u * WPH;
    const UnitGeo g = unit_geo<WPH>(u, c, H);
    const int py = (c ^ 2) * nq;
    const uint32_t S = rd_word<PK>(sp32, c * nq + t);
    const uint32_t YP = rd_word<PK>(sp32, py + g.uyp * WPH + w), YM = rd_word<PK>(sp32, py + g.uym * WPH + w);
    uint32_t Xp, Xm;
    x_pair<WPH, HX, PK>(sp32, (c ^ 1) * nq + u * WPH, w, Xp, Xm);
    const uint32_t ax = (S ^ Xp) + (S ^ Xm), ay = (S ^ YP) + (S ^ YM);
    int *lsp = reinterpret_cast<int *>(&ls), *lnp = reinterpret_cast<int *>(&lv);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        lsp[j] += (int)((S >> (8 * j + SH)) & 1u) * 2 - 1;
        lnp[j] += 2 * ((int)((ay >> (8 * j + SH)) & 0xffu) - (int)((ax >> (8 * j + SH)) & 0xffu));
    }
}

// The same over the words [gb, ge) of a lattice (word g = c nq + t: plane c,
// word t), a thread taking g = gb + t0, gb + t0 + stride, ... in batches of K
// words whose accumulators (in HBM) are all requested before the first is
// added to: one memory round trip per batch instead of one per word.
template <int WPH, int SH, int PK = 0, int K = 4>
__device__ __forceinline__ void locnem_range(const uint32_t *sp32, int H, int gb, int ge, int t0, int stride,
                                             int32_t *locspin, int32_t *locnem)
{
    const int nq = 4 * WPH * WPH * H;
    int4 *ls4 = reinterpret_cast<int4 *>(locspin), *lv4 = reinterpret_cast<int4 *>(locnem);
    for (int g0 = gb + t0; g0 < ge; g0 += K * stride) {
        int4 ls[K], lv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int g = g0 + k * stride;
            if (g < ge) {
                ls[k] = ls4[g];
                lv[k] = lv4[g];
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int g = g0 + k * stride;
            if (g >= ge) break;
            const int c = g / nq, t = g - c * nq;
            if (c & 1) locnem_word<WPH, 1, SH, PK>(sp32, c, H, t, ls[k], lv[k]);
            else locnem_word<WPH, 0, SH, PK>(sp32, c, H, t, ls[k], lv[k]);
            ls4[g] = ls[k];
            lv4[g] = lv[k];
        }
    }
}

// Every word of a lattice.
template <int WPH, int SH, int PK = 0>
__device__ __forceinline__ void locnem_units(const uint32_t *sp32, int H, int t0, int stride, int32_t *locspin,
                                             int32_t *locnem)
{
    locnem_range<WPH, SH, PK>(sp32, H, 0, 16 * WPH * WPH * H, t0, stride, locspin, locnem);
}

// Per-thread partial sums of the integer observables in int32 (a thread covers
// at most a few hundred sites) and the field term's sum q e (QSum).
struct ObsAcc {
    int32_t m, sx, sy, bx, by, bd, bz;
    QSum q;
};

// Integer observables (see k_observables) of one colour plane of the resident
// sweep's LDS copy (spin = bit 2 of its byte), forward bonds, one dword
// (4 sites) per step; the field term as sum q e (QSum).
template <int WPH, int HX>
__device__ __forceinline__ void observe_plane(const uint32_t *sp32, int c, int H, int t0, int stride,
                                              const QPlanes &qp, ObsAcc &a)
{
    const int nq = 4 * WPH * WPH * H;
    for (int t = t0; t < nq; t += stride) {
        const int u = t / WPH, w = t - u * WPH;
        const UnitGeo g = unit_geo<WPH>(u, c, H);
        const int py = (c ^ 2) * nq;
        const uint32_t S = sp32[c * nq + t];
        const uint32_t YP = sp32[py + g.uyp * WPH + w], YM = sp32[py + g.uym * WPH + w];
        const uint32_t ZP = sp32[py + g.uzp * WPH + w];
        uint32_t Xp0, Xm0, Xp1, Xm1;
        x_pair<WPH, HX>(sp32, (c ^ 1) * nq + u * WPH, w, Xp0, Xm0);
        x_pair<WPH, HX>(sp32, (c ^ 3) * nq + g.uyp * WPH, w, Xp1, Xm1);
        const int m4 = 2 * __builtin_popcount(S) - 4;
        a.m += m4;
        a.sx += HX ? -m4 : m4;
        a.sy += g.p ? -m4 : m4;
        a.bx += 4 - 2 * __builtin_popcount(S ^ Xp0);
        a.by += 4 - 2 * __builtin_popcount(S ^ YP);
        a.bz += 4 - 2 * __builtin_popcount(S ^ ZP);
        a.bd += 8 - 2 * (__builtin_popcount(S ^ Xp1) + __builtin_popcount(S ^ Xm1));
        const uint32_t ax = (S ^ Xp0) + (S ^ Xm0), ay = (S ^ YP) + (S ^ YM);   // bytes 0, 4, 8
        const uint32_t e = ((ay + 0x08080808u) - ax) >> 2;                     // bytes l/2 + 2
        qsum_word(a.q, qp, (uint32_t)(c * nq + t) * 4u, e);
    }
}

// ---------------------------------------------------------------------------
// LDS-resident FP32 sweep.  One block per lattice; the lattice's spins live in
// LDS for all PTSWAPEVER sweeps of the cycle, and the per-site FP32 constants
// stream from L2/HBM as float4 (one per plane per 4 sites).  In a colour pass
// each thread owns units of the colour's plane; lane i of a wave reads the
// 4*WPH bytes at unit u0+i of each neighbour plane.
// ---------------------------------------------------------------------------
// Offsets (in dwords, relative to the updated word) of the neighbour words of
// one colour-plane dword: +-y, +-z (plane c^2), the other half of the row
// (plane c^1, words w and w-+1) and of rows y+-1 (plane c^3).
struct NbOff {
    int yp, ym, zp, zm, xa, xb, d1a, d1b, d2a, d2b;
};

template <int WPH, int HX>
__device__ __forceinline__ NbOff nb_offsets(int c, int nq, int u, int w, const UnitGeo &g)
{
    const int wb = HX ? (w == WPH - 1 ? 0 : w + 1) : (w == 0 ? WPH - 1 : w - 1);
    const int dY = ((c ^ 2) - c) * nq, dX = ((c ^ 1) - c) * nq, dD = ((c ^ 3) - c) * nq;
    NbOff o;
    o.yp = dY + (g.uyp - u) * WPH;
    o.ym = dY + (g.uym - u) * WPH;
    o.zp = dY + (g.uzp - u) * WPH;
    o.zm = dY + (g.uzm - u) * WPH;
    o.xa = dX;
    o.xb = dX + (wb - w);
    o.d1a = dD + (g.uyp - u) * WPH;
    o.d1b = o.d1a + (wb - w);
    o.d2a = dD + (g.uym - u) * WPH;
    o.d2b = o.d2a + (wb - w);
    return o;
}

template <int HX>
__device__ __forceinline__ void x_from(uint32_t A, uint32_t B, uint32_t &Xp, uint32_t &Xm)
{
    if constexpr (HX == 0) {      // even sites: +x = odd m, -x = odd m-1
        Xp = A;
        Xm = __builtin_amdgcn_alignbyte(A, B, 3);
    } else {                      // odd sites: -x = even m, +x = even m+1
        Xm = A;
        Xp = __builtin_amdgcn_alignbyte(B, A, 1);
    }
}

// The same observables when a plane's words map onto the threads with a fixed
// (y/2, w) per thread and z stepping by an even count (fast_zs below: all
// sizes with 1024 threads, L <= 40 with 256 and 512, L != 40 with 576):
// neighbour offsets are then constant per thread, the
// y parity too, so the bond counts accumulate raw popcounts (turned into
// sums of s_i s_j once at the end) and the field term accumulates
// q_i e_i with e_i = l_i / 2 + 2 in 0..4 (sum q_i l_i = 2 sum q_i e_i -
// 4 sum q_i, the last sum being a per-config constant).

// The fixed-geometry stepping: a step covers ZS = NT / PER_Z z-layers of a
// colour plane and threads ZS * PER_Z.. of the block idle through the pass.
// An even ZS keeps each thread's y parity (one set of neighbour offsets); an
// odd ZS alternates it from step to step (two sets, used in turn).  Used when
// ZS >= 2 and at least 3/4 of the threads work (0 otherwise: the general path).
template <int WPH, int NT>
constexpr int fast_zs()
{
    constexpr int PER_Z = 4 * WPH * WPH, ZS = NT / PER_Z;
    return (ZS >= 2 && 4 * ZS * PER_Z >= 3 * NT) ? ZS : 0;
}

// With the lattice height HC known at compile time, a colour pass of the
// fixed-geometry stepping unrolls all of its ceil(HC / ZS) steps (at most 12).
// The sweep's colour passes work in z-layer pairs (one Philox call per pair
// of words): a step covers ZP = NT / PER_Z layer pairs, threads ZP * PER_Z..
// idle through the pass.  Used when ZP >= 1 and at least 3/4 of the threads
// work (0 otherwise: one pair of words per thread and trip, any geometry).
template <int WPH, int NT>
constexpr int pair_zs()
{
    constexpr int PER_Z = 4 * WPH * WPH, ZP = NT / PER_Z;
    return (ZP >= 1 && 4 * ZP * PER_Z >= 3 * NT) ? ZP : 0;
}

// With the lattice height HC known at compile time, a colour pass unrolls all
// of its ceil(HC/2 / ZP) steps (at most 6, i.e. 12 words).
template <int WPH, int NT, int HC>
constexpr bool unrolled_pairs()
{
    constexpr int ZP = pair_zs<WPH, NT>();
    return HC > 0 && ZP >= 1 && (HC / 2 + ZP - 1) / ZP <= 6;
}

// Steps of an unrolled colour pass (two words each).
template <int WPH, int NT, int HC>
constexpr int pair_steps()
{
    constexpr int ZP = pair_zs<WPH, NT>();
    return unrolled_pairs<WPH, NT, HC>() ? (HC / 2 + ZP - 1) / ZP : 0;
}

template <int WPH, int NT, int HX>
__device__ __forceinline__ void observe_fast(const uint32_t *lds32, int c, int H, const QPlanes &qp,
                                             ObsFast &a, int tid)
{
    constexpr int HALF = 4 * WPH, PER_Z = HALF * WPH, ZSTEP = fast_zs<WPH, NT>();
    const int nq = PER_Z * H;
    if (tid >= ZSTEP * PER_Z) return;
    const int tz = tid % PER_Z, w = tz % WPH;
    int z0 = tid / PER_Z;
    if constexpr (PER_Z % 64 == 0) z0 = __builtin_amdgcn_readfirstlane(z0);
    const int u0 = z0 * HALF + tz / WPH;
    const UnitGeo g = unit_geo<WPH>(u0, c, H);
    const NbOff oa = nb_offsets<WPH, HX>(c, nq, u0, w, g);
    // odd ZSTEP: every other step is in the other y parity (offsets ob)
    const int u1 = u0 + (ZSTEP & 1) * ZSTEP * HALF;
    const UnitGeo g1 = unit_geo<WPH>(u1, c, H);
    const NbOff ob = nb_offsets<WPH, HX>(c, nq, u1, w, g1);
    const int dY = ((c ^ 2) - c) * nq;
    const int own0 = c * nq + z0 * PER_Z + tz;
    uint32_t ps = 0, px = 0, py = 0, pz = 0, pd = 0, nw = 0, ps1 = 0, nw1 = 0;
#pragma unroll 2
    for (int z = z0; z < H; z += ZSTEP) {
        const bool alt = (ZSTEP & 1) && (((z - z0) / ZSTEP) & 1);
        const NbOff &o = alt ? ob : oa;
        const int own = own0 + (z - z0) * PER_Z;
        const uint32_t S = lds32[own];
        const uint32_t YP = lds32[own + o.yp], YM = lds32[own + o.ym];
        const uint32_t ZP = lds32[own + dY + (z == H - 1 ? -(H - 1) * PER_Z : PER_Z)];
        uint32_t Xp0, Xm0, Xp1, Xm1;
        x_from<HX>(lds32[own + o.xa], lds32[own + o.xb], Xp0, Xm0);
        x_from<HX>(lds32[own + o.d1a], lds32[own + o.d1b], Xp1, Xm1);
        const uint32_t sx0 = S ^ Xp0, sy0 = S ^ YP;
        if (alt) {
            ps1 += __builtin_popcount(S);
            nw1 += 1;
        } else {
            ps += __builtin_popcount(S);
            nw += 1;
        }
        px += __builtin_popcount(sx0);
        py += __builtin_popcount(sy0);
        pz += __builtin_popcount(S ^ ZP);
        pd += __builtin_popcount(S ^ Xp1) + __builtin_popcount(S ^ Xm1);
        const uint32_t ax = sx0 + (S ^ Xm0), ay = sy0 + (S ^ YM);          // bytes 0, 4, 8
        const uint32_t e = ((ay + 0x08080808u) - ax) >> 2;                 // bytes l/2 + 2
        qsum_word(a.q, qp, (uint32_t)own * 4u, e);
    }
    const int32_t m = 2 * (int32_t)ps - 4 * (int32_t)nw, m1 = 2 * (int32_t)ps1 - 4 * (int32_t)nw1;
    a.ps += ps + ps1; a.px += px; a.py += py; a.pz += pz; a.pd += pd; a.nw += nw + nw1;
    a.sxm += HX ? -(m + m1) : m + m1;
    a.sym += (g.p ? -m : m) + (g1.p ? -m1 : m1);
}

// Hide how an LDS address was formed from the optimiser: it stays one
// register (an empty asm statement; no instruction is emitted).
template <typename T>
__device__ __forceinline__ void opaque(T *&p)
{
    auto a = (__attribute__((address_space(3))) T *)p;
    asm volatile("" : "+v"(a));
    p = (T *)a;
}

// LDS addresses of a word and of its neighbour words.
struct NbPtr {
    uint32_t *own;
    const uint32_t *yp, *ym, *zp, *zm, *xa, *xb, *d1a, *d1b, *d2a, *d2b;
};

__device__ __forceinline__ NbPtr nb_ptrs(uint32_t *lds32, int own, const NbOff &o)
{
    NbPtr P;
    P.own = lds32 + own;
    P.yp = P.own + o.yp; P.ym = P.own + o.ym; P.zp = P.own + o.zp; P.zm = P.own + o.zm;
    P.xa = P.own + o.xa; P.xb = P.own + o.xb; P.d1a = P.own + o.d1a; P.d1b = P.own + o.d1b;
    P.d2a = P.own + o.d2a; P.d2b = P.own + o.d2b;
    return P;
}

__device__ __forceinline__ NbPtr nb_shift(const NbPtr &P, int d)
{
    NbPtr Q;
    Q.own = P.own + d;
    Q.yp = P.yp + d; Q.ym = P.ym + d; Q.zp = P.zp + d; Q.zm = P.zm + d; Q.xa = P.xa + d; Q.xb = P.xb + d;
    Q.d1a = P.d1a + d; Q.d1b = P.d1b + d; Q.d2a = P.d2a + d; Q.d2b = P.d2b + d;
    return Q;
}

// Update the 4 sites of colour-plane dword `own` in a layer of parity ODD,
// whose random words are the ODD halves of the words `hi` of Philox call m
// (the low 16 bits of a site's word from the call with LO_CTR, drawn only when
// the high ones cannot decide); P holds the LDS addresses of the word and its
// neighbours.
// The f16 views of the x neighbours straight from the two words of the
// other-parity half-row (A: the same word, B: the one before (HX = 0) or
// after (HX = 1)): the shifted neighbour word is [B3 A0 A1 A2] (HX = 0) or
// [A1 A2 A3 B0] (HX = 1), so its even (HX = 0) or odd (HX = 1) view is one
// byte permute of (A, B), and its other view is the unshifted word's other
// view (f16_odd([B3 A0 A1 A2]) = f16_even(A), f16_even([A1 A2 A3 B0]) =
// f16_odd(A)).
template <int HX>
__device__ __forceinline__ void x_views(uint32_t A, uint32_t B, uint32_t &xpe, uint32_t &xpo, uint32_t &xme,
                                        uint32_t &xmo)
{
    if constexpr (HX == 0) {
        xpe = f16_even(A);
        xpo = f16_odd(A);
        xme = __builtin_amdgcn_perm(A, B, 0x050c030cu);
        xmo = xpe;
    } else {
        xme = f16_even(A);
        xmo = f16_odd(A);
        xpo = __builtin_amdgcn_perm(B, A, 0x040c020cu);
        xpe = xmo;
    }
}

template <int WPH, int HX, bool SHADOW, int ODD, bool RC = false, bool CB = false>
__device__ __forceinline__ void update_word(uint32_t *lds32, const SweepCtx &cx, const ConfConst &cc, int own,
                                           uint32_t m, int w, const NbPtr &P, const U4 &hi,
                                           unsigned long long &sh_prop, unsigned long long &sh_mis,
                                           uint32_t *fa = nullptr)
{
    constexpr int HALF = 4 * WPH;
    // CB: a site's byte is 4 b + 8 c (c its flips in this launch): the
    // neighbours' spin bits alone enter the field
    constexpr uint32_t NM = CB ? 0x04040404u : 0xFFFFFFFFu;
    const uint32_t S = *P.own;
    const uint32_t YP = *P.yp & NM, YM = *P.ym & NM;
    const uint32_t ZP = *P.zp & NM, ZM = *P.zm & NM;
    const uint32_t XA = *P.xa & NM, XB = *P.xb & NM, D1A = *P.d1a & NM, D1B = *P.d1b & NM, D2A = *P.d2a & NM,
                   D2B = *P.d2b & NM;
    // diagonals: the two rows' words summed first (bytes 0, 4, 8: no carries),
    // so the x shift of the sum is one align
    uint32_t Dp, Dm;
    x_from<HX>(D1A + D2A, D1B + D2B, Dp, Dm);
    const uint32_t D = Dp + Dm;                         // bytewise 4 * (0..4)
    const uint32_t Z = ZP + ZM;                         // bytewise 4 * (0..2)
    const uint32_t de = lo_even(D), dod = lo_odd(D), ze = lo_even(Z), zod = lo_odd(Z);
    uint32_t xpe, xpo, xme, xmo;
    x_views<HX>(XA, XB, xpe, xpo, xme, xmo);
    const uint32_t ype = f16_even(YP), ypo = f16_odd(YP), yme = f16_even(YM), ymo = f16_odd(YM);
    // 32-bit byte offset from a uniform base, formed here: hidden from the
    // optimiser, which would otherwise hoist every word's five 64-bit plane
    // addresses out of the sweep loop (registers the colour passes need)
    uint32_t ob = (uint32_t)own * 16u;
    asm volatile("" : "+v"(ob));
    const float4 kc = ld16(cx.kc4, ob), kxp = ld16(cx.kxp4, ob), kxm = ld16(cx.kxm4, ob), kyp = ld16(cx.kyp4, ob),
                 kym = ld16(cx.kym4, ob);
    const float kca[4] = {kc.x, kc.y, kc.z, kc.w};
    const float kxpa[4] = {kxp.x, kxp.y, kxp.z, kxp.w};
    const float kxma[4] = {kxm.x, kxm.y, kxm.z, kxm.w};
    const float kypa[4] = {kyp.x, kyp.y, kyp.z, kyp.w};
    const float kyma[4] = {kym.x, kym.y, kym.z, kym.w};
    const uint32_t ha[4] = {hi.x, hi.y, hi.z, hi.w};
    uint32_t ra[4] = {0u, 0u, 0u, 0u};
    if constexpr (SHADOW) {                             // both rules on the whole words
        const U4 lo = philox4x32_10_u(m, cx.ctr1 | LO_CTR, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
        ra[0] = site_word(hi.x, lo.x, ODD); ra[1] = site_word(hi.y, lo.y, ODD);
        ra[2] = site_word(hi.z, lo.z, ODD); ra[3] = site_word(hi.w, lo.w, ODD);
    }
    uint32_t flip = 0;
    float ys[4];
    uint32_t open = 0;
    // the four sites' fields term by term (each site's sum in the same order):
    // independent neighbours in the issue stream, no dependent fma pairs
    float h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = __builtin_fmaf(f16_site(xpe, xpo, j), kxpa[j], kca[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = __builtin_fmaf(f16_site(xme, xmo, j), kxma[j], h[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = __builtin_fmaf(f16_site(ype, ypo, j), kypa[j], h[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = __builtin_fmaf(f16_site(yme, ymo, j), kyma[j], h[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = __builtin_fmaf(f16_site(de, dod, j), cx.j2d, h[j]);  // (2^-24 4D) (2^24 2 J2 / 4)
#pragma unroll
    for (int j = 0; j < 4; ++j) h[j] = __builtin_fmaf(f16_site(ze, zod, j), cx.j3d, h[j]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ys[j] = y_f32_word(h[j], S, 8 * j + 2, cx.b2);
        if constexpr (!SHADOW) {
            const Filt f = decide16_filter(ys[j], v16_float(ha[j], ODD));
            flip |= (f.acc ? 4u : 0u) << (8 * j);
            open |= f.open ? 1u << j : 0u;
            continue;
        }
        const bool acc = decide_f32_open(ys[j], ra[j]);
        if constexpr (SHADOW) {
            const bool up = lds_bit(S, j) != 0;
            const int mm = 4 * w + j;                         // position in the half-row
            const int mo_p = HX ? ((mm == HALF - 1) ? 0 : mm + 1) : mm;
            const int mo_m = HX ? mm : ((mm == 0) ? HALF - 1 : mm - 1);
            uint32_t Xp0, Xm0, Xp1, Xm1, Xp2, Xm2;
            x_from<HX>(XA, XB, Xp0, Xm0);
            x_from<HX>(D1A, D1B, Xp1, Xm1);
            x_from<HX>(D2A, D2B, Xp2, Xm2);
            const double sv[10] = {pm1(lds_bit(YP, j)), pm1(lds_bit(YM, j)), pm1(lds_bit(Xp0, j)),
                                   pm1(lds_bit(Xm0, j)), pm1(lds_bit(Xp1, j)), pm1(lds_bit(Xm1, j)),
                                   pm1(lds_bit(Xp2, j)), pm1(lds_bit(Xm2, j)), pm1(lds_bit(ZP, j)),
                                   pm1(lds_bit(ZM, j))};
            const int xbase = 4 * ((int)(P.xa - lds32) - w);  // first site of the other half-row
            const double pv[5] = {cx.p64[4 * own + j], cx.p64[4 * (int)(P.yp - lds32) + j],
                                  cx.p64[4 * (int)(P.ym - lds32) + j], cx.p64[xbase + mo_p], cx.p64[xbase + mo_m]};
            const bool a64 = decide_f64(field_f64(sv, pv, cc), up, cx.beta, ra[j]);
            sh_prop += 1;
            sh_mis += (a64 != acc);
        }
        flip |= (acc ? 4u : 0u) << (8 * j);
    }
    if (!SHADOW && open) {                                // rare: the low halves and the exact rule
        const U4 lo = philox4x32_10_u(m, cx.ctr1 | LO_CTR, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
        const uint32_t la[4] = {lo.x, lo.y, lo.z, lo.w};
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (((open >> j) & 1u) && decide_f32_open(ys[j], site_word(ha[j], la[j], ODD))) flip |= 4u << (8 * j);
    }
    if constexpr (CB) {
        *P.own = (S ^ flip) + (flip << 1);              // the spin bit toggled, c + 1 (no carry out: c < 32)
        return;
    }
    *P.own = S ^ flip;
    // flip counts (WRITENUMFLIPS): this thread alone touches these four sites'
    // counters in this pass, in a register (RC) or in the LDS after the lattice
    if constexpr (RC) *fa += flip >> 2;
    else if (cx.cnt_off) P.own[cx.cnt_off] += flip >> 2;
}

// The odd layer of a pair from the even layer's addresses P (layer 2 zp, y
// parity PY) and the odd layer's rows that the even layer does not read (yf:
// plane c^2, dfa and dfb: plane c^3, at row yh + 1 for PY = 0 and yh - 1 for
// PY = 1, all in the odd layer).  Every other neighbour word of the odd layer
// is the even layer's word of the same role or of the other y direction, one
// layer (PER_Z dwords) further.
template <int PY, int PER_Z>
__device__ __forceinline__ NbPtr odd_layer(const NbPtr &P, const uint32_t *yf, const uint32_t *dfa,
                                           const uint32_t *dfb)
{
    NbPtr Q;
    Q.own = P.own + PER_Z;
    Q.zp = P.zp + PER_Z; Q.zm = P.zm + PER_Z; Q.xa = P.xa + PER_Z; Q.xb = P.xb + PER_Z;
    if constexpr (PY == 0) {      // even layer: y+1 in row yh, y-1 in row yh-1; odd layer: y+1 in row yh+1, y-1 in row yh
        Q.yp = yf; Q.ym = P.yp + PER_Z;
        Q.d1a = dfa; Q.d1b = dfb; Q.d2a = P.d1a + PER_Z; Q.d2b = P.d1b + PER_Z;
    } else {                      // even layer: y+1 in row yh+1, y-1 in row yh; odd layer: y+1 in row yh, y-1 in row yh-1
        Q.yp = P.ym + PER_Z; Q.ym = yf;
        Q.d1a = P.d2a + PER_Z; Q.d1b = P.d2b + PER_Z; Q.d2a = dfa; Q.d2b = dfb;
    }
    return Q;
}

// One colour pass over plane C.  HC > 0: the lattice height is HC, known at
// compile time.  Work comes in layer pairs (2 zp, 2 zp + 1), one Philox call
// per pair of words (m = zp PER_Z + the word's index within its layer).
template <int WPH, int NT, int C, bool SHADOW, int HC = 0, bool RC = false, bool CB = false>
__device__ __forceinline__ void colour_pass(uint32_t *lds32, const SweepCtx &cx, const ConfConst &cc, int H,
                                            unsigned long long &sh_prop, unsigned long long &sh_mis, int tid,
                                            uint32_t *fa = nullptr)
{
    constexpr int HX = C & 1, PY = C >> 1;
    constexpr int HALF = 4 * WPH, PER_Z = HALF * WPH;    // dwords per z-layer of a plane
    const int nq = PER_Z * H, npairs = H >> 1;
    if constexpr (pair_zs<WPH, NT>() >= 1) {
        // Each thread keeps its (y/2, w) and steps over layer pairs by ZP
        // pairs: every neighbour offset is constant except at the z boundary
        // (layer 0, the even layer of the first pair, and layer H - 1, the odd
        // layer of the last), so a later step's LDS addresses are an earlier
        // one's plus a constant, which the reads take as an immediate.
        constexpr int ZP = pair_zs<WPH, NT>(), STEP = 2 * ZP * PER_Z, TRIP = 2;
        if (tid >= ZP * PER_Z) return;
        const int tz = tid % PER_Z, w = tz % WPH, yh = tz / WPH;
        int zp0 = tid / PER_Z;
        if constexpr (PER_Z % 64 == 0) zp0 = __builtin_amdgcn_readfirstlane(zp0);   // wave-uniform
        const int u0 = 2 * zp0 * HALF + yh;
        NbOff o = nb_offsets<WPH, HX>(C, nq, u0, w, unit_geo<WPH>(u0, C, H));
        const int dY = ((C ^ 2) - C) * nq, dD = ((C ^ 3) - C) * nq;
        o.zp = dY + PER_Z;                                // away from the z boundary
        o.zm = dY - PER_Z;
        const int own0 = C * nq + 2 * zp0 * PER_Z + tz;
        NbPtr P = nb_ptrs(lds32, own0, o);
        // the odd layer's far row (unit offset from the thread's unit, periodic in y)
        const int far = PY ? (yh == 0 ? HALF - 1 : -1) : (yh == HALF - 1 ? -(HALF - 1) : 1);
        const uint32_t *yf = P.own + PER_Z + dY + far * WPH;
        const uint32_t *dfa = P.own + PER_Z + dD + far * WPH;
        const uint32_t *dfb = dfa + (o.xb - o.xa);
        // materialise each address in one VGPR (otherwise the plane base is
        // re-added to every address on every trip)
        opaque(P.own); opaque(P.yp); opaque(P.ym); opaque(P.zp); opaque(P.zm); opaque(P.xa); opaque(P.xb);
        opaque(P.d1a); opaque(P.d1b); opaque(P.d2a); opaque(P.d2b); opaque(yf); opaque(dfa); opaque(dfb);
        const uint32_t m0 = (uint32_t)(zp0 * PER_Z + tz);
        const auto step = [&](int kk, int zp) {
            const uint32_t m = m0 + (uint32_t)(kk * ZP * PER_Z);
            const U4 hi = philox4x32_10_u(m, cx.ctr1, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
            NbPtr Q = nb_shift(P, kk * STEP);
            if (zp == 0) Q.zm += H * PER_Z;                // periodic in z
            update_word<WPH, HX, SHADOW, 0, RC, CB>(lds32, cx, cc, own0 + kk * STEP, m, w, Q, hi, sh_prop, sh_mis,
                                                RC ? fa + 2 * kk : nullptr);
            NbPtr Q1 = nb_shift(odd_layer<PY, PER_Z>(P, yf, dfa, dfb), kk * STEP);
            if (zp == npairs - 1) Q1.zp -= H * PER_Z;
            update_word<WPH, HX, SHADOW, 1, RC, CB>(lds32, cx, cc, own0 + kk * STEP + PER_Z, m, w, Q1, hi, sh_prop,
                                                sh_mis, RC ? fa + 2 * kk + 1 : nullptr);
        };
        if constexpr (unrolled_pairs<WPH, NT, HC>()) {
            // a thread (first pair zp0 < ZP) takes NK = ceil(HC/2 / ZP) steps,
            // the last one only if zp0 + (NK-1) ZP < HC/2: all of them unrolled
            constexpr int NK = (HC / 2 + ZP - 1) / ZP;
            constexpr bool RAGGED = (HC / 2) % ZP != 0;
#pragma unroll
            for (int kk = 0; kk < NK; ++kk) {
                const int zp = zp0 + kk * ZP;
                if (RAGGED && kk == NK - 1 && zp >= HC / 2) break;
                step(kk, zp);
            }
        } else {
            static_assert(!RC, "register flip counts need the unrolled pass");
            const int nk = (npairs - zp0 + ZP - 1) / ZP;
            for (int k = 0; k < nk; k += TRIP) {
#pragma unroll
                for (int i = 0; i < TRIP; ++i)
                    if (k + i < nk) step(k + i, zp0 + (k + i) * ZP);
            }
        }
    } else {
        static_assert(!RC, "register flip counts need the unrolled pass");
        for (int mm = tid; mm < nq / 2; mm += NT) {
            const int zp = mm / PER_Z, tz = mm - zp * PER_Z;
            const U4 hi = philox4x32_10_u((uint32_t)mm, cx.ctr1, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
            const int t0 = 2 * zp * PER_Z + tz, t1 = t0 + PER_Z;
            const int u0 = t0 / WPH, w = t0 - u0 * WPH, u1 = u0 + HALF;
            const NbOff o0 = nb_offsets<WPH, HX>(C, nq, u0, w, unit_geo<WPH>(u0, C, H));
            update_word<WPH, HX, SHADOW, 0, false, CB>(lds32, cx, cc, C * nq + t0, (uint32_t)mm, w,
                                            nb_ptrs(lds32, C * nq + t0, o0), hi, sh_prop, sh_mis);
            const NbOff o1 = nb_offsets<WPH, HX>(C, nq, u1, w, unit_geo<WPH>(u1, C, H));
            update_word<WPH, HX, SHADOW, 1, false, CB>(lds32, cx, cc, C * nq + t1, (uint32_t)mm, w,
                                            nb_ptrs(lds32, C * nq + t1, o1), hi, sh_prop, sh_mis);
        }
    }
}

// The counts a thread kept in registers through the unrolled colour passes
// (word kk, layer o of plane C: the pass's own0 + kk STEP + o PER_Z) go to
// the slot's pending byte per site (four sites a word: bytes < 128 plus at
// most PTSWAPEVER <= 127, no carries); a word whose bytes reach 128 is added
// to confnumflips(slot, site) and cleared.  So a launch moves 8 bytes per four
// sites, not the 32 of the 4-byte counters, and the counters once in about
// 128 / PTSWAPEVER launches.  All of a thread's pending words are requested
// before the first is added to (one memory round trip).
template <int WPH, int NT, int HC, int RCW>
__device__ __forceinline__ void flush_reg_counts(int32_t *flips, uint8_t *pend8, const uint32_t (&fa)[4][RCW],
                                                 int tid)
{
    constexpr int PER_Z = 4 * WPH * WPH, ZP = pair_zs<WPH, NT>(), STEP = 2 * ZP * PER_Z;
    constexpr int nq = PER_Z * HC;
    if (tid >= ZP * PER_Z) return;
    const int tz = tid % PER_Z, zp0 = tid / PER_Z;
    int4 *fl = reinterpret_cast<int4 *>(flips);
    uint32_t *pend = reinterpret_cast<uint32_t *>(pend8);
    uint32_t p[4][RCW];
#pragma unroll
    for (int C = 0; C < 4; ++C)
#pragma unroll
        for (int q = 0; q < RCW; ++q)
            if (zp0 + (q >> 1) * ZP < HC / 2 && fa[C][q])
                p[C][q] = pend[C * nq + 2 * zp0 * PER_Z + tz + (q >> 1) * STEP + (q & 1) * PER_Z];
#pragma unroll
    for (int C = 0; C < 4; ++C) {
        const int own0 = C * nq + 2 * zp0 * PER_Z + tz;
#pragma unroll
        for (int q = 0; q < RCW; ++q) {
            if (zp0 + (q >> 1) * ZP >= HC / 2 || !fa[C][q]) continue;
            const int i = own0 + (q >> 1) * STEP + (q & 1) * PER_Z;
            const uint32_t n4 = p[C][q] + fa[C][q];
            if (n4 & 0x80808080u) {
                int4 g = fl[i];
                g.x += (int)(n4 & 0xFFu);
                g.y += (int)((n4 >> 8) & 0xFFu);
                g.z += (int)((n4 >> 16) & 0xFFu);
                g.w += (int)(n4 >> 24);
                fl[i] = g;
                pend[i] = 0u;
            } else
                pend[i] = n4;
        }
    }
}

// The sweeps of one lattice by a group of NT threads (tid: index in the
// group; the whole block, or one of several groups that sweep a lattice each
// in their own part of the LDS).  active = false: the group repeats the work
// of another group's lattice in its own LDS (a bundle with fewer lattices
// than groups) and writes nothing to global memory; it takes every barrier.
template <int WPH, int NT, bool SHADOW, int HC = 0, bool RC = false, bool CB = false>
__device__ __forceinline__ void resident_body(const DeviceState &st, const LatDesc &d, int lat,
                                              uint32_t *lds32, int64_t (*red)[OBS_N], int tid = threadIdx.x,
                                              bool active = true)
{
    const int H = HC > 0 ? HC : d.H;
    const int N = 64 * WPH * WPH * H;
    uint8_t *gsp = st.spins + d.spin_off;

    // bits -> bytes 4b in LDS (32 sites per word); a thread's words are all
    // requested before the first is unpacked (one memory round trip, not
    // four), the first batch before anything else
    constexpr int LDU = 4;
    const uint32_t *gsp32 = reinterpret_cast<const uint32_t *>(gsp);
    uint32_t W0[LDU];
#pragma unroll
    for (int q = 0; q < LDU; ++q) {
        const int i = tid + q * NT;
        W0[q] = i < N / 32 ? gsp32[i] : 0u;
    }
    // The last measuring cycle left a replica of this config at slot
    // PTMEASINDEX, and its PT step kept that replica's packed spins: every
    // block of the config adds its share of the words to the local nematic
    // sums (J1J2_6tv.f90:1009-1025), reading the copy from LDS (PK = 1, before
    // the lattice takes the LDS), so no block carries the whole lattice's
    // accumulator traffic.  The first batch of accumulators is requested with
    // the copy and the lattice words above (one round trip for all three).
    if (st.fuse_locnem && st.locsnap && st.locnem_pending[d.conf]) {
        const int nw = N / 4, R = st.confs[d.conf].nrep;
        const int gb = (int)((int64_t)nw * d.rep / R), ge = (int)((int64_t)nw * (d.rep + 1) / R);
        int4 *ls4 = reinterpret_cast<int4 *>(st.locspin + d.site_off);
        int4 *lv4 = reinterpret_cast<int4 *>(st.locnem + d.site_off);
        constexpr int K = 4;
        int4 ls[K], lv[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int g = gb + tid + k * NT;
            if (g < ge) {
                ls[k] = ls4[g];
                lv[k] = lv4[g];
            }
        }
        const uint32_t *snap = st.locsnap + d.site_off / 32;
#pragma unroll 4
        for (int i = tid; i < N / 32; i += NT) lds32[i] = snap[i];
        __syncthreads();
        const int nq = N / 16;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int g = gb + tid + k * NT;
            if (g >= ge || !active) break;
            const int c = g / nq, t = g - c * nq;
            if (c & 1) locnem_word<WPH, 1, 0, 1>(lds32, c, H, t, ls[k], lv[k]);
            else locnem_word<WPH, 0, 0, 1>(lds32, c, H, t, ls[k], lv[k]);
            ls4[g] = ls[k];
            lv4[g] = lv[k];
        }
        // the rest of the share (blocks of configs with few replicas)
        if (active) {
            locnem_range<WPH, 0, 1>(lds32, H, gb + K * NT, ge, tid, NT, st.locspin + d.site_off,
                                    st.locnem + d.site_off);
            if (tid == 0 && d.rep == 0) st.locsamples[d.conf] += 1;
        }
        __syncthreads();                                // the copy's LDS becomes the lattice's
    }
    for (int i0 = tid; i0 < N / 32; i0 += LDU * NT) {
        uint32_t W[LDU];
        if (i0 == tid) {
#pragma unroll
            for (int q = 0; q < LDU; ++q) W[q] = W0[q];
        } else {
#pragma unroll
            for (int q = 0; q < LDU; ++q) {
                const int i = i0 + q * NT;
                W[q] = i < N / 32 ? gsp32[i] : 0u;
            }
        }
#pragma unroll
        for (int q = 0; q < LDU; ++q) {
            const int i = i0 + q * NT;
            if (i >= N / 32) break;
            uint32_t b[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) b[k] = (((W[q] >> (4 * k)) & 0xFu) * 0x00810204u) & 0x04040404u;
            reinterpret_cast<uint4 *>(lds32)[2 * i] = make_uint4(b[0], b[1], b[2], b[3]);
            reinterpret_cast<uint4 *>(lds32)[2 * i + 1] = make_uint4(b[4], b[5], b[6], b[7]);
        }
    }
    // flip counts of this launch, one byte per site right after the lattice
    // (at most PTSWAPEVER <= 255 per site), added to the slot's counters at the end
    const bool counting = st.flips != nullptr;
    if (counting && !RC && !CB)
        for (int i = tid; i < N / 4; i += NT) lds32[N / 4 + i] = 0u;
    // RC: the counts of the words a thread updates stay in its registers
    uint32_t fa[4][RC ? 2 * pair_steps<WPH, NT, HC>() : 1] = {};

    const ConfDesc cd = st.confs[d.conf];
    const ConfConst cc = st.cconst[d.conf];
    const int slot = st.temp_of_rep[d.slot_base + d.rep];
    SweepCtx cx;
    cx.kc4 = reinterpret_cast<const float4 *>(st.kd + 0 * st.nsites + d.site_off);
    cx.kxp4 = reinterpret_cast<const float4 *>(st.kd + 1 * st.nsites + d.site_off);
    cx.kxm4 = reinterpret_cast<const float4 *>(st.kd + 2 * st.nsites + d.site_off);
    cx.kyp4 = reinterpret_cast<const float4 *>(st.kd + 3 * st.nsites + d.site_off);
    cx.kym4 = reinterpret_cast<const float4 *>(st.kd + 4 * st.nsites + d.site_off);
    cx.p64 = st.p64 + d.site_off;
    cx.b2 = st.b2[d.slot_base + slot];
    cx.j2d = cc.j2x2 * 0x1p22f;
    cx.j3d = cc.j3x2 * 0x1p22f;
    cx.beta = SHADOW ? st.beta[d.slot_base + slot] : 0.0;
    cx.key0 = st.seed;
    cx.key1 = (uint32_t)cd.iconf;
    cx.cnt_off = counting && !RC && !CB ? N / 4 : 0;
    const uint64_t S0 = clock_cycle(st) * (uint64_t)st.ptswapever;
    unsigned long long sh_prop = 0, sh_mis = 0;
    __syncthreads();

    for (int k = 0; k < st.ptswapever; ++k) {
        const uint64_t sweep = S0 + (uint64_t)k;
        cx.sweep_lo = (uint32_t)sweep;
        cx.sweep_hi = (uint32_t)(sweep >> 32);
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            cx.ctr1 = (uint32_t)c | ((uint32_t)d.rep << 2);
            switch (c) {
                case 0: colour_pass<WPH, NT, 0, SHADOW, HC, RC, CB>(lds32, cx, cc, H, sh_prop, sh_mis, tid, fa[0]); break;
                case 1: colour_pass<WPH, NT, 1, SHADOW, HC, RC, CB>(lds32, cx, cc, H, sh_prop, sh_mis, tid, fa[1]); break;
                case 2: colour_pass<WPH, NT, 2, SHADOW, HC, RC, CB>(lds32, cx, cc, H, sh_prop, sh_mis, tid, fa[2]); break;
                default: colour_pass<WPH, NT, 3, SHADOW, HC, RC, CB>(lds32, cx, cc, H, sh_prop, sh_mis, tid, fa[3]); break;
            }
            __syncthreads();
        }
    }
    if constexpr (CB) {
        // confnumflips(slot, site) += the counts in bits 3..7 of the sites'
        // bytes, which are cleared before anything else reads the lattice
        int4 *fl = reinterpret_cast<int4 *>(st.flips + cd.flip_off + (int64_t)slot * N);
        for (int i = tid; i < N / 4; i += NT) {
            const uint32_t w = lds32[i], n4 = (w >> 3) & 0x1F1F1F1Fu;
            if (n4 && active && counting) {
                int4 g = fl[i];
                g.x += (int)(n4 & 0xFFu);
                g.y += (int)((n4 >> 8) & 0xFFu);
                g.z += (int)((n4 >> 16) & 0xFFu);
                g.w += (int)(n4 >> 24);
                fl[i] = g;
            }
            lds32[i] = w & 0x04040404u;
        }
        __syncthreads();
    } else if constexpr (RC) {
        if (counting && active)
            flush_reg_counts<WPH, NT, HC>(st.flips + cd.flip_off + (int64_t)slot * N,
                                          st.flips_pend + cd.flip_off + (int64_t)slot * N, fa, tid);
    } else if (counting && active) {
        // confnumflips(slot, site) += this launch's counts (internal site order)
        // (in batches of four words, the counters of a batch all requested
        // before the first is added to)
        int4 *fl = reinterpret_cast<int4 *>(st.flips + cd.flip_off + (int64_t)slot * N);
        constexpr int KF = 4;
        for (int i0 = tid; i0 < N / 4; i0 += KF * NT) {
            int4 g[KF];
            uint32_t n4[KF];
#pragma unroll
            for (int q = 0; q < KF; ++q) {
                const int i = i0 + q * NT;
                n4[q] = i < N / 4 ? lds32[N / 4 + i] : 0u;
                if (n4[q]) g[q] = fl[i];
            }
#pragma unroll
            for (int q = 0; q < KF; ++q) {
                if (!n4[q]) continue;
                g[q].x += (int)(n4[q] & 0xFFu);
                g[q].y += (int)((n4[q] >> 8) & 0xFFu);
                g[q].z += (int)((n4[q] >> 16) & 0xFFu);
                g[q].w += (int)(n4[q] >> 24);
                fl[i0 + q * NT] = g[q];
            }
        }
    }

    // ---- observables of the final state (integer sums; phi part fixed point)
    int64_t v[OBS_N];
    const QPlanes qp = qplanes(st.qb + d.site_off, st.nsites);
    if constexpr (fast_zs<WPH, NT>() >= 2) {
        ObsFast a = {};
        for (int c = 0; c < 4; ++c) {
            if (c & 1) observe_fast<WPH, NT, 1>(lds32, c, H, qp, a, tid);
            else observe_fast<WPH, NT, 0>(lds32, c, H, qp, a, tid);
        }
        const int32_t nw = (int32_t)a.nw;
        v[OBS_M] = 2 * (int64_t)a.ps - 4 * (int64_t)nw;
        v[OBS_SX] = a.sxm;
        v[OBS_SY] = a.sym;
        v[OBS_BX] = 4 * (int64_t)nw - 2 * (int64_t)a.px;
        v[OBS_BY] = 4 * (int64_t)nw - 2 * (int64_t)a.py;
        v[OBS_BD] = 8 * (int64_t)nw - 2 * (int64_t)a.pd;
        v[OBS_BZ] = 4 * (int64_t)nw - 2 * (int64_t)a.pz;
        v[OBS_EQ] = 2 * qsum_total(a.q);             // the -4 sum q term is added after the reduction
    } else {
        ObsAcc a = {};
        for (int c = 0; c < 4; ++c) {
            if (c & 1) observe_plane<WPH, 1>(lds32, c, H, tid, NT, qp, a);
            else observe_plane<WPH, 0>(lds32, c, H, tid, NT, qp, a);
        }
        v[OBS_M] = a.m; v[OBS_SX] = a.sx; v[OBS_SY] = a.sy; v[OBS_BX] = a.bx; v[OBS_BY] = a.by;
        v[OBS_BD] = a.bd; v[OBS_BZ] = a.bz; v[OBS_EQ] = 2 * qsum_total(a.q);
    }
    if (SHADOW) {
        sh_prop = wave_sum<unsigned long long>(sh_prop);
        sh_mis = wave_sum<unsigned long long>(sh_mis);
        if ((tid & 63) == 0 && sh_prop && active) {
            atomicAdd(&st.shadow[2 * d.conf + 0], sh_prop);
            atomicAdd(&st.shadow[2 * d.conf + 1], sh_mis);
        }
    }
    for (int i = tid; active && i < N / 32; i += NT) {  // bytes 4b -> bits
        const uint4 v0 = reinterpret_cast<const uint4 *>(lds32)[2 * i], v1 = reinterpret_cast<const uint4 *>(lds32)[2 * i + 1];
        const uint32_t b[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        uint32_t W = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) W |= (((((b[k] >> 2) & 0x01010101u) * 0x01020408u) >> 24) & 0xFu) << (4 * k);
        reinterpret_cast<uint32_t *>(gsp)[i] = W;
    }
    block_reduce_obs<NT>(v, red, tid);
    if (tid == 0 && active) {
        v[OBS_EQ] -= 4 * cc.qsum;
        int64_t *o = st.obs + (int64_t)lat * OBS_N;
#pragma unroll
        for (int k = 0; k < OBS_N; ++k) o[k] = v[k];
    }
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8).
// The host's work list puts all replicas of a config (which share its
// disorder planes) on one XCD, so those planes are served from that XCD's L2,
// balances the XCDs by work and orders each XCD's lattices largest first.


// Waves per SIMD the resident kernel is compiled for (its register budget:
// 512 / waves VGPRs): for cubic lattices, the count at which one more block
// of NT threads fits on a CU when the lattices' LDS would allow it -- 512
// threads at L <= 32: 6 (three blocks; four L = 32 lattices fit the LDS), 576
// threads at L <= 40: 5 (two blocks), 256 threads at L <= 24: 6; otherwise 4
// (128 VGPRs).
template <int WPH, int NT, int HC>
constexpr int res_waves()
{
    if (HC == 0) return 4;
    if (NT == 512 && WPH <= 4) return 6;
    if (NT == 576 && WPH <= 5) return 5;
    if (NT == 256 && WPH <= 3) return 6;
    return 4;
}

// With flip counts, the 256-thread blocks of the unrolled passes keep each
// thread's counts in registers (8 per step and colour pass, at most 32):
// those blocks run at four waves per SIMD with room for them, and the LDS
// then holds the lattice alone, so as many blocks fit a CU as without counts.
template <int WPH, int NT, int HC>
constexpr bool reg_counts()
{
    return NT == 256 && unrolled_pairs<WPH, NT, HC>() && 8 * pair_steps<WPH, NT, HC>() <= 32 &&
           res_waves<WPH, NT, HC>() == 4;
}

// One geometry.  HC > 0: every lattice of the launch has height HC (cubic
// lattices, H = L).  RC: flip counts in registers (reg_counts), none in the
// LDS.  CB: flip counts in the sites' own bytes (byte_counts).
template <int WPH, int NT, bool SHADOW, int HC = 0, bool RC = false, bool CB = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SHADOW ? 1 : res_waves<WPH, NT, HC>(), 8)))
void k_sweep_resident(DeviceState st, int32_t ord0, int32_t nord)
{
    // the lattice at LDS address 0 (so neighbour addresses need no base), the
    // reduction scratch right after it
    extern __shared__ uint32_t lds32[];
    const int lat = st.order[ord0 + blockIdx.x];
    if (lat < 0) return;
    const LatDesc d = st.lats[lat];
    const int lds_words = (st.flips && !RC && !CB ? 2 : 1) * (d.N / 4);     // lattice (and flip counts)
    resident_body<WPH, NT, SHADOW, HC, RC, CB>(st, d, lat, lds32,
                                           reinterpret_cast<int64_t (*)[OBS_N]>(lds32 + lds_words));
}

// A bundle: G lattices of size L = 8 WPH swept side by side by one block of
// 1024 threads, a group of 1024 / G threads and a part of the LDS each (small
// lattices keep few of 1024 threads busy: L = 16 has 128 words per layer pair
// of a colour plane).  The bundle's lattices are lat .. lat + cnt - 1; groups
// from cnt on repeat lattice lat without writing to global memory.
template <int WPH, int G, int HC>
__device__ __forceinline__ void bundle_body(const DeviceState &st, int lat, int cnt, uint32_t *lds32, int words)
{
    constexpr int NTG = 1024 / G;
    static_assert(NTG % 64 == 0, "a group is whole waves");
    // the group is wave-uniform (whole waves), so its lattice's descriptor
    // and everything derived from it stay scalar
    const int g = __builtin_amdgcn_readfirstlane((int)threadIdx.x / NTG), tid = threadIdx.x - g * NTG;
    const bool active = g < cnt;
    const int my = active ? lat + g : lat;
    const LatDesc d = st.lats[my];
    resident_body<WPH, NTG, false, HC>(st, d, my, lds32 + g * words,
                                       reinterpret_cast<int64_t (*)[OBS_N]>(lds32 + G * words), tid, active);
}

// Uneven batches: lattices of several sizes (L in {8, ..., 48}) in one launch.
// CUBIC: every lattice of the launch has H = L.  With 1024 threads (FP32) a
// work-list entry may be a bundle of 2 .. res_bundle(L) lattices of one
// config (entry = first lattice + (count - 1) << BUNDLE_SHIFT), which needs
// res_bundle(L) lattices' LDS.
template <int NT, bool SHADOW, bool CUBIC = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(SHADOW ? 1 : 4, 8)))
void k_sweep_resident_mixed(DeviceState st, int32_t ord0, int32_t nord)
{
    extern __shared__ uint32_t lds32[];
    const int e = st.order[ord0 + blockIdx.x];
    if (e < 0) return;
    const int lat = e & BUNDLE_LAT_MASK;
    const LatDesc d = st.lats[lat];
    const int words = (st.flips ? 2 : 1) * (d.N / 4);   // a lattice (and its flip counts) in LDS
    if constexpr (NT == 1024 && !SHADOW) {
        // a bundle of one lattice takes the single-lattice path below (its
        // LDS is one lattice's, as for any list without bundles)
        const int cnt = (e >> BUNDLE_SHIFT) + 1;
        if (cnt > 1) switch (d.L) {
            case 16: bundle_body<2, res_bundle(16), CUBIC ? 16 : 0>(st, lat, cnt, lds32, words); return;
            case 24: bundle_body<3, res_bundle(24), CUBIC ? 24 : 0>(st, lat, cnt, lds32, words); return;
            case 32: bundle_body<4, res_bundle(32), CUBIC ? 32 : 0>(st, lat, cnt, lds32, words); return;
            default: break;
        }
    }
    int64_t (*red)[OBS_N] = reinterpret_cast<int64_t (*)[OBS_N]>(lds32 + words);
    switch (d.L) {
        case 8: resident_body<1, NT, SHADOW, CUBIC ? 8 : 0>(st, d, lat, lds32, red); break;
        case 16: resident_body<2, NT, SHADOW, CUBIC ? 16 : 0>(st, d, lat, lds32, red); break;
        case 24: resident_body<3, NT, SHADOW, CUBIC ? 24 : 0>(st, d, lat, lds32, red); break;
        case 32: resident_body<4, NT, SHADOW, CUBIC ? 32 : 0>(st, d, lat, lds32, red); break;
        case 40: resident_body<5, NT, SHADOW, CUBIC ? 40 : 0>(st, d, lat, lds32, red); break;
        case 48: resident_body<6, NT, SHADOW, CUBIC ? 48 : 0>(st, d, lat, lds32, red); break;
        default: break;   // the host only sends L in {8,...,48}, L % 8 == 0, here
    }
}

// reduction scratch after the lattice in the resident sweep's dynamic LDS
constexpr size_t RES_STATIC_LDS = sizeof(int64_t) * (1024 / 64) * OBS_N;

// The FP32 kernel of cubic lattices of size 8 WPH; with flip counts the one
// that keeps them in registers where there is one.
template <int WPH, int NT>
ResKernel cubic_kernel(bool flips)
{
    if constexpr (reg_counts<WPH, NT, 8 * WPH>())
        if (flips) return k_sweep_resident<WPH, NT, false, 8 * WPH, true>;
    if constexpr (byte_counts(8 * WPH))
        if (flips) return k_sweep_resident<WPH, NT, false, 8 * WPH, false, true>;
    return k_sweep_resident<WPH, NT, false, 8 * WPH>;
}

// The kernel for lattices of size L (0: a launch of several sizes).  cubic:
// every lattice of the launch has H = L (the FP32 kernels then take the
// height as a compile-time constant).  flips: the engine counts flips.
template <int NT, bool SHADOW>
ResKernel resident_kernel(int32_t L, bool cubic, bool flips)
{
    if (cubic && !SHADOW) {
        switch (L) {
            case 8: return cubic_kernel<1, NT>(flips);
            case 16: return cubic_kernel<2, NT>(flips);
            case 24: return cubic_kernel<3, NT>(flips);
            case 32: return cubic_kernel<4, NT>(flips);
            case 40: return cubic_kernel<5, NT>(flips);
            case 48: return cubic_kernel<6, NT>(flips);
            default: return k_sweep_resident_mixed<NT, false, true>;
        }
    }
    switch (L) {
        case 8: return k_sweep_resident<1, NT, SHADOW>;
        case 16: return k_sweep_resident<2, NT, SHADOW>;
        case 24: return k_sweep_resident<3, NT, SHADOW>;
        case 32: return k_sweep_resident<4, NT, SHADOW>;
        case 40: return k_sweep_resident<5, NT, SHADOW>;
        case 48: return k_sweep_resident<6, NT, SHADOW>;
        default: return k_sweep_resident_mixed<NT, SHADOW>;
    }
}

}  // namespace

}  // namespace j1j2
