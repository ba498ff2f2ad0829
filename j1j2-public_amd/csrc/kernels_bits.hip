// kernels_bits.hip -- LDS-resident FP32 sweep with one bit per site, for
// lattices whose byte layout does not fit one CU's LDS (L % 8 == 0,
// 48 < L <= 96; L = 96, H = 96 takes 144 KiB).  Same update rule, Philox
// counters, disorder planes and observables as the byte-layout kernel in
// kernels.hip, so the results are bit-identical to it and to the oracle.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "sweep_common.h"

namespace j1j2 {

namespace {

// LDS layout of one lattice: plane c (colour), unit u = z*HW + y/2 (a half-row
// of HW = L/2 same-colour sites), one 64-bit word per unit at dwords
// 2*(c*NU + u), NU = HW*H units per plane.  Bit 1 + m holds site m (m = x/2)
// of the half-row, bit 0 repeats site HW-1 and bit HW+1 repeats site 0, so the
// four sites m0+s .. m0+s+3 (m0 = 4j, s in {-1, 0, +1}) are always bits
// m0+1+s .. m0+4+s of the word, wrap included.  One spare dword follows the
// last word (the two-dword reads of the last unit touch it).
template <int L>
struct BitGeo {
    static constexpr int HW = L / 2;                 // sites per half-row
    static constexpr int G = HW / 4;                 // 4-site groups per half-row (= Philox groups)
    static constexpr int TPL = G * HW;               // groups per z-layer of a plane: one thread each
    static constexpr int NT = (TPL + 63) / 64 * 64;  // threads per block (idle tail lanes)
};

size_t bits_lds_bytes(int32_t L, int32_t H) { return (size_t)32 * (L / 2) * H + 8; }

// Four sites starting at bit `pos` of the unit word at dword `dw`.
__device__ __forceinline__ uint32_t nib(const uint32_t *lds32, int dw, int pos)
{
    const int d = dw + (pos >> 5);
    return __builtin_amdgcn_alignbit(lds32[d + 1], lds32[d], (uint32_t)(pos & 31)) & 0xFu;
}

// Site k of a nibble into bit 8k+2 (the byte layout of the LDS copy in
// kernels.hip: byte value 4b).
__device__ __forceinline__ uint32_t spread4(uint32_t n) { return (n * 0x00810204u) & 0x04040404u; }

// Neighbour units of unit u = z*HW + yh of plane c (y parity p = ((c>>1) ^ z) & 1).
struct BitNb {
    int uyp, uym, uzp, uzm, p;
};

template <int HW>
__device__ __forceinline__ BitNb bit_nb(int c, int z, int yh, int H)
{
    BitNb n;
    const int u = z * HW + yh;
    n.p = ((c >> 1) ^ z) & 1;
    if (n.p) {                        // y odd: y+1 is the next unit, y-1 this one
        n.uyp = (yh == HW - 1) ? u - (HW - 1) : u + 1;
        n.uym = u;
    } else {                          // y even: y+1 is this unit, y-1 the previous one
        n.uyp = u;
        n.uym = (yh == 0) ? u + (HW - 1) : u - 1;
    }
    n.uzp = (z == H - 1) ? u - (H - 1) * HW : u + HW;
    n.uzm = (z == 0) ? u + (H - 1) * HW : u - HW;
    return n;
}

// The neighbour nibbles of one 4-site group, spread to the byte layout.
struct BitSites {
    uint32_t S;                       // own spins, nibble (bit k = site k)
    uint32_t xp, xm, yp, ym, zp, zm;  // spread: bit 8k+2 = neighbour of site k
    uint32_t d1p, d1m, d2p, d2m;      // diagonals of rows y+1 (d1) and y-1 (d2)
};

// HX: x parity of the plane's sites; for even sites +x is the other plane's
// site m and -x its site m-1, for odd sites -x is m and +x is m+1.
template <int HW, int HX>
__device__ __forceinline__ BitSites bit_sites(const uint32_t *lds32, int c, int NU, int u, const BitNb &n,
                                              int m0)
{
    const int pos0 = m0 + 1, posx = m0 + 1 + (HX ? 1 : -1);
    const int p1 = 2 * (c ^ 1) * NU, p2 = 2 * (c ^ 2) * NU, p3 = 2 * (c ^ 3) * NU;
    BitSites s;
    s.S = nib(lds32, 2 * (c * NU + u), pos0);
    const uint32_t x0 = spread4(nib(lds32, p1 + 2 * u, pos0)), x1 = spread4(nib(lds32, p1 + 2 * u, posx));
    s.xp = HX ? x1 : x0;
    s.xm = HX ? x0 : x1;
    s.yp = spread4(nib(lds32, p2 + 2 * n.uyp, pos0));
    s.ym = spread4(nib(lds32, p2 + 2 * n.uym, pos0));
    s.zp = spread4(nib(lds32, p2 + 2 * n.uzp, pos0));
    s.zm = spread4(nib(lds32, p2 + 2 * n.uzm, pos0));
    const uint32_t a0 = spread4(nib(lds32, p3 + 2 * n.uyp, pos0)), a1 = spread4(nib(lds32, p3 + 2 * n.uyp, posx));
    const uint32_t b0 = spread4(nib(lds32, p3 + 2 * n.uym, pos0)), b1 = spread4(nib(lds32, p3 + 2 * n.uym, posx));
    s.d1p = HX ? a1 : a0;
    s.d1m = HX ? a0 : a1;
    s.d2p = HX ? b1 : b0;
    s.d2m = HX ? b0 : b1;
    return s;
}

__device__ __forceinline__ uint32_t nib_bit(uint32_t n, int k) { return (n >> k) & 1u; }
__device__ __forceinline__ uint32_t sp_bit(uint32_t w, int k) { return (w >> (8 * k + 2)) & 1u; }

// One 4-site group of a colour pass of the bit layout: thread (yh, j) owns
// group j of unit (z, yh); its flips go to the unit word with an LDS atomic
// xor (the G threads of a unit share the word; nobody reads plane c in pass c).
template <int L, int HX, bool SHADOW>
__device__ __forceinline__ void bit_group(uint32_t *lds32, const SweepCtx &cx, const ConfConst &cc, int c, int H,
                                          int z, int yh, int j, unsigned long long &sh_prop,
                                          unsigned long long &sh_mis)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x, m0 = 4 * j;
    const int NU = HW * H, nq = NU * G;
    unsigned long long *lds64 = reinterpret_cast<unsigned long long *>(lds32);
    const int u = z * HW + yh;
    const BitNb n = bit_nb<HW>(c, z, yh, H);
    const BitSites s = bit_sites<HW, HX>(lds32, c, NU, u, n, m0);
    const uint32_t D = (s.d1p + s.d1m) + (s.d2p + s.d2m);     // bytewise 4 * (0..4)
    const uint32_t Z = s.zp + s.zm;                           // bytewise 4 * (0..2)
    const uint32_t de = lo_even(D), dod = lo_odd(D), ze = lo_even(Z), zod = lo_odd(Z);
    const uint32_t xpe = f16_even(s.xp), xpo = f16_odd(s.xp), xme = f16_even(s.xm), xmo = f16_odd(s.xm);
    const uint32_t ype = f16_even(s.yp), ypo = f16_odd(s.yp), yme = f16_even(s.ym), ymo = f16_odd(s.ym);
    const int t = z * TPL + tid;                              // plane-relative group = Philox counter
    const int own = c * nq + t;
    const uint32_t ob = (uint32_t)own * 16u;
    const float4 kc = ld16(cx.kc4, ob), kxp = ld16(cx.kxp4, ob), kxm = ld16(cx.kxm4, ob),
                 kyp = ld16(cx.kyp4, ob), kym = ld16(cx.kym4, ob);
    const U4 rn = philox4x32_10_u((uint32_t)t, cx.ctr1, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
    const float kca[4] = {kc.x, kc.y, kc.z, kc.w};
    const float kxpa[4] = {kxp.x, kxp.y, kxp.z, kxp.w};
    const float kxma[4] = {kxm.x, kxm.y, kxm.z, kxm.w};
    const float kypa[4] = {kyp.x, kyp.y, kyp.z, kyp.w};
    const float kyma[4] = {kym.x, kym.y, kym.z, kym.w};
    const uint32_t ra[4] = {rn.x, rn.y, rn.z, rn.w};
    uint32_t f = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float h = kca[k];
        h = __builtin_fmaf(f16_site(xpe, xpo, k), kxpa[k], h);
        h = __builtin_fmaf(f16_site(xme, xmo, k), kxma[k], h);
        h = __builtin_fmaf(f16_site(ype, ypo, k), kypa[k], h);
        h = __builtin_fmaf(f16_site(yme, ymo, k), kyma[k], h);
        h = __builtin_fmaf(f16_site(de, dod, k), cx.j2d, h);
        h = __builtin_fmaf(f16_site(ze, zod, k), cx.j3d, h);
        const bool acc = decide_f32_word(h, s.S, k, cx.b2, ra[k]);
        if constexpr (SHADOW) {
            const bool up = nib_bit(s.S, k) != 0;
            const int m = m0 + k;                             // position in the half-row
            const int mo_p = HX ? ((m == HW - 1) ? 0 : m + 1) : m;
            const int mo_m = HX ? m : ((m == 0) ? HW - 1 : m - 1);
            const double sv[10] = {pm1(sp_bit(s.yp, k)), pm1(sp_bit(s.ym, k)), pm1(sp_bit(s.xp, k)),
                                   pm1(sp_bit(s.xm, k)), pm1(sp_bit(s.d1p, k)), pm1(sp_bit(s.d1m, k)),
                                   pm1(sp_bit(s.d2p, k)), pm1(sp_bit(s.d2m, k)), pm1(sp_bit(s.zp, k)),
                                   pm1(sp_bit(s.zm, k))};
            // site indices in the colour-plane order of the disorder planes
            const int64_t q1 = (int64_t)((c ^ 1) * NU + u) * HW, q2 = (int64_t)(c ^ 2) * NU * HW;
            const double pv[5] = {cx.p64[4 * (int64_t)own + k], cx.p64[q2 + (int64_t)n.uyp * HW + m],
                                  cx.p64[q2 + (int64_t)n.uym * HW + m], cx.p64[q1 + mo_p], cx.p64[q1 + mo_m]};
            const bool a64 = decide_f64(field_f64(sv, pv, cc), up, cx.beta, ra[k]);
            sh_prop += 1;
            sh_mis += (a64 != acc);
        }
        f |= (acc ? 1u : 0u) << k;
    }
    unsigned long long mask = (unsigned long long)f << (m0 + 1);
    if (j == 0) mask |= (unsigned long long)(f & 1u) << (HW + 1);     // copy of site 0
    if (j == G - 1) mask |= (unsigned long long)(f >> 3);             // copy of site HW-1
    atomicXor(&lds64[c * NU + u], mask);
}

// One colour pass: every thread steps through the z-layers, two per trip (the
// two groups are independent, which the scheduler interleaves).
template <int L, int HX, bool SHADOW>
__device__ __forceinline__ void bit_pass(uint32_t *lds32, const SweepCtx &cx, const ConfConst &cc, int c, int H,
                                         unsigned long long &sh_prop, unsigned long long &sh_mis)
{
    using Gm = BitGeo<L>;
    constexpr int G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x;
    if (tid >= TPL) return;
    const int yh = tid / G, j = tid - yh * G;
    for (int z = 0; z < H; z += 2) {                 // H is even
        bit_group<L, HX, SHADOW>(lds32, cx, cc, c, H, z, yh, j, sh_prop, sh_mis);
        bit_group<L, HX, SHADOW>(lds32, cx, cc, c, H, z + 1, yh, j, sh_prop, sh_mis);
    }
}

// Per-thread partial observables of the bit layout (same sums as ObsFast in
// kernels.hip, one 4-site group per step).
template <int L, int HX>
__device__ __forceinline__ void bit_observe(const uint32_t *lds32, int c, int H, const int32_t *qh,
                                            const int32_t *ql, ObsFast &a)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x;
    if (tid >= TPL) return;
    const int yh = tid / G, j = tid - yh * G, m0 = 4 * j;
    const int NU = HW * H, nq = NU * G;
    uint32_t ps = 0, px = 0, py = 0, pz = 0, pd = 0, nw = 0;
    int32_t sxm = 0, sym = 0, el = 0;
    int64_t eh = 0;
    for (int z = 0; z < H; ++z) {
        const int u = z * HW + yh;
        const BitNb n = bit_nb<HW>(c, z, yh, H);
        const BitSites s = bit_sites<HW, HX>(lds32, c, NU, u, n, m0);
        const uint32_t S = spread4(s.S);
        const uint32_t sx0 = S ^ s.xp, sy0 = S ^ s.yp;
        const int pcs = __builtin_popcount(s.S);
        ps += pcs;
        px += __builtin_popcount(sx0);
        py += __builtin_popcount(sy0);
        pz += __builtin_popcount(S ^ s.zp);
        pd += __builtin_popcount(S ^ s.d1p) + __builtin_popcount(S ^ s.d1m);
        nw += 1;
        const int m = 2 * pcs - 4;
        sxm += HX ? -m : m;
        sym += n.p ? -m : m;
        const uint32_t ax = sx0 + (S ^ s.xm), ay = sy0 + (S ^ s.ym);   // bytes 0, 4, 8
        const uint32_t e = ((ay + 0x08080808u) - ax) >> 2;             // bytes l/2 + 2
        const uint32_t ob = (uint32_t)(c * nq + z * TPL + tid) * 16u;
        const int4 h4 = *reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(qh) + ob);
        const int4 l4 = *reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(ql) + ob);
        eh += __mul24(h4.x, (int)(e & 0xffu)) + __mul24(h4.y, (int)((e >> 8) & 0xffu)) +
              __mul24(h4.z, (int)((e >> 16) & 0xffu)) + __mul24(h4.w, (int)(e >> 24));
        el += __mul24(l4.x, (int)(e & 0xffu)) + __mul24(l4.y, (int)((e >> 8) & 0xffu)) +
              __mul24(l4.z, (int)((e >> 16) & 0xffu)) + __mul24(l4.w, (int)(e >> 24));
    }
    a.ps += ps; a.px += px; a.py += py; a.pz += pz; a.pd += pd; a.nw += nw;
    a.sxm += sxm; a.sym += sym; a.eh += eh; a.el += el;
}

// Accumulate_local_nematic_average (J1J2_6tv.f90:1009-1025) from the bit layout.
template <int L, int HX>
__device__ __forceinline__ void bit_locnem(const uint32_t *lds32, int c, int H, int32_t *locspin, int32_t *locnem)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x;
    if (tid >= TPL) return;
    const int yh = tid / G, j = tid - yh * G, m0 = 4 * j;
    const int NU = HW * H, nq = NU * G;
    for (int z = 0; z < H; ++z) {
        const int u = z * HW + yh;
        const BitNb n = bit_nb<HW>(c, z, yh, H);
        const BitSites s = bit_sites<HW, HX>(lds32, c, NU, u, n, m0);
        const uint32_t S = spread4(s.S);
        const uint32_t ax = (S ^ s.xp) + (S ^ s.xm), ay = (S ^ s.yp) + (S ^ s.ym);
        const int dw = c * nq + z * TPL + tid;
        int4 ls = reinterpret_cast<int4 *>(locspin)[dw], lv = reinterpret_cast<int4 *>(locnem)[dw];
        int *lsp = reinterpret_cast<int *>(&ls), *lnp = reinterpret_cast<int *>(&lv);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lsp[k] += (int)nib_bit(s.S, k) * 2 - 1;
            lnp[k] += 2 * ((int)((ay >> (8 * k + 2)) & 0xffu) - (int)((ax >> (8 * k + 2)) & 0xffu));
        }
        reinterpret_cast<int4 *>(locspin)[dw] = ls;
        reinterpret_cast<int4 *>(locnem)[dw] = lv;
    }
}

template <int L, bool SHADOW>
__device__ __forceinline__ void bits_body(const DeviceState &st, const LatDesc &d, int lat, uint32_t *lds32,
                                          int64_t (*red)[OBS_N])
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, NT = Gm::NT;
    const int H = d.H, tid = threadIdx.x;
    const int NU = HW * H;
    uint8_t *gsp = st.spins + d.spin_off;

    // bytes (colour-plane order, unit U = c*NU + u at byte U*HW) -> unit words
    for (int U = tid; U < 4 * NU; U += NT) {
        const uint8_t *src = gsp + (int64_t)U * HW;
        unsigned long long w = 0;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const uint32_t b = *reinterpret_cast<const uint32_t *>(src + 4 * k);
            w |= (unsigned long long)((b * 0x10204080u) >> 28) << (1 + 4 * k);   // bytes 0..3 -> bits
        }
        w |= (w >> HW) & 1ull;                                 // site HW-1 at bit 0
        w |= (w & 2ull) << HW;                                 // site 0 at bit HW+1
        reinterpret_cast<unsigned long long *>(lds32)[U] = w;
    }
    if (tid == 0) lds32[8 * NU] = 0;                            // the spare dword

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
    const uint64_t S0 = st.clk->cycle * (uint64_t)st.ptswapever;
    unsigned long long sh_prop = 0, sh_mis = 0;
    __syncthreads();
    if (st.fuse_locnem && slot == st.ptmeasindex - 1 && st.locnem_pending[d.conf]) {
        for (int c = 0; c < 4; ++c) {
            if (c & 1) bit_locnem<L, 1>(lds32, c, H, st.locspin + d.site_off, st.locnem + d.site_off);
            else bit_locnem<L, 0>(lds32, c, H, st.locspin + d.site_off, st.locnem + d.site_off);
        }
        if (tid == 0) st.locsamples[d.conf] += 1;
        __syncthreads();
    }

    for (int k = 0; k < st.ptswapever; ++k) {
        const uint64_t sweep = S0 + (uint64_t)k;
        cx.sweep_lo = (uint32_t)sweep;
        cx.sweep_hi = (uint32_t)(sweep >> 32);
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            cx.ctr1 = (uint32_t)c | ((uint32_t)d.rep << 2);
            if (c & 1) bit_pass<L, 1, SHADOW>(lds32, cx, cc, c, H, sh_prop, sh_mis);
            else bit_pass<L, 0, SHADOW>(lds32, cx, cc, c, H, sh_prop, sh_mis);
            __syncthreads();
        }
    }

    ObsFast a = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 4; ++c) {
        if (c & 1) bit_observe<L, 1>(lds32, c, H, st.qh + d.site_off, st.ql + d.site_off, a);
        else bit_observe<L, 0>(lds32, c, H, st.qh + d.site_off, st.ql + d.site_off, a);
    }
    const int64_t nw = (int64_t)a.nw;
    int64_t v[OBS_N];
    v[OBS_M] = 2 * (int64_t)a.ps - 4 * nw;
    v[OBS_SX] = a.sxm;
    v[OBS_SY] = a.sym;
    v[OBS_BX] = 4 * nw - 2 * (int64_t)a.px;
    v[OBS_BY] = 4 * nw - 2 * (int64_t)a.py;
    v[OBS_BD] = 8 * nw - 2 * (int64_t)a.pd;
    v[OBS_BZ] = 4 * nw - 2 * (int64_t)a.pz;
    v[OBS_EQ] = 2 * (a.eh * 65536 + a.el);
    if (SHADOW) {
        sh_prop = wave_sum<unsigned long long>(sh_prop);
        sh_mis = wave_sum<unsigned long long>(sh_mis);
        if ((tid & 63) == 0 && sh_prop) {
            atomicAdd(&st.shadow[2 * d.conf + 0], sh_prop);
            atomicAdd(&st.shadow[2 * d.conf + 1], sh_mis);
        }
    }
    // unit words -> bytes
    for (int U = tid; U < 4 * NU; U += NT) {
        const unsigned long long w = reinterpret_cast<const unsigned long long *>(lds32)[U];
        uint8_t *dst = gsp + (int64_t)U * HW;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const uint32_t n4 = (uint32_t)(w >> (1 + 4 * k)) & 0xFu;
            *reinterpret_cast<uint32_t *>(dst + 4 * k) = (n4 * 0x00204081u) & 0x01010101u;
        }
    }
    block_reduce_obs<NT>(v, red);
    if (tid == 0) {
        v[OBS_EQ] -= 4 * cc.qsum;
        int64_t *o = st.obs + (int64_t)lat * OBS_N;
#pragma unroll
        for (int k = 0; k < OBS_N; ++k) o[k] = v[k];
    }
}

// At most 168 VGPRs (three waves per SIMD): one lattice of L >= 64 fills most
// of a CU's LDS, so a CU holds one or two blocks of at most 9 waves.
template <int L, bool SHADOW>
__global__ __launch_bounds__(BitGeo<L>::NT) __attribute__((amdgpu_waves_per_eu(SHADOW ? 1 : 3, 8)))
void k_sweep_bits(DeviceState st, int32_t ord0, int32_t nord, int32_t lds_bytes)
{
    // the lattice's unit words from LDS address 0, the reduction scratch after them
    extern __shared__ uint32_t lds32[];
    const int lat = st.order[ord0 + blockIdx.x];
    if (lat < 0) return;
    const LatDesc d = st.lats[lat];
    int64_t (*red)[OBS_N] = reinterpret_cast<int64_t (*)[OBS_N]>(reinterpret_cast<char *>(lds32) + lds_bytes);
    bits_body<L, SHADOW>(st, d, lat, lds32, red);
}

using BitKernel = void (*)(DeviceState, int32_t, int32_t, int32_t);

template <bool SHADOW>
BitKernel bits_kernel(int32_t L)
{
    switch (L) {
        case 56: return k_sweep_bits<56, SHADOW>;
        case 64: return k_sweep_bits<64, SHADOW>;
        case 72: return k_sweep_bits<72, SHADOW>;
        case 80: return k_sweep_bits<80, SHADOW>;
        case 88: return k_sweep_bits<88, SHADOW>;
        case 96: return k_sweep_bits<96, SHADOW>;
        default: return nullptr;
    }
}

int bits_threads(int32_t L)
{
    switch (L) {
        case 56: return BitGeo<56>::NT;
        case 64: return BitGeo<64>::NT;
        case 72: return BitGeo<72>::NT;
        case 80: return BitGeo<80>::NT;
        case 88: return BitGeo<88>::NT;
        case 96: return BitGeo<96>::NT;
        default: return 0;
    }
}

constexpr size_t BITS_RED_LDS = sizeof(int64_t) * (1024 / 64) * OBS_N;

}  // namespace

bool bits_supported(int32_t L) { return L > 48 && L <= 96 && (L % 8) == 0; }

size_t bits_lds_need(int32_t L, int32_t H) { return (bits_lds_bytes(L, H) + 7) / 8 * 8 + BITS_RED_LDS; }

hipError_t setup_bits(size_t max_lds)
{
    for (int L : {56, 64, 72, 80, 88, 96})
        for (bool sh : {false, true}) {
            hipError_t e = hipFuncSetAttribute((const void *)(sh ? bits_kernel<true>(L) : bits_kernel<false>(L)),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

hipError_t launch_sweep_bits(const DeviceState &st, int32_t ord0, int32_t nord, int32_t L, int32_t H, bool shadow,
                             hipStream_t s)
{
    if (nord <= 0) return hipSuccess;
    const BitKernel k = shadow ? bits_kernel<true>(L) : bits_kernel<false>(L);
    if (!k) return hipErrorInvalidValue;
    const int32_t lds = (int32_t)((bits_lds_bytes(L, H) + 7) / 8 * 8);
    hipLaunchKernelGGL(k, dim3(nord), dim3(bits_threads(L)), (size_t)lds + BITS_RED_LDS, s, st, ord0, nord, lds);
    return hipGetLastError();
}

}  // namespace j1j2
