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
// of HW = L/2 same-colour sites), two dwords per unit at dwords 2*(c*NU + u),
// NU = HW*H units per plane.  Dword 0 holds sites 0..N0-1 (m = x/2) at bits
// 1..N0, dword 1 sites N0..HW-1 at bits 1..N1 (N0 = 4 ceil(G/2) and N1 = HW-N0
// sites, G = HW/4 groups); bit 0 of each dword repeats the site before its
// first one (dword 0: site HW-1, dword 1: site N0-1) and bit N+1 the site
// after its last one (dword 0: site N0, dword 1: site 0).  So the four sites
// of a group and the same window shifted by one site either way (its x
// neighbours, wrap included) are one 4-bit field of one dword.  The lattice's
// spins in HBM are the same unit words (unit_word_bit), so loading and storing
// a lattice is a copy of 8 bytes per half-row.
//
// A lattice may instead be split into S z-slabs, one block each (all of them
// resident at once): a slab holds its own layers plus one halo layer below and
// above per plane; after colour pass c it publishes its two boundary layers of
// plane c as exchange words in global memory, each carrying a version number,
// and during the pass before c' it waits for its neighbours' words of plane
// c'^2 (the only plane whose z-neighbours pass c' reads) with the version of
// that plane's last update.  A slot is never overwritten before it is read: a slab
// republishes plane c only after a pass that waits for the reader's later
// publication.  64-bit agent-scope atomic loads and stores carry each word
// whole, so no fences are needed.
template <int L>
struct BitGeo {
    static constexpr int HW = L / 2;                 // sites per half-row
    static constexpr int G = HW / 4;                 // 4-site groups per half-row
    static constexpr int TPL = G * HW;               // groups per z-layer of a plane: one thread each
    static constexpr int NT = (TPL + 63) / 64 * 64;  // threads per block (idle tail lanes)
    static constexpr int G0 = (G + 1) / 2;           // groups in dword 0 of a unit
    static constexpr int N0 = 4 * G0, N1 = HW - N0;  // sites in dwords 0 and 1 (<= 24 each)
};

// The two dwords of a unit from its sites (bit m of v = site m, m < HW).
template <int L>
__device__ __forceinline__ uint2 unit_from_sites(unsigned long long v)
{
    constexpr int N0 = BitGeo<L>::N0, N1 = BitGeo<L>::N1;
    const uint32_t lo = (uint32_t)v & ((1u << N0) - 1u), hi = (uint32_t)(v >> N0) & ((1u << N1) - 1u);
    return make_uint2((lo << 1) | (hi >> (N1 - 1)) | ((hi & 1u) << (N0 + 1)),
                      (hi << 1) | (lo >> (N0 - 1)) | ((lo & 1u) << (N1 + 1)));
}

// The sites of a unit (bit m = site m) from its two dwords.
template <int L>
__device__ __forceinline__ unsigned long long sites_from_unit(uint2 d)
{
    constexpr int N0 = BitGeo<L>::N0, N1 = BitGeo<L>::N1;
    return (unsigned long long)((d.x >> 1) & ((1u << N0) - 1u)) |
           ((unsigned long long)((d.y >> 1) & ((1u << N1) - 1u)) << N0);
}

// Four sites starting at bit `pos` of LDS dword `dw`.
__device__ __forceinline__ uint32_t nib(const uint32_t *lds32, int dw, int pos)
{
    return __builtin_amdgcn_ubfe(lds32[dw], (uint32_t)pos, 4u);
}

// Site k of a nibble into bit 8k+2 (the byte layout of the LDS copy in
// kernels.hip: byte value 4b).
__device__ __forceinline__ uint32_t spread4(uint32_t n) { return (n * 0x00810204u) & 0x04040404u; }

// Neighbour units of LDS unit u of plane c (global layer zg sets the y parity
// p = ((c>>1) ^ zg) & 1).  Whole lattice: u = zg*HW + yh, z wraps; slab: u =
// (zl+1)*HW + yh (zl the slab-local layer), z neighbours are the next and
// previous layers (halos at the slab's ends).
struct BitNb {
    int u, uyp, uym, uzp, uzm, p;
};

template <int HW, bool SLAB>
__device__ __forceinline__ BitNb bit_nb(int c, int zg, int zl, int yh, int H)
{
    BitNb n;
    const int u = (SLAB ? zl + 1 : zl) * HW + yh;
    n.u = u;
    n.p = ((c >> 1) ^ zg) & 1;
    if (n.p) {                        // y odd: y+1 is the next unit, y-1 this one
        n.uyp = (yh == HW - 1) ? u - (HW - 1) : u + 1;
        n.uym = u;
    } else {                          // y even: y+1 is this unit, y-1 the previous one
        n.uyp = u;
        n.uym = (yh == 0) ? u + (HW - 1) : u - 1;
    }
    if constexpr (SLAB) {
        n.uzp = u + HW;
        n.uzm = u - HW;
    } else {
        n.uzp = (zl == H - 1) ? u - (H - 1) * HW : u + HW;
        n.uzm = (zl == 0) ? u + (H - 1) * HW : u - HW;
    }
    return n;
}

// The neighbour nibbles of one 4-site group, spread to the byte layout.
struct BitSites {
    uint32_t S;                       // own spins, nibble (bit k = site k)
    uint32_t xp, xm, yp, ym, zp, zm;  // spread: bit 8k+2 = neighbour of site k
    uint32_t d1p, d1m, d2p, d2m;      // diagonals of rows y+1 (d1) and y-1 (d2)
};

// Where group j of a unit sits: dword h of the unit, sites at bits pos..pos+3.
struct GroupPos {
    int h, pos;
};

template <int L>
__device__ __forceinline__ GroupPos group_pos(int j)
{
    constexpr int G0 = BitGeo<L>::G0;
    const int h = j >= G0 ? 1 : 0;
    return GroupPos{h, 4 * (j - h * G0) + 1};
}

// HX: x parity of the plane's sites; for even sites +x is the other plane's
// site m and -x its site m-1, for odd sites -x is m and +x is m+1.  NU: units
// per plane in LDS; gp: the group's place in its units.
template <int HW, int HX>
__device__ __forceinline__ BitSites bit_sites(const uint32_t *lds32, int c, int NU, const BitNb &n, GroupPos gp)
{
    const int pos0 = gp.pos, posx = gp.pos + (HX ? 1 : -1);
    const int p1 = 2 * (c ^ 1) * NU + gp.h, p2 = 2 * (c ^ 2) * NU + gp.h, p3 = 2 * (c ^ 3) * NU + gp.h;
    const int u = n.u;
    BitSites s;
    s.S = nib(lds32, 2 * (c * NU + u) + gp.h, pos0);
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

// The operands of the field sums of one 4-site group (fast path): the own
// nibble, the f16 views of the x and y neighbours (site bit at bit 10 of its
// half, as f16_even / f16_odd of the byte form) and of the bond counts Z and
// D.  The x neighbours and the diagonals come from 5-site windows of one dword
// starting at the lower x (bit k of a window is the -x side of site k, bit
// k + 1 its +x side), so each is one extract: a multiply by 0x00800200 puts
// window bit k at bits 9 + k and 23 + k (the views' bits 10 and 26 then hold
// window bits 1 and 3; shifted by one, bits 0 and 2 or 2 and 4); a multiply by 0x108421
// puts window bit b at bit 6b (no two products overlap, so no carries), the
// two windows' sum u has the per-bit counts 0..2 in 6-bit fields and u + (u <<
// 6) the diagonal count D_k = u_k + u_{k+1} in field k+1.  Views of D: D_0
// and D_1 at bit 6 of the low half (f16 value D 2^-18), D_2 and D_3 at bit 2
// of the high half (D 2^-22), against the coefficients 2^18 (2 J2) and 2^22
// (2 J2): the products are exactly D 2 J2, as in the byte form (4 D per byte
// against 2^22 (2 J2) / 4), so every fma rounds as before.
struct BitView {
    uint32_t S, xpe, xpo, xme, xmo, ype, ypo, yme, ymo, ze, zod, de, dod;
};

__device__ __forceinline__ uint32_t spread6(uint32_t w5) { return (w5 * 0x108421u) & 0x01041041u; }

template <int HW, int HX>
__device__ __forceinline__ BitView bit_view(const uint32_t *lds32, int c, int NU, const BitNb &n, GroupPos gp)
{
    const int pos0 = gp.pos, pmin = HX ? gp.pos : gp.pos - 1;
    const int p1 = 2 * (c ^ 1) * NU + gp.h, p2 = 2 * (c ^ 2) * NU + gp.h, p3 = 2 * (c ^ 3) * NU + gp.h;
    BitView v;
    v.S = nib(lds32, 2 * (c * NU + n.u) + gp.h, pos0);
    // x window bit k at bits 9 + k and 23 + k: the f16 views take bits 10
    // and 26 (sites 0 and 2 of an even view, 1 and 3 of an odd one), and the
    // +x view of sites 0, 2 is the -x view of sites 1, 3 (window bits 1, 3)
    uint32_t tx = __builtin_amdgcn_ubfe(lds32[p1 + 2 * n.u], (uint32_t)pmin, 5u) * 0x00800200u;
    // (hidden from the optimiser, which would otherwise fold tx << 1 into a
    // second multiply, by a constant too wide for the 24-bit multiplier)
    asm volatile("" : "+v"(tx));
    v.xmo = tx & 0x04000400u;
    v.xpe = v.xmo;
    v.xme = (tx << 1) & 0x04000400u;
    v.xpo = (tx >> 1) & 0x04000400u;
    const uint32_t yp = spread4(nib(lds32, p2 + 2 * n.uyp, pos0)), ym = spread4(nib(lds32, p2 + 2 * n.uym, pos0));
    v.ype = f16_even(yp);
    v.ypo = f16_odd(yp);
    v.yme = f16_even(ym);
    v.ymo = f16_odd(ym);
    const uint32_t Z = spread4(nib(lds32, p2 + 2 * n.uzp, pos0)) + spread4(nib(lds32, p2 + 2 * n.uzm, pos0));
    v.ze = lo_even(Z);
    v.zod = lo_odd(Z);
    const uint32_t u = spread6(__builtin_amdgcn_ubfe(lds32[p3 + 2 * n.uyp], (uint32_t)pmin, 5u)) +
                       spread6(__builtin_amdgcn_ubfe(lds32[p3 + 2 * n.uym], (uint32_t)pmin, 5u));
    const uint32_t Dw = u + (u << 6);
    v.de = Dw & 0x001C01C0u;
    v.dod = (Dw >> 6) & 0x001C01C0u;
    return v;
}

// The same operands from the spread neighbours (FP64 shadow path: D in the
// byte form, coefficient 2^22 (2 J2) for every site).
__device__ __forceinline__ BitView view_of_sites(const BitSites &s)
{
    const uint32_t D = (s.d1p + s.d1m) + (s.d2p + s.d2m);     // bytewise 4 * (0..4)
    const uint32_t Z = s.zp + s.zm;                           // bytewise 4 * (0..2)
    return BitView{s.S,           f16_even(s.xp), f16_odd(s.xp), f16_even(s.xm), f16_odd(s.xm),
                   f16_even(s.yp), f16_odd(s.yp), f16_even(s.ym), f16_odd(s.ym), lo_even(Z),
                   lo_odd(Z),      lo_even(D),    lo_odd(D)};
}

__device__ __forceinline__ uint32_t nib_bit(uint32_t n, int k) { return (n >> k) & 1u; }
__device__ __forceinline__ uint32_t sp_bit(uint32_t w, int k) { return (w >> (8 * k + 2)) & 1u; }

// A block's share of its lattice: global layers z0 .. z0+nz-1, NU units per
// plane in LDS (HW*H for the whole lattice, HW*(nz+2) for a slab).
struct Share {
    int z0, nz, NU;
};

// One 4-site group of a colour pass of the bit layout: thread (yh, j) owns
// group j of the half-row (zg, yh); its flips go to the unit word with an LDS
// atomic xor (the G threads of a unit share the word; nobody reads plane c in
// pass c).  Its random words are the ODD (= zg & 1) halves of the words `hi`
// of Philox call m = (zg / 2) TPL + tid (low halves from the call with LO_CTR
// when needed).  Philox counters and disorder planes are indexed by the global
// layer zg, so a slab updates its sites exactly as the whole lattice would.
struct KPlanes {
    float4 kc, kxp, kxm, kyp, kym;
};

__device__ __forceinline__ KPlanes load_kplanes(const SweepCtx &cx, uint32_t ob)
{
    return KPlanes{ld16(cx.kc4, ob), ld16(cx.kxp4, ob), ld16(cx.kxm4, ob), ld16(cx.kyp4, ob), ld16(cx.kym4, ob)};
}

template <int L, int HX, bool SHADOW, bool SLAB, int ODD>
__device__ __forceinline__ void bit_group(uint32_t *lds32, const SweepCtx &cx, const ConfConst &cc, int c, int H,
                                          int NU, int zg, int zl, int yh, int j, uint32_t m, const U4 &hi,
                                          const KPlanes &kp, unsigned long long &sh_prop, unsigned long long &sh_mis)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x, m0 = 4 * j;
    const int NUg = HW * H, nq = NUg * G;
    const BitNb n = bit_nb<HW, SLAB>(c, zg, zl, yh, H);
    const GroupPos gp = group_pos<L>(j);
    BitSites s{};                                             // spread neighbours: FP64 shadow only
    BitView v;
    if constexpr (SHADOW) {
        s = bit_sites<HW, HX>(lds32, c, NU, n, gp);
        v = view_of_sites(s);
    } else {
        v = bit_view<HW, HX>(lds32, c, NU, n, gp);
    }
    // D coefficients per site (the fast path's D_0, D_1 sit 2^4 higher)
    const float dlo = SHADOW ? cx.j2d : cx.j2d * 0x1p-4f;
    const float dk[4] = {dlo, dlo, cx.j2d, cx.j2d};
    const int own = c * nq + zg * TPL + tid;                  // the group's dword in the colour-plane order
    const float4 kc = kp.kc, kxp = kp.kxp, kxm = kp.kxm, kyp = kp.kyp, kym = kp.kym;
    const float kca[4] = {kc.x, kc.y, kc.z, kc.w};
    const float kxpa[4] = {kxp.x, kxp.y, kxp.z, kxp.w};
    const float kxma[4] = {kxm.x, kxm.y, kxm.z, kxm.w};
    const float kypa[4] = {kyp.x, kyp.y, kyp.z, kyp.w};
    const float kyma[4] = {kym.x, kym.y, kym.z, kym.w};
    const uint32_t ha[4] = {hi.x, hi.y, hi.z, hi.w};
    uint32_t ra[4] = {0u, 0u, 0u, 0u};
    if constexpr (SHADOW) {                                   // both rules on the whole words
        const U4 lo = philox4x32_10_u(m, cx.ctr1 | LO_CTR, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
        ra[0] = site_word(hi.x, lo.x, ODD); ra[1] = site_word(hi.y, lo.y, ODD);
        ra[2] = site_word(hi.z, lo.z, ODD); ra[3] = site_word(hi.w, lo.w, ODD);
    }
    uint32_t f = 0, open = 0;
    float ys[4], h[4];
    // the four sites' fields term by term (each site's sum in the same order):
    // independent neighbours in the issue stream, no dependent fma pairs
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = __builtin_fmaf(f16_site(v.xpe, v.xpo, k), kxpa[k], kca[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = __builtin_fmaf(f16_site(v.xme, v.xmo, k), kxma[k], h[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = __builtin_fmaf(f16_site(v.ype, v.ypo, k), kypa[k], h[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = __builtin_fmaf(f16_site(v.yme, v.ymo, k), kyma[k], h[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = __builtin_fmaf(f16_site(v.de, v.dod, k), dk[k], h[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = __builtin_fmaf(f16_site(v.ze, v.zod, k), cx.j3d, h[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ys[k] = y_f32_word(h[k], v.S, k, cx.b2);
        if constexpr (!SHADOW) {
            const Filt fl = decide16_filter(ys[k], v16_float(ha[k], ODD));
            f |= (fl.acc ? 1u : 0u) << k;
            open |= fl.open ? 1u << k : 0u;
            continue;
        }
        const bool acc = decide_f32_open(ys[k], ra[k]);
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
            // (global units: the y neighbours sit in the same layer)
            const int ug = zg * HW + yh;
            const int64_t q1 = (int64_t)((c ^ 1) * NUg + ug) * HW, q2 = (int64_t)(c ^ 2) * NUg * HW;
            const int64_t uyp = ug + (n.uyp - n.u), uym = ug + (n.uym - n.u);
            const double pv[5] = {cx.p64[4 * (int64_t)own + k], cx.p64[q2 + uyp * HW + m],
                                  cx.p64[q2 + uym * HW + m], cx.p64[q1 + mo_p], cx.p64[q1 + mo_m]};
            const bool a64 = decide_f64(field_f64(sv, pv, cc), up, cx.beta, ra[k]);
            sh_prop += 1;
            sh_mis += (a64 != acc);
        }
        f |= (acc ? 1u : 0u) << k;
    }
    if (!SHADOW && open) {                                // rare: the low halves and the exact rule
        const U4 lo = philox4x32_10_u(m, cx.ctr1 | LO_CTR, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
        const uint32_t la[4] = {lo.x, lo.y, lo.z, lo.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (((open >> k) & 1u) && decide_f32_open(ys[k], site_word(ha[k], la[k], ODD))) f |= 1u << k;
    }
    // flips into the group's dword; the first and last group of each dword
    // also flip their site's copy in the other dword of the unit
    constexpr int G0 = Gm::G0, N0 = Gm::N0, N1 = Gm::N1;
    uint32_t *uw = lds32 + 2 * (c * NU + n.u);
    atomicXor(uw + gp.h, f << gp.pos);
    // (one bit-field extract and one shift with operands fixed per thread --
    // a first group's bit 0 to its copy's place, a last group's bit 3 to bit 0,
    // a width of 0 for the others -- so one conditional LDS op and no branch tree)
    const bool first = j == 0 || j == G0, last = j == G0 - 1 || j == G - 1;
    const uint32_t cb = last ? 3u : 0u, cw = (first || last) ? 1u : 0u;
    const uint32_t csh = first ? (j == 0 ? N1 + 1 : N0 + 1) : 0u;
    const uint32_t cm = __builtin_amdgcn_ubfe(f, cb, cw) << csh;
    if (cm) atomicXor(uw + (1 - gp.h), cm);
}

// One colour pass over the block's layers, a layer pair per trip (one Philox
// call for the two groups; the block's first layer and its layer count are
// even).
template <int L, int HX, bool SHADOW, bool SLAB>
__device__ __forceinline__ void bit_pass(uint32_t *lds32, const SweepCtx &cx, const ConfConst &cc, int c, int H,
                                         const Share &sh, unsigned long long &sh_prop, unsigned long long &sh_mis)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x;
    if (tid >= TPL) return;
    const int yh = tid / G, j = tid - yh * G;
    const int nq = HW * H * G;
    for (int zl = 0; zl < sh.nz; zl += 2) {
        const int zg = sh.z0 + zl;
        const uint32_t m = (uint32_t)((zg >> 1) * TPL + tid);
        // the first group's planes are requested before its Philox call, the
        // second group's after the first group's fields (their latency hides
        // behind that work)
        uint32_t ob = (uint32_t)(c * nq + zg * TPL + tid) * 16u;   // formed here, not hoisted out of the loop
        asm volatile("" : "+v"(ob));
        const KPlanes k0 = load_kplanes(cx, ob);
        const U4 hi = philox4x32_10_u(m, cx.ctr1, cx.sweep_lo, cx.sweep_hi, cx.key0, cx.key1);
        bit_group<L, HX, SHADOW, SLAB, 0>(lds32, cx, cc, c, H, sh.NU, zg, zl, yh, j, m, hi, k0, sh_prop, sh_mis);
        const KPlanes k1 = load_kplanes(cx, ob + 16u * TPL);
        bit_group<L, HX, SHADOW, SLAB, 1>(lds32, cx, cc, c, H, sh.NU, zg + 1, zl + 1, yh, j, m, hi, k1, sh_prop,
                                          sh_mis);
    }
}

// Per-thread partial observables of the bit layout (same sums as ObsFast in
// kernels.hip, one 4-site group per step; +z bonds, so a slab reads its upper
// halo), for colour C.  The bond counts are popcounts of nibble xors (the x
// and diagonal neighbours from 5-site windows as in the sweep); the field
// term's e_k = a(+y) + a(-y) - a(+x) - a(-x) + 2 is formed in bytes from the
// four anti-aligned nibbles.  Layers go in pairs: the first of a pair has y
// parity C >> 1 (a block's first layer is even), the second the other one.
__device__ __forceinline__ uint32_t spread_b0(uint32_t n) { return (n * 0x00204081u) & 0x01010101u; }

template <int L, int C, bool SLAB>
__device__ __forceinline__ void bit_observe(const uint32_t *lds32, int H, const Share &sh, const QPlanes &qp,
                                            ObsFast &a)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL, HX = C & 1, P0 = C >> 1;
    const int tid = threadIdx.x;
    if (tid >= TPL) return;
    const int yh = tid / G, j = tid - yh * G;
    const int nq = HW * H * G, NU = sh.NU;
    const GroupPos gp = group_pos<L>(j);
    const uint32_t pos0 = (uint32_t)gp.pos, pmin = HX ? pos0 : pos0 - 1u;
    const int q0 = 2 * C * NU + gp.h, q1 = 2 * (C ^ 1) * NU + gp.h, q2 = 2 * (C ^ 2) * NU + gp.h,
              q3 = 2 * (C ^ 3) * NU + gp.h;
    uint32_t ps0 = 0, ps1 = 0, px = 0, py = 0, pz = 0, pd = 0;
    const auto layer = [&](int zl, int par, uint32_t &ps) {
        const int zg = sh.z0 + zl;
        const int u = (SLAB ? zl + 1 : zl) * HW + yh;
        const int yp = (P0 ^ par) ? (yh == HW - 1 ? u - (HW - 1) : u + 1) : u;
        const int ym = (P0 ^ par) ? u : (yh == 0 ? u + (HW - 1) : u - 1);
        const int zp = SLAB ? u + HW : (zl == H - 1 ? u - (H - 1) * HW : u + HW);
        const uint32_t S = __builtin_amdgcn_ubfe(lds32[q0 + 2 * u], pos0, 4u);
        const uint32_t wx = __builtin_amdgcn_ubfe(lds32[q1 + 2 * u], pmin, 5u);
        const uint32_t wd = __builtin_amdgcn_ubfe(lds32[q3 + 2 * yp], pmin, 5u);
        const uint32_t A = S ^ __builtin_amdgcn_ubfe(lds32[q2 + 2 * yp], pos0, 4u);   // +y anti-aligned
        const uint32_t B = S ^ __builtin_amdgcn_ubfe(lds32[q2 + 2 * ym], pos0, 4u);   // -y
        const uint32_t X = S ^ (wx >> 1), Y = S ^ (wx & 15u);                         // +x, -x
        ps += __builtin_popcount(S);
        px += __builtin_popcount(X);
        py += __builtin_popcount(A);
        pz += __builtin_popcount(S ^ __builtin_amdgcn_ubfe(lds32[q2 + 2 * zp], pos0, 4u));
        pd += __builtin_popcount(S ^ (wd >> 1)) + __builtin_popcount(S ^ (wd & 15u));
        const uint32_t e = (spread_b0(A) + spread_b0(B) + 0x02020202u) - (spread_b0(X) + spread_b0(Y));
        qsum_word(a.q, qp, (uint32_t)(C * nq + zg * TPL + tid) * 4u, e);
    };
    for (int zl = 0; zl < sh.nz; zl += 2) {
        layer(zl, 0, ps0);
        layer(zl + 1, 1, ps1);
    }
    const int32_t half = sh.nz / 2;
    const int32_t m0 = 2 * (int32_t)ps0 - 4 * half, m1 = 2 * (int32_t)ps1 - 4 * half;
    a.ps += ps0 + ps1; a.px += px; a.py += py; a.pz += pz; a.pd += pd; a.nw += (uint32_t)sh.nz;
    a.sxm += HX ? -(m0 + m1) : m0 + m1;
    a.sym += P0 ? m1 - m0 : m0 - m1;
}

// Accumulate_local_nematic_average (J1J2_6tv.f90:1009-1025) from the bit layout
// (x and y neighbours only: no halo involved).
template <int L, int HX, bool SLAB>
__device__ __forceinline__ void bit_locnem(const uint32_t *lds32, int c, int H, const Share &sh, int32_t *locspin,
                                           int32_t *locnem)
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, G = Gm::G, TPL = Gm::TPL;
    const int tid = threadIdx.x;
    if (tid >= TPL) return;
    const int yh = tid / G, j = tid - yh * G;
    const int nq = HW * H * G;
    int4 *ls4 = reinterpret_cast<int4 *>(locspin), *lv4 = reinterpret_cast<int4 *>(locnem);
    // layers in batches of K whose accumulators (in HBM) are all requested
    // before the first is added to: one memory round trip per batch
    constexpr int K = 4;
    for (int zb = 0; zb < sh.nz; zb += K) {
        int4 ls[K], lv[K];
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (zb + q < sh.nz) {
                const int dw = c * nq + (sh.z0 + zb + q) * TPL + tid;
                ls[q] = ls4[dw];
                lv[q] = lv4[dw];
            }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int zl = zb + q;
            if (zl >= sh.nz) break;
            const int zg = sh.z0 + zl;
            const BitNb n = bit_nb<HW, SLAB>(c, zg, zl, yh, H);
            const BitSites s = bit_sites<HW, HX>(lds32, c, sh.NU, n, group_pos<L>(j));
            const uint32_t S = spread4(s.S);
            const uint32_t ax = (S ^ s.xp) + (S ^ s.xm), ay = (S ^ s.yp) + (S ^ s.ym);
            int *lsp = reinterpret_cast<int *>(&ls[q]), *lnp = reinterpret_cast<int *>(&lv[q]);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                lsp[k] += (int)nib_bit(s.S, k) * 2 - 1;
                lnp[k] += 2 * ((int)((ay >> (8 * k + 2)) & 0xffu) - (int)((ax >> (8 * k + 2)) & 0xffu));
            }
            const int dw = c * nq + zg * TPL + tid;
            ls4[dw] = ls[q];
            lv4[dw] = lv[q];
        }
    }
}

// ---- slab exchange words
constexpr int XV_SHIFT = 50;                             // a unit's HW <= 48 sites below, the version above
constexpr unsigned long long XW_MASK = (1ull << XV_SHIFT) - 1;
constexpr uint32_t XV_MASK = (1u << (64 - XV_SHIFT)) - 1;

__device__ __forceinline__ unsigned long long xload(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void xstore(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait until the word at p carries version `ver` in its bits from
// `shift` up; after spin_max polls (or once another wait has run out) the
// error flag is set and the wait ends, so every wave always finishes.
// spin_max < 0 makes every wait give up at once (the host's test of the
// error path).
__device__ unsigned long long xwait(const unsigned long long *p, int shift, uint32_t ver, int32_t *err,
                                    int spin_max)
{
    unsigned long long w = xload(p);
    if (spin_max < 0) {
        atomicOr(err, 1);
        return w;
    }
    for (int it = 1; (uint32_t)(w >> shift) != ver; ++it) {
        if ((it & 255) == 0) {
            if (it >= spin_max) {
                atomicOr(err, 1);
                break;
            }
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        }
        __builtin_amdgcn_s_sleep(1);
        w = xload(p);
    }
    return w;
}

// Exchange word of (lattice, slab, plane, boundary w: 0 lower layer, 1 upper), unit yh.
template <int HW>
__device__ __forceinline__ unsigned long long *xword(const DeviceState &st, int lat, int s, int c, int w, int yh)
{
    return st.slab_x + ((((int64_t)lat * BIT_SLAB_MAX + s) * 4 + c) * 2 + w) * HW + yh;
}

// After pass c: the slab's boundary layers of plane c, version ver.
template <int L>
__device__ __forceinline__ void slab_publish(const DeviceState &st, const uint32_t *lds32, const Share &sh, int lat,
                                             int s, int c, uint32_t ver)
{
    constexpr int HW = BitGeo<L>::HW;
    const int tid = threadIdx.x;
    if (tid < 2 * HW) {
        const int w = tid >= HW, yh = tid - w * HW;
        const unsigned long long v =
            sites_from_unit<L>(reinterpret_cast<const uint2 *>(lds32)[c * sh.NU + (w ? sh.nz : 1) * HW + yh]);
        xstore(xword<HW>(st, lat, s, c, w, yh), v | ((unsigned long long)ver << XV_SHIFT));
    }
}

// The halo layers of plane c from the slabs below (its upper layer) and above
// (its lower layer), version ver; the caller synchronises before reading them.
template <int L>
__device__ __forceinline__ void slab_fetch(const DeviceState &st, uint32_t *lds32, const Share &sh, int lat, int s,
                                           int S, int c, uint32_t ver)
{
    constexpr int HW = BitGeo<L>::HW;
    const int tid = threadIdx.x;
    if (tid < 2 * HW) {
        const int w = tid >= HW, yh = tid - w * HW;            // w = 0: lower halo, 1: upper
        const int src = w ? (s + 1 == S ? 0 : s + 1) : (s == 0 ? S - 1 : s - 1);
        const unsigned long long x = xwait(xword<HW>(st, lat, src, c, 1 - w, yh), XV_SHIFT, ver, st.slab_err,
                                           st.slab_spin_max);
        reinterpret_cast<uint2 *>(lds32)[c * sh.NU + (w ? sh.nz + 1 : 0) * HW + yh] = unit_from_sites<L>(x & XW_MASK);
    }
}

template <int L, bool SHADOW, bool SLAB>
__device__ __forceinline__ void bits_body(const DeviceState &st, const LatDesc &d, int lat, int s, int S,
                                          uint32_t *lds32, int64_t (*red)[OBS_N])
{
    using Gm = BitGeo<L>;
    constexpr int HW = Gm::HW, NT = Gm::NT;
    const int H = d.H, tid = threadIdx.x;
    const int NUg = HW * H;
    Share sh;
    sh.nz = SLAB ? H / S : H;
    sh.z0 = SLAB ? s * sh.nz : 0;
    sh.NU = SLAB ? HW * (sh.nz + 2) : NUg;
    // the lattice in HBM is its unit words (global unit c*NUg + z*HW + yh);
    // a slab takes its layers and the two halo layers
    uint2 *gu = reinterpret_cast<uint2 *>(st.spins + d.spin_off);
    for (int U = tid; U < 4 * sh.NU; U += NT) {
        const int c = U / sh.NU, r = U - c * sh.NU, li = r / HW, yh = r - li * HW;
        int zg = li;
        if constexpr (SLAB) {
            zg = sh.z0 - 1 + li;
            zg = zg < 0 ? zg + H : (zg >= H ? zg - H : zg);
        }
        reinterpret_cast<uint2 *>(lds32)[U] = gu[(int64_t)c * NUg + (int64_t)zg * HW + yh];
    }
    if (tid == 0) lds32[8 * sh.NU] = 0;                         // the spare dword

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
    cx.cnt_off = 0;                                            // flip counts take the global-memory sweep
    const uint64_t S0 = clock_cycle(st) * (uint64_t)st.ptswapever;
    // exchange versions count the lattice's slab launches (word 0 of slab 0's
    // observable slots, advanced by slab 0 once every slab has published its
    // last words): pass c of sweep k is version 4k + c + 1 after vbase, so a
    // slot's stale version never equals the one awaited
    unsigned long long *ox = SLAB ? st.slab_obs + (int64_t)lat * BIT_SLAB_MAX * 16 : nullptr;
    if (SLAB && tid == 0) red[0][0] = (int64_t)xload(ox);
    unsigned long long sh_prop = 0, sh_mis = 0;
    __syncthreads();
    const uint32_t lc = SLAB ? (uint32_t)red[0][0] : 0u;
    const uint32_t vbase = lc * (uint32_t)(4 * st.ptswapever + 1);
    if (st.fuse_locnem && slot == st.ptmeasindex - 1 && st.locnem_pending[d.conf]) {
        for (int c = 0; c < 4; ++c) {
            if (c & 1) bit_locnem<L, 1, SLAB>(lds32, c, H, sh, st.locspin + d.site_off, st.locnem + d.site_off);
            else bit_locnem<L, 0, SLAB>(lds32, c, H, sh, st.locspin + d.site_off, st.locnem + d.site_off);
        }
        if (tid == 0 && s == 0) st.locsamples[d.conf] += 1;
        __syncthreads();
    }

    for (int k = 0; k < st.ptswapever; ++k) {
        const uint64_t sweep = S0 + (uint64_t)k;
        cx.sweep_lo = (uint32_t)sweep;
        cx.sweep_hi = (uint32_t)(sweep >> 32);
#pragma unroll 1
        for (int c = 0; c < 4; ++c) {
            cx.ctr1 = (uint32_t)c | ((uint32_t)d.rep << 2);
            if constexpr (SLAB) {
                // The halos of the next pass (kn, cn) are fetched during this one:
                // it reads plane cn^2 only in its halo layers, which this pass
                // neither reads (its z-neighbours are in plane c^2) nor writes,
                // and the barrier after it orders the writes before pass cn.
                // Plane cn^2 was last updated in pass cn^2 of sweep kn (cn >= 2)
                // or kn - 1 (cn < 2; at kn = 0 the halos loaded with the slab),
                // published by the neighbours before they could need this
                // pass's words, so the wait always ends.
                const int cn = (c + 1) & 3, kn = c == 3 ? k + 1 : k;
                if (kn < st.ptswapever && (cn >= 2 || kn > 0)) {
                    const uint32_t ver = vbase + 4u * (uint32_t)(cn >= 2 ? kn : kn - 1) + (uint32_t)(cn ^ 2) + 1u;
                    slab_fetch<L>(st, lds32, sh, lat, s, S, cn ^ 2, ver & XV_MASK);
                }
            }
            if (c & 1) bit_pass<L, 1, SHADOW, SLAB>(lds32, cx, cc, c, H, sh, sh_prop, sh_mis);
            else bit_pass<L, 0, SHADOW, SLAB>(lds32, cx, cc, c, H, sh, sh_prop, sh_mis);
            __syncthreads();
            if constexpr (SLAB) slab_publish<L>(st, lds32, sh, lat, s, c, (vbase + 4u * k + c + 1u) & XV_MASK);
        }
    }
    if constexpr (SLAB) {
        // the final planes 2 and 3 of the neighbours (planes 0 and 1 were
        // fetched after their last update); the observables read the upper halo
        const uint32_t kl = (uint32_t)st.ptswapever - 1u;
        slab_fetch<L>(st, lds32, sh, lat, s, S, 2, (vbase + 4u * kl + 3u) & XV_MASK);
        slab_fetch<L>(st, lds32, sh, lat, s, S, 3, (vbase + 4u * kl + 4u) & XV_MASK);
        __syncthreads();
    }

    ObsFast a = {};
    const QPlanes qp = qplanes(st.qb + d.site_off, st.nsites);
    bit_observe<L, 0, SLAB>(lds32, H, sh, qp, a);
    bit_observe<L, 1, SLAB>(lds32, H, sh, qp, a);
    bit_observe<L, 2, SLAB>(lds32, H, sh, qp, a);
    bit_observe<L, 3, SLAB>(lds32, H, sh, qp, a);
    const int64_t nw = (int64_t)a.nw;
    int64_t v[OBS_N];
    v[OBS_M] = 2 * (int64_t)a.ps - 4 * nw;
    v[OBS_SX] = a.sxm;
    v[OBS_SY] = a.sym;
    v[OBS_BX] = 4 * nw - 2 * (int64_t)a.px;
    v[OBS_BY] = 4 * nw - 2 * (int64_t)a.py;
    v[OBS_BD] = 8 * nw - 2 * (int64_t)a.pd;
    v[OBS_BZ] = 4 * nw - 2 * (int64_t)a.pz;
    v[OBS_EQ] = 2 * qsum_total(a.q);
    if (SHADOW) {
        sh_prop = wave_sum<unsigned long long>(sh_prop);
        sh_mis = wave_sum<unsigned long long>(sh_mis);
        if ((tid & 63) == 0 && sh_prop) {
            atomicAdd(&st.shadow[2 * d.conf + 0], sh_prop);
            atomicAdd(&st.shadow[2 * d.conf + 1], sh_mis);
        }
    }
    // unit words back to HBM (a slab stores its own layers)
    const int off = SLAB ? HW : 0;
    for (int U = tid; U < 4 * sh.nz * HW; U += NT) {
        const int c = U / (sh.nz * HW), r = U - c * sh.nz * HW;
        gu[(int64_t)c * NUg + (int64_t)sh.z0 * HW + r] = reinterpret_cast<const uint2 *>(lds32)[c * sh.NU + off + r];
    }
    block_reduce_obs<NT>(v, red);
    if (tid == 0) {
        if constexpr (SLAB) {
            // partial sums as 32-bit halves under the 32-bit version lc + 1;
            // slab 0 adds the others' and writes the lattice's observables
            const uint32_t ov = lc + 1u;
            if (s != 0) {
#pragma unroll
                for (int k = 0; k < OBS_N; ++k) {
                    const uint64_t u = (uint64_t)v[k];
                    xstore(ox + s * 16 + 2 * k, ((unsigned long long)ov << 32) | (uint32_t)u);
                    xstore(ox + s * 16 + 2 * k + 1, ((unsigned long long)ov << 32) | (uint32_t)(u >> 32));
                }
                return;
            }
            for (int q = 1; q < S; ++q)
#pragma unroll
                for (int k = 0; k < OBS_N; ++k) {
                    const uint32_t lo = (uint32_t)xwait(ox + q * 16 + 2 * k, 32, ov, st.slab_err, st.slab_spin_max);
                    const uint32_t hi =
                        (uint32_t)xwait(ox + q * 16 + 2 * k + 1, 32, ov, st.slab_err, st.slab_spin_max);
                    v[k] += (int64_t)(((uint64_t)hi << 32) | lo);
                }
            xstore(ox, (unsigned long long)(lc + 1u));   // every slab has read lc and published
        }
        v[OBS_EQ] -= 4 * cc.qsum;
        int64_t *o = st.obs + (int64_t)lat * OBS_N;
#pragma unroll
        for (int k = 0; k < OBS_N; ++k) o[k] = v[k];
    }
}

// Whole lattice.  At most 168 VGPRs (three waves per SIMD): one lattice of
// L >= 64 fills most of a CU's LDS, so a CU holds one or two blocks of at most
// 9 waves.
template <int L, bool SHADOW>
__global__ __launch_bounds__(BitGeo<L>::NT) __attribute__((amdgpu_waves_per_eu(SHADOW ? 1 : 3, 8)))
void k_sweep_bits(DeviceState st, int32_t ord0, int32_t nord, int32_t lds_bytes, int32_t slabs)
{
    // the lattice's unit words from LDS address 0, the reduction scratch after them
    extern __shared__ uint32_t lds32[];
    const int lat = st.order[ord0 + blockIdx.x];
    if (lat < 0) return;
    const LatDesc d = st.lats[lat];
    int64_t (*red)[OBS_N] = reinterpret_cast<int64_t (*)[OBS_N]>(reinterpret_cast<char *>(lds32) + lds_bytes);
    bits_body<L, SHADOW, false>(st, d, lat, 0, 1, lds32, red);
}

// Slabs.  At most 72 VGPRs (seven waves per SIMD), so that three slab blocks
// of up to 9 waves share a CU.
template <int L, bool SHADOW>
__global__ __launch_bounds__(BitGeo<L>::NT) __attribute__((amdgpu_waves_per_eu(SHADOW ? 1 : 7, 8)))
void k_sweep_bits_slab(DeviceState st, int32_t ord0, int32_t nord, int32_t lds_bytes, int32_t slabs)
{
    extern __shared__ uint32_t lds32[];
    const int e = st.order[ord0 + blockIdx.x];
    if (e < 0) return;
    const int lat = e / BIT_SLAB_MAX, s = e - lat * BIT_SLAB_MAX;
    const LatDesc d = st.lats[lat];
    int64_t (*red)[OBS_N] = reinterpret_cast<int64_t (*)[OBS_N]>(reinterpret_cast<char *>(lds32) + lds_bytes);
    bits_body<L, SHADOW, true>(st, d, lat, s, slabs, lds32, red);
}

using BitKernel = void (*)(DeviceState, int32_t, int32_t, int32_t, int32_t);

template <bool SHADOW, bool SLAB>
BitKernel bits_kernel(int32_t L)
{
    switch (L) {
        case 56: return SLAB ? k_sweep_bits_slab<56, SHADOW> : k_sweep_bits<56, SHADOW>;
        case 64: return SLAB ? k_sweep_bits_slab<64, SHADOW> : k_sweep_bits<64, SHADOW>;
        case 72: return SLAB ? k_sweep_bits_slab<72, SHADOW> : k_sweep_bits<72, SHADOW>;
        case 80: return SLAB ? k_sweep_bits_slab<80, SHADOW> : k_sweep_bits<80, SHADOW>;
        case 88: return SLAB ? k_sweep_bits_slab<88, SHADOW> : k_sweep_bits<88, SHADOW>;
        case 96: return SLAB ? k_sweep_bits_slab<96, SHADOW> : k_sweep_bits<96, SHADOW>;
        default: return nullptr;
    }
}

BitKernel pick_bits(int32_t L, bool shadow, bool slab)
{
    return shadow ? (slab ? bits_kernel<true, true>(L) : bits_kernel<true, false>(L))
                  : (slab ? bits_kernel<false, true>(L) : bits_kernel<false, false>(L));
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

// LDS of a block: its unit words and the spare dword, rounded to 8 bytes
size_t bits_block_lds(int32_t L, int32_t H, int32_t slabs)
{
    const size_t words = (size_t)4 * (L / 2) * (slabs > 1 ? H / slabs + 2 : H);
    return (8 * words + 4 + 7) / 8 * 8;
}

}  // namespace

bool bits_supported(int32_t L) { return L > 48 && L <= 96 && (L % 8) == 0; }

size_t bits_lds_need(int32_t L, int32_t H) { return bits_block_lds(L, H, 1) + BITS_RED_LDS; }

bool bits_slab_valid(int32_t L, int32_t H, int32_t slabs)
{
    return bits_supported(L) && slabs > 1 && slabs <= BIT_SLAB_MAX && H % (2 * slabs) == 0;
}

bool bits_slab_fits(int32_t L, int32_t H, int32_t slabs, int32_t nblocks, bool shadow)
{
    if (!bits_slab_valid(L, H, slabs)) return false;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    const size_t lds = bits_block_lds(L, H, slabs) + BITS_RED_LDS;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)pick_bits(L, shadow, true),
                                                     bits_threads(L), lds) != hipSuccess)
        return false;
    return (int64_t)per * ncu >= nblocks;
}

size_t bits_slab_x_words(int32_t nlat) { return (size_t)nlat * BIT_SLAB_MAX * 4 * 2 * 48; }

hipError_t setup_bits(size_t max_lds)
{
    for (int L : {56, 64, 72, 80, 88, 96})
        for (bool sh : {false, true})
            for (bool sl : {false, true}) {
                hipError_t e = hipFuncSetAttribute((const void *)pick_bits(L, sh, sl),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
                if (e != hipSuccess) return e;
            }
    return hipSuccess;
}

hipError_t launch_sweep_bits(const DeviceState &st, int32_t ord0, int32_t nord, int32_t L, int32_t H, int32_t slabs,
                             bool shadow, hipStream_t s)
{
    if (nord <= 0) return hipSuccess;
    const bool slab = slabs > 1;
    if (slab && !bits_slab_valid(L, H, slabs)) return hipErrorInvalidValue;
    const BitKernel k = pick_bits(L, shadow, slab);
    if (!k) return hipErrorInvalidValue;
    const int32_t lds = (int32_t)bits_block_lds(L, H, slab ? slabs : 1);
    hipLaunchKernelGGL(k, dim3(nord), dim3(bits_threads(L)), (size_t)lds + BITS_RED_LDS, s, st, ord0, nord, lds,
                       slab ? slabs : 1);
    return hipGetLastError();
}

}  // namespace j1j2
