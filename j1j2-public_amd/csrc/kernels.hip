// kernels.hip -- HIP kernels of the J1-J2 Metropolis / parallel-tempering
// engine for MI355X (gfx950, wave64): the global-memory sweeps, observables,
// PT step and local-nematic sums, and the launchers of the LDS-resident
// sweep (its device code in resident.h, instantiated per block size in
// kernels_res<NT>.hip).
//
// Update rule (the oracle restates it on the CPU):
//   * a sweep is four colour passes, colour(x,y,z) = (x&1) | ((y+z)&1)<<1;
//     no stencil neighbour of a site (+-x, +-y, +-z, +-x+-y) shares its
//     colour, so each pass updates N/4 independent sites (needs L, H even);
//   * the c-th colour's sites of layer z are numbered q = (y/2)*L/2 + x/2;
//     site q of layer z of replica r of config iconf in sweep S takes the
//     32-bit word v 2^16 + l: v is one 16-bit half of word q&3 of
//     Philox4x32-10(counter = (m, c | r<<2, S lo, S hi), key = (seed, iconf))
//     with m = (z/2) GL + q/4 (GL groups of four sites per layer; even layers
//     the high halves, odd layers the low halves), l the same half of the
//     call with bit 30 set in the second counter word (device_common.h);
//   * FP32 form: h = C_i + sum_j 2K_j b_j + 2 J2 D + 2 J3 Z (b in {0,1});
//     FP64 form: Compute_local_field (J1J2_6tv.f90:2498-2541) in its order;
//     Metropolis acceptance of Metro_sweep (J1J2_6tv.f90:2572-2604).
#include <hip/hip_runtime.h>

#include "resident.h"

namespace j1j2 {

namespace {

constexpr int GLB_THREADS = 256;

// ---------------------------------------------------------------------------
// Global-memory sweeps (any even L): one colour pass per launch, one thread
// per Philox call (two 4-site groups, one in each layer of a pair),
// neighbours read from L2/HBM.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void colour_site(int L, int c, int q, int &x, int &y, int &z)
{
    const int half = L >> 1;
    const int xh = q % half, rr = q / half;
    z = rr / half;
    const int yh = rr - z * half;
    x = 2 * xh + (c & 1);
    y = 2 * yh + (((c >> 1) ^ z) & 1);
}

struct Nbr {               // internal indices of a site and its ten neighbours
    int64_t i, m[10];      // m1..m10 = +y,-y,+x,-x,+x+y,-x+y,+x-y,-x-y,+z,-z
};

__device__ __forceinline__ Nbr neighbours(int L, int H, int x, int y, int z)
{
    const int xp = (x == L - 1) ? 0 : x + 1, xm = (x == 0) ? L - 1 : x - 1;
    const int yp = (y == L - 1) ? 0 : y + 1, ym = (y == 0) ? L - 1 : y - 1;
    const int zp = (z == H - 1) ? 0 : z + 1, zm = (z == 0) ? H - 1 : z - 1;
    Nbr n;
    n.i = internal_index(L, H, x, y, z);
    n.m[0] = internal_index(L, H, x, yp, z);  n.m[1] = internal_index(L, H, x, ym, z);
    n.m[2] = internal_index(L, H, xp, y, z);  n.m[3] = internal_index(L, H, xm, y, z);
    n.m[4] = internal_index(L, H, xp, yp, z); n.m[5] = internal_index(L, H, xm, yp, z);
    n.m[6] = internal_index(L, H, xp, ym, z); n.m[7] = internal_index(L, H, xm, ym, z);
    n.m[8] = internal_index(L, H, x, y, zp);  n.m[9] = internal_index(L, H, x, y, zm);
    return n;
}

// Philox calls of one colour plane: a call per 4-site group of a layer pair
// (GL = ceil((L/2)^2 / 4) groups per layer; the last one of a layer is short
// when (L/2)^2 % 4 != 0).
__host__ __device__ inline int global_calls(int L, int H)
{
    const int half = L >> 1;
    return (H >> 1) * ((half * half + 3) / 4);
}

template <int PREC, bool SHADOW>   // PREC 1 = FP32 rule, 0 = FP64 plain rule
__global__ __launch_bounds__(GLB_THREADS) void k_sweep_global(DeviceState st, int32_t lat_begin,
                                                              int32_t L, int32_t H, int32_t k, int32_t c)
{
    const int lat = lat_begin + blockIdx.y;
    const LatDesc d = st.lats[lat];
    const int half = L >> 1, lay = half * half, GL = (lay + 3) / 4;
    const int m = blockIdx.x * GLB_THREADS + threadIdx.x;
    if (m >= (H >> 1) * GL) return;
    const int zp = m / GL, gl = m - zp * GL;
    const ConfDesc cd = st.confs[d.conf];
    const ConfConst cc = st.cconst[d.conf];
    const int slot = st.temp_of_rep[d.slot_base + d.rep];
    const float b2 = st.b2[d.slot_base + slot];
    const double beta = st.beta[d.slot_base + slot];
    const uint64_t sweep = clock_cycle(st) * (uint64_t)st.ptswapever + (uint64_t)k;
    const uint32_t ctr1 = (uint32_t)c | ((uint32_t)d.rep << 2);
    const U4 rn = philox4x32_10((uint32_t)m, ctr1, (uint32_t)sweep, (uint32_t)(sweep >> 32), st.seed,
                                (uint32_t)cd.iconf);
    const uint32_t ha[4] = {rn.x, rn.y, rn.z, rn.w};
    uint32_t la[4] = {0u, 0u, 0u, 0u};
    bool have_lo = false;
    const auto draw_lo = [&]() {      // the low halves, drawn once and only when needed
        if (have_lo) return;
        const U4 lo = philox4x32_10((uint32_t)m, ctr1 | LO_CTR, (uint32_t)sweep, (uint32_t)(sweep >> 32), st.seed,
                                    (uint32_t)cd.iconf);
        la[0] = lo.x; la[1] = lo.y; la[2] = lo.z; la[3] = lo.w;
        have_lo = true;
    };
    if (SHADOW) draw_lo();
    uint8_t *sp = st.spins + d.spin_off;
    const int64_t so = d.site_off;
    const float *kc = st.kd + 0 * st.nsites + so, *kxp = st.kd + 1 * st.nsites + so,
                *kxm = st.kd + 2 * st.nsites + so, *kyp = st.kd + 3 * st.nsites + so,
                *kym = st.kd + 4 * st.nsites + so;
    const double *p = st.p64 + so;
    const int N = L * L * H;
    unsigned long long sh_prop = 0, sh_mis = 0;
    for (int odd = 0; odd < 2; ++odd)
        for (int j = 0; j < 4; ++j) {
            const int qz = 4 * gl + j;
            if (qz >= lay) break;
            int x, y, z;
            colour_site(L, c, (2 * zp + odd) * lay + qz, x, y, z);
            const Nbr nb = neighbours(L, H, x, y, z);
            uint32_t b[10];
#pragma unroll
            for (int t = 0; t < 10; ++t) b[t] = sp[nb.m[t]];
            const uint32_t up = sp[nb.i];
            bool acc, other = false;
            if (PREC == 1 || SHADOW) {
                // planes hold 2^14 * 2K: (2^-14 b) (2^14 2K) is exactly b 2K
                float h = kc[nb.i];
                h = __builtin_fmaf((float)b[2] * 0x1p-14f, kxp[nb.i], h);
                h = __builtin_fmaf((float)b[3] * 0x1p-14f, kxm[nb.i], h);
                h = __builtin_fmaf((float)b[0] * 0x1p-14f, kyp[nb.i], h);
                h = __builtin_fmaf((float)b[1] * 0x1p-14f, kym[nb.i], h);
                h = __builtin_fmaf((float)(b[4] + b[5] + b[6] + b[7]), cc.j2x2, h);
                h = __builtin_fmaf((float)(b[8] + b[9]), cc.j3x2, h);
                const float yv = y_f32_word(h, up, 0, b2);
                bool a32;
                if (SHADOW) {
                    a32 = decide_f32_open(yv, site_word(ha[j], la[j], odd));
                } else {
                    const Filt f = decide16_filter(yv, v16_float(ha[j], odd));
                    a32 = f.acc;
                    if (f.open) {
                        draw_lo();
                        a32 = decide_f32_open(yv, site_word(ha[j], la[j], odd));
                    }
                }
                if (PREC == 1) acc = a32; else other = a32;
            }
            if (PREC == 0 || SHADOW) {
                double sv[10];
#pragma unroll
                for (int t = 0; t < 10; ++t) sv[t] = pm1(b[t]);
                const double pv[5] = {p[nb.i], p[nb.m[0]], p[nb.m[1]], p[nb.m[2]], p[nb.m[3]]};
                const double field = field_f64(sv, pv, cc);
                bool a64;
                if (SHADOW) {
                    a64 = decide_f64(field, up != 0, beta, site_word(ha[j], la[j], odd));
                } else {
                    const Dec64 dd = decide16_f64(field, up != 0, beta, half16(ha[j], odd));
                    a64 = dd.acc;
                    if (dd.open) {
                        draw_lo();
                        a64 = decide_f64(field, up != 0, beta, site_word(ha[j], la[j], odd));
                    }
                }
                if (PREC == 0) acc = a64; else other = a64;
            }
            if (SHADOW) { sh_prop += 1; sh_mis += (other != acc); }
            if (acc) {
                sp[nb.i] = (uint8_t)(up ^ 1u);
                if (st.flips) st.flips[cd.flip_off + (int64_t)slot * N + nb.i] += 1;
            }
        }
    if (SHADOW) {
        sh_prop = wave_sum<unsigned long long>(sh_prop);
        sh_mis = wave_sum<unsigned long long>(sh_mis);
        if ((threadIdx.x & 63) == 0 && sh_prop) {
            atomicAdd(&st.shadow[2 * d.conf + 0], sh_prop);
            atomicAdd(&st.shadow[2 * d.conf + 1], sh_mis);
        }
    }
}

// ---------------------------------------------------------------------------
// Integer observables from global memory (paths without the fused reduction).
// ---------------------------------------------------------------------------
constexpr int OBS_THREADS = 256;

__global__ __launch_bounds__(OBS_THREADS) void k_observables(DeviceState st, int32_t lat_begin)
{
    __shared__ int64_t red[OBS_THREADS / 64][OBS_N];
    const int lat = lat_begin + blockIdx.x;
    const LatDesc d = st.lats[lat];
    const int L = d.L, H = d.H, N = d.N;
    const uint8_t *sp = st.spins + d.spin_off;
    const uint8_t *qb = st.qb + d.site_off;
    int64_t v[OBS_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < N; i += OBS_THREADS) {
        int x, y, z;
        site_of_internal(L, H, i, x, y, z);
        const Nbr nb = neighbours(L, H, x, y, z);
        const int si = spin_at(sp, d.packed, L, i) ? 1 : -1;
        int s[10];
#pragma unroll
        for (int t = 0; t < 10; ++t) s[t] = spin_at(sp, d.packed, L, nb.m[t]) ? 1 : -1;
        v[OBS_M] += si;
        v[OBS_SX] += (x & 1) ? -si : si;
        v[OBS_SY] += (y & 1) ? -si : si;
        v[OBS_BX] += si * s[2];
        v[OBS_BY] += si * s[0];
        v[OBS_BD] += si * s[4] + si * s[5];
        v[OBS_BZ] += si * s[8];
        int64_t q = -QOFF;
#pragma unroll
        for (int b = 0; b < 5; ++b) q += (int64_t)qb[b * st.nsites + i] << (8 * b);
        v[OBS_EQ] += q * (int64_t)(si * (s[2] + s[3] - s[0] - s[1]));
    }
    block_reduce_obs<OBS_THREADS>(v, red);
    if (threadIdx.x == 0) {
        int64_t *o = st.obs + (int64_t)lat * OBS_N;
        for (int k = 0; k < OBS_N; ++k) o[k] = v[k];
    }
}

// ---------------------------------------------------------------------------
// Per-config PT step (one wave per config): energies from the integer
// observables, one pass of adjacent swaps (Pt_attempt_swaps, J1J2_6tv.f90:
// 1303-1338; pairs are disjoint so they are decided in parallel), then the
// per-slot accumulation of Pt_measurement_loop (J1J2_6tv.f90:1480-1524).
// ---------------------------------------------------------------------------
constexpr int PT_THREADS = 64;
constexpr int MAX_NREP = 256;

__global__ __launch_bounds__(PT_THREADS) void k_pt_step(DeviceState st, int32_t mode)
{
    // mode bit 0: accumulate measurements; bit 1: energies and values only
    const bool measure = (mode & 1) != 0, swaps = (mode & 2) == 0;
    __shared__ int32_t rat[MAX_NREP], tor[MAX_NREP];
    __shared__ double et[MAX_NREP], vv[MAX_NREP][5];
    const int conf = blockIdx.x, tid = threadIdx.x;
    // the config's offsets: computed when every config has the same shape
    // (nothing on the critical path waits for its descriptor), else loaded
    int nrep, sb, lat0, pair_base;
    int64_t hist_off;
    double N;
    if (st.uni_nrep > 0) {
        nrep = st.uni_nrep;
        sb = conf * nrep;
        lat0 = sb;
        pair_base = conf * (nrep - 1);
        hist_off = (int64_t)sb * st.bins;
        N = (double)st.uni_n;
    } else {
        const ConfDesc cd = st.confs[conf];
        nrep = cd.nrep;
        sb = cd.slot_base;
        lat0 = cd.lat0;
        pair_base = cd.pair_base;
        hist_off = cd.hist_off;
        N = (double)cd.N;
    }
    const ConfConst cc = st.cconst[conf];
    const uint64_t ticks = st.clk->ticks;
    const uint64_t cycle = ticks / (uint64_t)st.nconf;
    const int phase = (((ticks - st.clk->phase_start) / (uint64_t)st.nconf) & 1ull) ? 2 : 1;
    const int64_t slen = st.local ? st.series_len[conf] : 0;
    // The kernel is a chain of dependent memory round trips; everything that
    // does not depend on this cycle's energies is requested up front: the
    // betas of this thread's first swap pair and its first slot's running
    // sums, and counters are bumped with atomics that return nothing.
    const int start = (phase == 1) ? 0 : 1;
    const int s0 = start + 2 * tid;
    const bool pair0 = !st.local && swaps && s0 < nrep - 1;
    const double b_lo = pair0 ? st.beta[sb + s0] : 0.0, b_hi = pair0 ? st.beta[sb + s0 + 1] : 0.0;
    const uint32_t key1 = pair0 ? (uint32_t)st.confs[conf].iconf : 0u;   // the swap draws' key word
    const bool slot0 = measure && !st.local && tid < nrep;
    double S0[NSUMS], best0 = 0.0, hmin0 = 0.0, hmax0 = 0.0, hbin0 = 0.0;
    if (slot0) {
        const double *S = st.sums + (int64_t)(sb + tid) * NSUMS;
#pragma unroll
        for (int q = 0; q < NSUMS; ++q) S0[q] = S[q];
        best0 = st.best[sb + tid];
        hmin0 = st.hmin[sb + tid];
        hmax0 = st.hmax[sb + tid];
        hbin0 = st.hbin[sb + tid];
    }
    for (int r = tid; r < nrep; r += PT_THREADS) {
        const int64_t *o = st.obs + (int64_t)(lat0 + r) * OBS_N;
        const double ej = ((cc.J1h * (double)o[OBS_BX]) + (cc.J1v * (double)o[OBS_BY])) +
                          ((cc.J2 * (double)o[OBS_BD]) + (cc.J3 * (double)o[OBS_BZ]));
        const double E = -ej - (double)o[OBS_EQ] * 0x1p-32;
        et[r] = E;
        vv[r][0] = (double)o[OBS_M] / N;
        vv[r][1] = E / N;
        vv[r][2] = (double)o[OBS_SX] / N;
        vv[r][3] = (double)o[OBS_SY] / N;
        vv[r][4] = (double)(2 * (o[OBS_BX] - o[OBS_BY])) / N;
        st.etot[sb + r] = E;
        for (int q = 0; q < 5; ++q) st.vals[(int64_t)(sb + r) * 5 + q] = vv[r][q];
        rat[r] = st.rep_at_temp[sb + r];
        tor[r] = st.temp_of_rep[sb + r];
    }
    __syncthreads();
    if (!st.local && swaps) {
        for (int s = s0; s < nrep - 1; s += 2 * PT_THREADS) {
            const int rlo = rat[s], rhi = rat[s + 1];
            atomicAdd(reinterpret_cast<unsigned long long *>(st.att + pair_base + s), 1ull);
            const double blo = s == s0 ? b_lo : st.beta[sb + s], bhi = s == s0 ? b_hi : st.beta[sb + s + 1];
            const double delta = (blo - bhi) * (et[rlo] - et[rhi]);
            const U4 rn = philox4x32_10((uint32_t)s, 0x80000000u, (uint32_t)cycle, (uint32_t)(cycle >> 32),
                                        st.seed, key1);
            const double u = (double)rn.x * 0x1p-32;
            if (delta >= 0.0 || u < exp64(delta)) {
                atomicAdd(reinterpret_cast<unsigned long long *>(st.acc + pair_base + s), 1ull);
                rat[s] = rhi;
                rat[s + 1] = rlo;
                tor[rlo] = s + 1;
                tor[rhi] = s;
            }
        }
        __syncthreads();
        for (int r = tid; r < nrep; r += PT_THREADS) {
            st.rep_at_temp[sb + r] = rat[r];
            st.temp_of_rep[sb + r] = tor[r];
        }
    }
    if (measure) {
        for (int sl = tid; sl < nrep; sl += PT_THREADS) {
            if (!st.local) {
                if (tor[sl] == 0) st.vcold[sb + sl] = 1;           // sl as a replica index
                if (tor[sl] == nrep - 1) st.vhot[sb + sl] = 1;
            }
            const int r = rat[sl];
            const double am = vv[r][0], ae = vv[r][1], ax = vv[r][2], ay = vv[r][3], an = vv[r][4];
            if (st.local) {
                if (slen < st.series_cap) {
                    double *sr = st.series + st.confs[conf].series_off + (int64_t)sl * 5 * st.series_cap;
                    for (int q = 0; q < 5; ++q) sr[(int64_t)q * st.series_cap + slen] = vv[r][q];
                }
                continue;
            }
            double Sv[NSUMS], best, hmin, hmax, hbin;
            double *S = st.sums + (int64_t)(sb + sl) * NSUMS;
            if (sl == tid) {
#pragma unroll
                for (int q = 0; q < NSUMS; ++q) Sv[q] = S0[q];
                best = best0; hmin = hmin0; hmax = hmax0; hbin = hbin0;
            } else {
#pragma unroll
                for (int q = 0; q < NSUMS; ++q) Sv[q] = S[q];
                best = st.best[sb + sl]; hmin = st.hmin[sb + sl]; hmax = st.hmax[sb + sl]; hbin = st.hbin[sb + sl];
            }
            const double ax2 = ax * ax, ay2 = ay * ay, m2 = am * am, e2 = ae * ae, n2 = an * an;
            const double str2 = ax2 + ay2;
            Sv[0] += am;      Sv[1] += fabs(am);      Sv[2] += m2;          Sv[3] += m2 * m2;
            Sv[4] += ae;      Sv[5] += e2;            Sv[6] += fabs(ax);    Sv[7] += fabs(ay);
            Sv[8] += ax2;     Sv[9] += ax2 * ax2;     Sv[10] += ay2;        Sv[11] += ay2 * ay2;
            Sv[12] += str2;   Sv[13] += str2 * str2;  Sv[14] += sqrt(str2); Sv[15] += fabs(ay2 - ax2);
            Sv[16] += an;     Sv[17] += fabs(an);     Sv[18] += n2;         Sv[19] += n2 * n2;
#pragma unroll
            for (int q = 0; q < NSUMS; ++q) S[q] = Sv[q];
            if (ae < best) st.best[sb + sl] = ae;
            st.final_en[sb + sl] = ae;
            if (hbin <= 0.0) {
                hmin = ae - (fabs(ae) + 1.0);
                hmax = ae + (fabs(ae) + 1.0);
                hbin = (hmax - hmin) / (double)st.bins;
                st.hmin[sb + sl] = hmin;
                st.hmax[sb + sl] = hmax;
                st.hbin[sb + sl] = hbin;
            }
            const double tq = (ae - hmin) / hbin;
            int idx;
            if (tq < 0.0) idx = 1;
            else if (tq >= (double)st.bins) idx = st.bins;
            else idx = 1 + (int)tq;
            idx = idx < 1 ? 1 : (idx > st.bins ? st.bins : idx);
            atomicAdd(reinterpret_cast<unsigned long long *>(st.hist + hist_off + (int64_t)sl * st.bins + idx - 1),
                      1ull);
        }
        if (tid == 0) {
            atomicAdd(reinterpret_cast<unsigned long long *>(st.nmeasure + conf), 1ull);
            if (st.local) st.series_len[conf] = slen + 1;
        }
    }
    // The replica now at slot PTMEASINDEX: a copy of its packed spins (byte-
    // layout configs), from which every block of the config's next sweep
    // launch adds its share to the local nematic sums (the sweep overwrites
    // the spins themselves)
    if (swaps && measure && st.fuse_locnem && st.locsnap) {
        const LatDesc d = st.lats[lat0 + rat[st.ptmeasindex - 1]];
        if (d.packed == 1) {
            // eight words per thread requested before the first is stored
            const uint2 *src = reinterpret_cast<const uint2 *>(st.spins + d.spin_off);
            uint2 *dst = reinterpret_cast<uint2 *>(st.locsnap + d.site_off / 32);
            constexpr int CB = 8;
            for (int i0 = tid; i0 < d.N / 64; i0 += CB * PT_THREADS) {
                uint2 v[CB];
#pragma unroll
                for (int q = 0; q < CB; ++q) {
                    const int i = i0 + q * PT_THREADS;
                    if (i < d.N / 64) v[q] = src[i];
                }
#pragma unroll
                for (int q = 0; q < CB; ++q) {
                    const int i = i0 + q * PT_THREADS;
                    if (i < d.N / 64) dst[i] = v[q];
                }
            }
        }
    }
    if (swaps && tid == 0) {
        if (st.fuse_locnem) st.locnem_pending[conf] = measure ? 1 : 0;
        // one tick per block (nconf per cycle, see clock_cycle): the block's
        // single wave read the clock above, and the next kernel sees every
        // tick across the launch boundary, so no fence and no returned value
        // are needed
        atomicAdd(reinterpret_cast<unsigned long long *>(&st.clk->ticks), 1ull);
    }
}

// ---------------------------------------------------------------------------
// Local nematic accumulation (J1J2_6tv.f90:1009-1025) for the replica at slot
// PTMEASINDEX, as exact integer sums, from global memory.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_locnem(DeviceState st, int32_t only_pending)
{
    const int conf = blockIdx.y;
    if (only_pending && !st.locnem_pending[conf]) return;
    const ConfDesc cd = st.confs[conf];
    const int r = st.rep_at_temp[cd.slot_base + st.ptmeasindex - 1];
    const LatDesc d = st.lats[cd.lat0 + r];
    const int L = d.L, H = d.H, N = d.N;
    const uint8_t *sp = st.spins + d.spin_off;
    const uint32_t *sp32 = reinterpret_cast<const uint32_t *>(sp);
    int32_t *ls = st.locspin + d.site_off, *ln = st.locnem + d.site_off;
    const int t0 = blockIdx.x * 256 + threadIdx.x, stride = gridDim.x * 256;
    const int pk = d.packed;
    switch (pk == 1 ? L : (pk == 0 ? L + 1 : 0)) {   // bits: L % 8 == 0, L <= 48; unit words: the loop
        case 8: locnem_units<1, 0, 1>(sp32, H, t0, stride, ls, ln); break;
        case 16: locnem_units<2, 0, 1>(sp32, H, t0, stride, ls, ln); break;
        case 24: locnem_units<3, 0, 1>(sp32, H, t0, stride, ls, ln); break;
        case 32: locnem_units<4, 0, 1>(sp32, H, t0, stride, ls, ln); break;
        case 40: locnem_units<5, 0, 1>(sp32, H, t0, stride, ls, ln); break;
        case 48: locnem_units<6, 0, 1>(sp32, H, t0, stride, ls, ln); break;
        case 9: locnem_units<1, 0>(sp32, H, t0, stride, ls, ln); break;
        case 17: locnem_units<2, 0>(sp32, H, t0, stride, ls, ln); break;
        case 25: locnem_units<3, 0>(sp32, H, t0, stride, ls, ln); break;
        case 33: locnem_units<4, 0>(sp32, H, t0, stride, ls, ln); break;
        case 41: locnem_units<5, 0>(sp32, H, t0, stride, ls, ln); break;
        case 49: locnem_units<6, 0>(sp32, H, t0, stride, ls, ln); break;
        default:
            for (int i = t0; i < N; i += stride) {
                int x, y, z;
                site_of_internal(L, H, i, x, y, z);
                const Nbr nb = neighbours(L, H, x, y, z);
                const int si = spin_at(sp, pk, L, i) ? 1 : -1;
                const int l = si * ((spin_at(sp, pk, L, nb.m[2]) ? 1 : -1) - (spin_at(sp, pk, L, nb.m[0]) ? 1 : -1) +
                                    (spin_at(sp, pk, L, nb.m[3]) ? 1 : -1) - (spin_at(sp, pk, L, nb.m[1]) ? 1 : -1));
                ls[i] += si;
                ln[i] += l;
            }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st.locsamples[conf] += 1;
}

__global__ void k_begin_phase(DeviceClock *clk) { clk->phase_start = clk->ticks; }

}  // namespace

size_t resident_lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    size_t lim = (size_t)v;
    return lim > RES_STATIC_LDS ? lim - RES_STATIC_LDS : 0;
}

ResKernel pick_resident(int32_t L, int32_t nt, bool shadow, bool cubic = false, bool flips = false)
{
    switch (nt) {
        case 256: return resident_kernels<256>(L, shadow, cubic, flips);
        case 512: return resident_kernels<512>(L, shadow, cubic, flips);
        case 576: return resident_kernels<576>(L, shadow, cubic, flips);
        case 1024: return resident_kernels<1024>(L, shadow, cubic, flips);
        default: return nullptr;
    }
}

hipError_t setup_resident(size_t max_lds)
{
    for (int L : {8, 16, 24, 32, 40, 48, 0})
        for (int nt : {256, 512, 576, 1024})
            for (bool sh : {false, true})
                for (int v = 0; v < 4; ++v) {
                    const bool cubic = v & 1, flips = v >> 1;
                    const ResKernel k = pick_resident(L, nt, sh, cubic, flips);
                    if (!k) continue;
                    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)(max_lds + RES_STATIC_LDS));
                    if (e != hipSuccess) return e;
                }
    return hipSuccess;
}

bool resident_nt_valid(int32_t L, int32_t nt) { return pick_resident(L, nt, false) != nullptr; }

bool resident_supported(int32_t L) { return L >= 8 && L <= 48 && (L % 8) == 0; }

hipError_t launch_sweep_resident(const DeviceState &st, int32_t ord0, int32_t nord, int32_t max_n,
                                 int32_t L, int32_t H, int32_t nt, bool shadow, hipStream_t s)
{
    if (nord <= 0) return hipSuccess;
    const bool cubic = L > 0 ? H == L : H < 0, flips = st.flips != nullptr;
    const ResKernel k = pick_resident(L, nt, shadow, cubic, flips && fast_counts_ok(L, cubic, shadow, st.ptswapever));
    if (!k) return hipErrorInvalidValue;
    // with flip counts: a byte per site more, unless the kernel keeps them in registers
    const bool lds_counts = flips && k == pick_resident(L, nt, shadow, cubic, false);
    const size_t lattice = (size_t)max_n * (lds_counts ? 2 : 1);
    hipLaunchKernelGGL(k, dim3(nord), dim3(nt), lattice + RES_STATIC_LDS, s, st, ord0, nord);
    return hipGetLastError();
}

hipError_t launch_sweep_global(const DeviceState &st, int32_t lat_begin, int32_t nlat, int32_t L, int32_t H,
                               int32_t k, int32_t colour, bool shadow, hipStream_t s)
{
    if (nlat <= 0) return hipSuccess;
    const int ncalls = global_calls(L, H);
    const dim3 grid((ncalls + GLB_THREADS - 1) / GLB_THREADS, nlat), block(GLB_THREADS);
    if (shadow)
        hipLaunchKernelGGL((k_sweep_global<1, true>), grid, block, 0, s, st, lat_begin, L, H, k, colour);
    else
        hipLaunchKernelGGL((k_sweep_global<1, false>), grid, block, 0, s, st, lat_begin, L, H, k, colour);
    return hipGetLastError();
}

hipError_t launch_sweep_plain64(const DeviceState &st, int32_t lat_begin, int32_t nlat, int32_t L, int32_t H,
                                int32_t k, int32_t colour, hipStream_t s)
{
    if (nlat <= 0) return hipSuccess;
    const int ncalls = global_calls(L, H);
    const dim3 grid((ncalls + GLB_THREADS - 1) / GLB_THREADS, nlat), block(GLB_THREADS);
    hipLaunchKernelGGL((k_sweep_global<0, false>), grid, block, 0, s, st, lat_begin, L, H, k, colour);
    return hipGetLastError();
}

hipError_t launch_observables(const DeviceState &st, int32_t lat_begin, int32_t nlat, hipStream_t s)
{
    if (nlat <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_observables, dim3(nlat), dim3(OBS_THREADS), 0, s, st, lat_begin);
    return hipGetLastError();
}

hipError_t launch_pt_step(const DeviceState &st, int32_t mode, hipStream_t s)
{
    hipLaunchKernelGGL(k_pt_step, dim3(st.nconf), dim3(PT_THREADS), 0, s, st, mode);
    return hipGetLastError();
}

hipError_t launch_locnem(const DeviceState &st, int32_t max_n, int32_t only_pending, hipStream_t s)
{
    const int chunks = (max_n / 4 + 256 * 4 - 1) / (256 * 4);
    hipLaunchKernelGGL(k_locnem, dim3(chunks, st.nconf), dim3(256), 0, s, st, only_pending);
    return hipGetLastError();
}

hipError_t launch_begin_phase(const DeviceState &st, hipStream_t s)
{
    hipLaunchKernelGGL(k_begin_phase, dim3(1), dim3(1), 0, s, st.clk);
    return hipGetLastError();
}

}  // namespace j1j2
