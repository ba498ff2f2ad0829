// engine.cpp -- host side of the J1-J2 MI355X engine: parameter validation,
// device buffers, per-cycle launch schedule (eager or replayed hipGraph),
// spin-layout conversion and the exported C-ABI (include/j1j2_engine.h).
#include "j1j2_engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_model.h"
#include "kernels.h"

using namespace j1j2;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? J1J2_ENOMEM : J1J2_EDEVICE,               \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

constexpr double LOG2E = 0x1.71547652b82fep+0;

// Work-list segment of a lattice range (one group, or all lattices).
struct OrderSeg {
    int32_t ord0 = 0, nord = 0;
};

struct Group {
    j1j2_params p;
    int32_t conf0, lat0, nlat, N;
    bool resident;             // swept by an LDS-resident kernel (observables fused)
    int kind = 0;              // 0 global-memory passes, 1 resident bytes, 2 resident bits
    OrderSeg slab_seg[BIT_SLAB_MAX + 1];   // kind 2: work lists of the lattices split into S slabs
};

struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> live, pool;
    double total_ms = 0.0;
    int64_t launches = 0;
};

// One launch of the LDS-resident sweep: the blocks of work-list segment
// [ord0, ord0 + nord), lattices of common size L (0: mixed sizes), nt threads
// per block, nlat lattices.
struct ResLaunch {
    int32_t ord0, nord, L, max_n, nt, nlat;
    int32_t kind = 1, H = 0;   // kind 2: bit layout (nt fixed by L)
    int32_t slabs = 1;         // kind 2: z-slabs per lattice (one block each)
    bool bundled = false;      // work-list entries are bundles (the launch over several sizes)
    // the L the kernel is picked by: a bundled launch takes the kernel over
    // several sizes (L = 0, its H then -1 when every lattice is cubic)
    int32_t kernel_L() const { return bundled ? 0 : L; }
};

// How a cycle's resident sweeps are launched.  The result of a cycle does not
// depend on the plan (site updates are keyed by lattice, colour and word, not
// by thread), so the autotuner may time candidates on real cycles.
struct Plan {
    std::vector<ResLaunch> launches;
    bool concurrent = false;   // one parallel graph branch (side stream) per launch
    std::string describe() const
    {
        std::string d = concurrent ? "concurrent" : "sequential";
        for (const ResLaunch &r : launches)
            d += " [L" + std::to_string(r.L) + " nt" + std::to_string(r.nt) + " x" + std::to_string(r.nlat) +
                 (r.slabs > 1 ? " s" + std::to_string(r.slabs) : std::string()) + (r.bundled ? " bundled" : "") + "]";
        return d;
    }
};

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Group> groups;
    j1j2_params common{};
    std::vector<ConfDesc> confs;
    std::vector<ConfConst> cconst;
    std::vector<LatDesc> lats;
    std::vector<std::vector<double>> temps;   // per config
    std::vector<double> hbound;               // per config: bound on |h| of the FP32 rule
    int64_t nsites = 0, nslots = 0, npairs = 0, nhist = 0, nseries = 0, nflips = 0, spin_bytes = 0;
    int32_t max_n = 0;
    int32_t path = 0;                          // 0 FP64 plain, 1 FP32 resident, 2 FP32 global, 3 resident incl. bits
    bool all_resident = false, shadow = false, locnem = false, use_graph = true, timing = false;
    bool all_byte = false;                     // every group on the byte-layout resident kernel
    bool all_one = false;                      // ... and one launch over all lattices fits the LDS
    int32_t res_nt = 256;                      // threads per block of the resident sweep (default plan)
    int32_t bit_slabs = 1;                     // z-slabs of the bit-layout lattices (default plan)
    int32_t uniform_L = 0;                     // common L of all groups, 0 if mixed
    int32_t uniform_H = 0;                     // common H of all groups, 0 if mixed
    bool all_cubic = false;                    // every group has H = L
    Plan plan;                                 // resident launches of a cycle
    std::vector<OrderSeg> group_seg;           // work-list segment of each resident group
    OrderSeg all_seg;                          // of all lattices (when all are resident)
    OrderSeg all_seg_b;                        // the same in bundles (1024-thread FP32 launch)
    int32_t all_max_b = 0;                     // the bundles' largest LDS lattice bytes (without flip counts)
    int32_t order_len = 0;                     // entries of the device work list (st.order)
    bool tuned = false, autotune = true;
    std::string tune_key;
    std::vector<hipStream_t> side;             // one per concurrent launch
    hipEvent_t fork_ev = nullptr;
    std::vector<hipEvent_t> join_ev;
    uint64_t cycles_done = 0;
    bool observed = false;                     // observables valid for the current state
    std::vector<void *> allocs;
    DeviceState st{};
    // replayed schedules: graph[measure][k] = 2^k cycles (k < GRAPH_SIZES)
    hipGraphExec_t graph[2][5] = {};
    Timer timers[4];

    ~Engine()
    {
        for (auto &gm : graph)
            for (auto &g : gm)
                if (g) (void)hipGraphExecDestroy(g);
        for (auto &t : timers) {
            for (auto &e : t.live) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
            for (auto &e : t.pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        }
        for (void *p : allocs) (void)hipFree(p);
        for (hipEvent_t e : join_ev) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        for (hipStream_t q : side) (void)hipStreamDestroy(q);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T>
    int alloc(T *&ptr, int64_t n)
    {
        void *p = nullptr;
        const size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
        HIPCHK(hipMalloc(&p, bytes));
        allocs.push_back(p);
        HIPCHK(hipMemsetAsync(p, 0, bytes, stream));
        ptr = static_cast<T *>(p);
        return 0;
    }
};

int validate_group(const j1j2_params &p)
{
    // Validate_run_options (J1J2_6tv.f90:841-925) plus the engine's own limits.
    if (p.L < 2 || p.H < 2 || (p.L & 1) || (p.H & 1))
        return fail(J1J2_EINVAL, "L and H must be even and >= 2 (four-colour sweep)");
    if ((int64_t)p.L * p.L * p.H > (1 << 22)) return fail(J1J2_EINVAL, "L*L*H must not exceed 2^22");
    if (p.sampler != J1J2_SAMPLER_PT && p.sampler != J1J2_SAMPLER_LOCAL)
        return fail(J1J2_EINVAL, "SAMPLERMODE must be LOCAL or PT");
    if (p.precision != J1J2_FP64_PLAIN && p.precision != J1J2_FP32_BATCHED)
        return fail(J1J2_EINVAL, "precision must be FP64_PLAIN or FP32_BATCHED");
    if (p.ladder != J1J2_LADDER_LINEAR && p.ladder != J1J2_LADDER_SPLIT && p.ladder != J1J2_LADDER_LOCAL)
        return fail(J1J2_EINVAL, "PTLADDERMODE must be LINEAR or SPLIT (or LOCAL spacing)");
    if (p.nreplica < 1 || p.nreplica > 256) return fail(J1J2_EINVAL, "NREPLICA must be in [1, 256]");
    if (p.nconf < 1) return fail(J1J2_EINVAL, "NCONF must be >= 1");
    if (p.bins < 1) return fail(J1J2_EINVAL, "BINS must be >= 1");
    if (p.ptswapever <= 0) return fail(J1J2_EINVAL, "PTSWAPEVER must be positive");
    if (p.sampler == J1J2_SAMPLER_LOCAL && p.ptswapever != 1)
        return fail(J1J2_EINVAL, "LOCAL sampler measures every sweep: PTSWAPEVER must be 1");
    if (p.ptmeasindex < 1 || p.ptmeasindex > p.nreplica) return fail(J1J2_EINVAL, "PTMEASINDEX out of range");
    if (p.neq < 0 || p.nmess < 0) return fail(J1J2_EINVAL, "NEQ and NMESS must be >= 0");
    if (p.tmax < p.tmin) return fail(J1J2_EINVAL, "TMAX must be >= TMIN");
    if (p.ladder == J1J2_LADDER_LOCAL && !(p.dt > 0.0)) return fail(J1J2_EINVAL, "DT must be positive");
    if (p.ladder == J1J2_LADDER_SPLIT) {
        if (p.nreplica < 3) return fail(J1J2_EINVAL, "SPLIT ladder requires at least 3 temperatures");
        if (p.tladdermid <= p.tmin || p.tladdermid >= p.tmax)
            return fail(J1J2_EINVAL, "TLADDERMID must lie strictly inside (TMIN,TMAX)");
        if (p.tladderfrac <= 0.0 || p.tladderfrac >= 1.0)
            return fail(J1J2_EINVAL, "TLADDERFRAC must lie strictly inside (0,1)");
    }
    if (!(p.tmin > 0.0)) return fail(J1J2_EINVAL, "temperatures must be positive");
    return 0;
}

int ref_site(int L, int x, int y, int z) { return x + L * y + L * L * z; }

// host-side FP32 constants, fixed-point disorder and FP64 field of one config,
// written in the internal site order
void build_site_planes(const ConfDesc &cd, const ConfConst &cc, const double *p, std::vector<float> kd[5],
                       std::vector<double> &p64, std::vector<uint8_t> &qb, int64_t &qsum)
{
    const int L = cd.L, H = cd.H, N = cd.N;
    std::vector<float> Kx(N), Ky(N);
    for (int z = 0; z < H; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int i = ref_site(L, x, y, z);
                Kx[i] = (float)((cc.J1h + p[i]) + p[ref_site(L, (x + 1) % L, y, z)]);
                Ky[i] = (float)((cc.J1v - p[i]) - p[ref_site(L, x, (y + 1) % L, z)]);
            }
    for (int P = 0; P < 5; ++P) kd[P].assign(N, 0.0f);
    p64.assign(N, 0.0);
    qb.assign((size_t)5 * N, 0);
    qsum = 0;
    for (int z = 0; z < H; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) {
                const int i = ref_site(L, x, y, z);
                const int64_t ii = internal_index(L, H, x, y, z);
                const float kxp = Kx[i], kxm = Kx[ref_site(L, (x + L - 1) % L, y, z)];
                const float kyp = Ky[i], kym = Ky[ref_site(L, x, (y + L - 1) % L, z)];
                const double sK = (((double)kxp + (double)kxm) + (double)kyp) + (double)kym;
                kd[0][ii] = (float)((-sK - 4.0 * cc.J2) - 2.0 * cc.J3);
                // 2K scaled by 2^14 (exact): the sweep multiplies them by 2^-14 b
                kd[1][ii] = std::ldexp(2.0f * kxp, 14);
                kd[2][ii] = std::ldexp(2.0f * kxm, 14);
                kd[3][ii] = std::ldexp(2.0f * kyp, 14);
                kd[4][ii] = std::ldexp(2.0f * kym, 14);
                p64[ii] = p[i];
                const int64_t q = (int64_t)std::llrint(p[i] * 0x1p32);     // |q| < 2^38 (|p| < 64)
                const uint64_t qo = (uint64_t)(q + QOFF);                   // five byte planes of q + 2^38
                for (int b = 0; b < 5; ++b) qb[(size_t)b * N + ii] = (uint8_t)(qo >> (8 * b));
                qsum += q;
            }
}

int upload_site_planes(Engine &E, int conf, const double *p)
{
    const ConfDesc &cd = E.confs[conf];
    std::vector<float> kd[5];
    std::vector<double> p64;
    std::vector<uint8_t> qb;
    int64_t qsum = 0;
    build_site_planes(cd, E.cconst[conf], p, kd, p64, qb, qsum);
    E.cconst[conf].qsum = qsum;
    HIPCHK(hipMemcpyAsync(const_cast<ConfConst *>(E.st.cconst) + conf, &E.cconst[conf], sizeof(ConfConst),
                          hipMemcpyHostToDevice, E.stream));
    for (int P = 0; P < 5; ++P)
        HIPCHK(hipMemcpyAsync(const_cast<float *>(E.st.kd) + P * E.nsites + cd.site_off, kd[P].data(),
                              sizeof(float) * cd.N, hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipMemcpyAsync(const_cast<double *>(E.st.p64) + cd.site_off, p64.data(), sizeof(double) * cd.N,
                          hipMemcpyHostToDevice, E.stream));
    for (int b = 0; b < 5; ++b)
        HIPCHK(hipMemcpyAsync(const_cast<uint8_t *>(E.st.qb) + b * E.nsites + cd.site_off, qb.data() + (size_t)b * cd.N,
                              cd.N, hipMemcpyHostToDevice, E.stream));
    HIPCHK(hipStreamSynchronize(E.stream));
    return 0;
}

template <typename T>
int h2d(Engine &E, T *dst, const std::vector<T> &src)
{
    HIPCHK(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, E.stream));
    return 0;
}

template <typename T>
int d2h(Engine &E, std::vector<T> &dst, const T *src, int64_t n)
{
    dst.resize((size_t)n);
    HIPCHK(hipMemcpyAsync(dst.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, E.stream));
    return 0;
}

// Zero the measurement accumulators of one config (start of Pt_measurement_loop,
// J1J2_6tv.f90:1375-1399).
int reset_measurement(Engine &E, int conf)
{
    const ConfDesc &cd = E.confs[conf];
    const int nrep = cd.nrep, sb = cd.slot_base;
    HIPCHK(hipMemsetAsync(E.st.sums + (int64_t)sb * NSUMS, 0, sizeof(double) * nrep * NSUMS, E.stream));
    int rc;
    if ((rc = h2d(E, E.st.best + sb, std::vector<double>(nrep, 1.7976931348623157e308)))) return rc;
    HIPCHK(hipMemsetAsync(E.st.final_en + sb, 0, sizeof(double) * nrep, E.stream));
    HIPCHK(hipMemsetAsync(E.st.hist + cd.hist_off, 0, sizeof(int64_t) * nrep * E.common.bins, E.stream));
    HIPCHK(hipMemsetAsync(E.st.hmin + sb, 0, sizeof(double) * nrep, E.stream));
    HIPCHK(hipMemsetAsync(E.st.hmax + sb, 0, sizeof(double) * nrep, E.stream));
    HIPCHK(hipMemsetAsync(E.st.hbin + sb, 0, sizeof(double) * nrep, E.stream));
    HIPCHK(hipMemsetAsync(E.st.vcold + sb, 0, nrep, E.stream));
    HIPCHK(hipMemsetAsync(E.st.vhot + sb, 0, nrep, E.stream));
    HIPCHK(hipMemsetAsync(E.st.nmeasure + conf, 0, sizeof(int64_t), E.stream));
    HIPCHK(hipMemsetAsync(E.st.locspin + cd.site_off, 0, sizeof(int32_t) * cd.N, E.stream));
    HIPCHK(hipMemsetAsync(E.st.locnem + cd.site_off, 0, sizeof(int32_t) * cd.N, E.stream));
    HIPCHK(hipMemsetAsync(E.st.locsamples + conf, 0, sizeof(int64_t), E.stream));
    HIPCHK(hipMemsetAsync(E.st.series_len + conf, 0, sizeof(int64_t), E.stream));
    return 0;
}

// Initialize_pt_state (J1J2_6tv.f90:1218-1247): identity maps, cleared PT
// diagnostics, flip counts and FP32/FP64 mismatch counters.
int reset_config(Engine &E, int conf)
{
    const ConfDesc &cd = E.confs[conf];
    const int nrep = cd.nrep, sb = cd.slot_base;
    std::vector<int32_t> id(nrep);
    for (int r = 0; r < nrep; ++r) id[r] = r;
    int rc;
    if ((rc = h2d(E, E.st.rep_at_temp + sb, id))) return rc;
    if ((rc = h2d(E, E.st.temp_of_rep + sb, id))) return rc;
    if (nrep > 1) {
        HIPCHK(hipMemsetAsync(E.st.att + cd.pair_base, 0, sizeof(int64_t) * (nrep - 1), E.stream));
        HIPCHK(hipMemsetAsync(E.st.acc + cd.pair_base, 0, sizeof(int64_t) * (nrep - 1), E.stream));
    }
    if (E.st.flips) {
        HIPCHK(hipMemsetAsync(E.st.flips + cd.flip_off, 0, sizeof(int32_t) * nrep * cd.N, E.stream));
        HIPCHK(hipMemsetAsync(E.st.flips_pend + cd.flip_off, 0, (size_t)nrep * cd.N, E.stream));
    }
    HIPCHK(hipMemsetAsync(E.st.shadow + 2 * conf, 0, 2 * sizeof(unsigned long long), E.stream));
    return reset_measurement(E, conf);
}

int set_ladder(Engine &E, int conf, const double *t)
{
    const ConfDesc &cd = E.confs[conf];
    std::vector<double> beta(cd.nrep);
    std::vector<float> b2(cd.nrep);
    for (int s = 0; s < cd.nrep; ++s) {
        if (!(t[s] > 0.0)) return fail(J1J2_EINVAL, "temperatures must be positive");
        beta[s] = 1.0 / t[s];
        b2[s] = (float)(2.0 * beta[s] * LOG2E);
    }
    E.temps[conf].assign(t, t + cd.nrep);
    int rc;
    if ((rc = h2d(E, const_cast<double *>(E.st.beta) + cd.slot_base, beta))) return rc;
    if ((rc = h2d(E, const_cast<float *>(E.st.b2) + cd.slot_base, b2))) return rc;
    HIPCHK(hipStreamSynchronize(E.stream));
    return 0;
}

// ---------------------------------------------------------------- schedule
int timed(Engine &E, int kind, hipError_t (*fn)(Engine &, const void *), const void *arg)
{
    if (!E.timing) {
        HIPCHK(fn(E, arg));
        return 0;
    }
    Timer &T = E.timers[kind];
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!T.pool.empty()) {
        ev = T.pool.back();
        T.pool.pop_back();
    } else {
        HIPCHK(hipEventCreate(&ev.first));
        HIPCHK(hipEventCreate(&ev.second));
    }
    HIPCHK(hipEventRecord(ev.first, E.stream));
    HIPCHK(fn(E, arg));
    HIPCHK(hipEventRecord(ev.second, E.stream));
    T.live.push_back(ev);
    return 0;
}

int collect_timers(Engine &E)
{
    for (auto &T : E.timers) {
        for (auto &ev : T.live) {
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, ev.first, ev.second));
            T.total_ms += ms;
            T.launches += 1;
            T.pool.push_back(ev);
        }
        T.live.clear();
    }
    return 0;
}

struct SweepArg { int g; int k; int c; };

// Events the autotuner records around the resident part of one cycle: per
// launch (sequential plans) and at the end of the whole resident phase.
struct Probe {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> per;
    hipEvent_t phase_end = nullptr;
};

// Every resident launch must read its blocks' lattices inside the device work
// list: a segment past its end would send blocks to arbitrary lattices.
int check_plan(const Engine &E, const Plan &P)
{
    for (const ResLaunch &r : P.launches)
        if (r.ord0 < 0 || r.nord < 0 || (int64_t)r.ord0 + r.nord > E.order_len || (r.nord & 7))
            return fail(J1J2_EINVAL, "internal error: a resident launch's work-list segment [" +
                                         std::to_string(r.ord0) + ", +" + std::to_string(r.nord) +
                                         ") lies outside the work list of " + std::to_string(E.order_len) +
                                         " entries");
    return 0;
}

int enqueue_cycle(Engine &E, int mode, const Plan *plan = nullptr, Probe *probe = nullptr)
{
    // mode bit 0: measure; bit 1: values only (no sweeps, no swaps)
    const bool measure = mode & 1, values_only = mode & 2;
    int rc;
    if (!values_only) {
        const Plan &P = plan ? *plan : E.plan;
        if ((rc = check_plan(E, P))) return rc;
        if (P.concurrent && P.launches.size() > 1) {
            // fork: every launch on its own branch, join before the swap step
            auto fn = [](Engine &e, const void *a) {
                const Plan &pl = *static_cast<const Plan *>(a);
                hipError_t r = hipEventRecord(e.fork_ev, e.stream);
                for (size_t i = 0; i < pl.launches.size() && r == hipSuccess; ++i) {
                    const ResLaunch &L = pl.launches[i];
                    if ((r = hipStreamWaitEvent(e.side[i], e.fork_ev, 0)) != hipSuccess) break;
                    r = L.kind == 2 ? launch_sweep_bits(e.st, L.ord0, L.nord, L.L, L.H, L.slabs, e.shadow, e.side[i])
                                    : launch_sweep_resident(e.st, L.ord0, L.nord, L.max_n, L.kernel_L(), L.H, L.nt,
                                                            e.shadow, e.side[i]);
                    if (r != hipSuccess) break;
                    if ((r = hipEventRecord(e.join_ev[i], e.side[i])) != hipSuccess) break;
                    r = hipStreamWaitEvent(e.stream, e.join_ev[i], 0);
                }
                return r;
            };
            if ((rc = timed(E, 0, fn, &P))) return rc;
        } else {
            for (const ResLaunch &L : P.launches) {
                auto fn = [](Engine &e, const void *a) {
                    const ResLaunch &r = *static_cast<const ResLaunch *>(a);
                    return r.kind == 2 ? launch_sweep_bits(e.st, r.ord0, r.nord, r.L, r.H, r.slabs, e.shadow, e.stream)
                                       : launch_sweep_resident(e.st, r.ord0, r.nord, r.max_n, r.kernel_L(), r.H,
                                                               r.nt, e.shadow, e.stream);
                };
                const size_t i = (size_t)(&L - P.launches.data());
                if (probe && i < probe->per.size()) HIPCHK(hipEventRecord(probe->per[i].first, E.stream));
                if ((rc = timed(E, 0, fn, &L))) return rc;
                if (probe && i < probe->per.size()) HIPCHK(hipEventRecord(probe->per[i].second, E.stream));
            }
        }
        if (probe && probe->phase_end) HIPCHK(hipEventRecord(probe->phase_end, E.stream));
        {
            for (int g = 0; g < (int)E.groups.size(); ++g) {
                const Group &G = E.groups[g];
                if (G.resident) continue;
                for (int k = 0; k < E.common.ptswapever; ++k)
                    for (int c = 0; c < 4; ++c) {
                        SweepArg a{g, k, c};
                        auto fn = [](Engine &e, const void *p) {
                            const SweepArg &s = *static_cast<const SweepArg *>(p);
                            const Group &gg = e.groups[s.g];
                            if (e.common.precision == J1J2_FP64_PLAIN)
                                return launch_sweep_plain64(e.st, gg.lat0, gg.nlat, gg.p.L, gg.p.H, s.k, s.c,
                                                            e.stream);
                            return launch_sweep_global(e.st, gg.lat0, gg.nlat, gg.p.L, gg.p.H, s.k, s.c,
                                                       e.shadow, e.stream);
                        };
                        if ((rc = timed(E, 0, fn, &a))) return rc;
                    }
            }
        }
    }
    for (int g = 0; g < (int)E.groups.size(); ++g) {
        if (E.groups[g].resident && !values_only) continue;
        auto fn = [](Engine &e, const void *a) {
            const Group &gg = e.groups[*static_cast<const int *>(a)];
            return launch_observables(e.st, gg.lat0, gg.nlat, e.stream);
        };
        if ((rc = timed(E, 2, fn, &g))) return rc;
    }
    auto ptfn = [](Engine &e, const void *a) { return launch_pt_step(e.st, *static_cast<const int *>(a), e.stream); };
    const int ptmode = values_only ? 2 : (measure ? 1 : 0);
    if ((rc = timed(E, 1, ptfn, &ptmode))) return rc;
    if (values_only) return 0;
    if (measure && E.locnem && !E.st.fuse_locnem) {
        auto fn = [](Engine &e, const void *) { return launch_locnem(e.st, e.max_n, 0, e.stream); };
        if ((rc = timed(E, 3, fn, nullptr))) return rc;
    }
    return 0;
}

// A call's cycles replay as graphs of 16 cycles and then one graph each of
// 8, 4, 2, 1 cycles for the remainder (at most 4 more launches: each graph
// launch costs tens of microseconds of GPU-side gap, which single-cycle
// graphs would pay per cycle).
constexpr int GRAPH_SIZES = 5;

int build_graph(Engine &E, int measure, int lg)
{
    if (E.graph[measure][lg]) return 0;
    HIPCHK(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int i = 0; i < (1 << lg) && rc == 0; ++i) rc = enqueue_cycle(E, measure);
    hipGraph_t cap = nullptr;
    const hipError_t e2 = hipStreamEndCapture(E.stream, &cap);
    if (rc) {
        if (cap) (void)hipGraphDestroy(cap);
        return rc;
    }
    HIPCHK(e2);
    hipGraphExec_t ex = nullptr;
    const hipError_t e3 = hipGraphInstantiate(&ex, cap, nullptr, nullptr, 0);
    (void)hipGraphDestroy(cap);
    HIPCHK(e3);
    E.graph[measure][lg] = ex;
    return 0;
}

// ---------------------------------------------------------------- autotuner
// Picks, per engine shape, the threads per block of each resident launch and
// whether the lattices of different sizes run as one launch or as parallel
// per-group launches.  Candidates are timed on real cycles of the run (the
// result does not depend on the plan) and the choice is cached per process
// by (device architecture, groups' L, H, lattice count, shadow).
std::mutex &tune_mutex()
{
    static std::mutex m;
    return m;
}

// What the autotuner chose for an engine shape.  Only the choices are cached,
// never work-list offsets: every engine rebuilds its launches from its own
// work-list segments (plan_from_choice), so an engine whose configs split the
// same lattices differently still launches exactly its own lists.
struct TuneChoice {
    std::vector<int32_t> pick;   // per resident group, in group order: threads per block (byte layout) or slabs (bits)
    bool concurrent = false;     // the per-group launches on parallel graph branches
    int32_t all_nt = 0;          // > 0: one launch over all lattices with this many threads per block
};

std::map<std::string, TuneChoice> &tune_cache()
{
    static std::map<std::string, TuneChoice> m;
    return m;
}

// nt: threads per block of a byte-layout group; slabs: z-slabs of a bit-layout
// group (its block size is fixed by L)
ResLaunch group_launch(const Engine &E, int g, int32_t nt, int32_t slabs = 1)
{
    const Group &G = E.groups[g];
    const OrderSeg &o = (G.kind == 2 && slabs > 1) ? G.slab_seg[slabs] : E.group_seg[g];
    ResLaunch r{o.ord0, o.nord, G.p.L, G.N, G.kind == 2 ? 0 : nt, G.nlat};
    r.kind = G.kind;
    r.H = G.p.H;
    r.slabs = G.kind == 2 ? slabs : 1;
    return r;
}

ResLaunch all_launch(const Engine &E, int32_t nt)
{
    if (nt == 1024 && E.all_seg_b.nord > 0) {                  // bundles of small lattices
        ResLaunch r{E.all_seg_b.ord0, E.all_seg_b.nord, E.uniform_L, E.all_max_b, nt, E.st.nlat};
        r.H = E.all_cubic ? -1 : 0;
        r.bundled = true;
        return r;
    }
    ResLaunch r{E.all_seg.ord0, E.all_seg.nord, E.uniform_L, E.max_n, nt, E.st.nlat};
    r.H = E.uniform_L ? E.uniform_H : (E.all_cubic ? -1 : 0);
    return r;
}

Plan plan_from_choice(const Engine &E, const TuneChoice &t)
{
    Plan P;
    if (t.all_nt > 0 && E.all_one) {
        P.launches.push_back(all_launch(E, t.all_nt));
        return P;
    }
    size_t i = 0;
    for (int g = 0; g < (int)E.groups.size(); ++g) {
        const Group &G = E.groups[g];
        if (!G.resident) continue;
        int32_t v = i < t.pick.size() ? t.pick[i] : (G.kind == 2 ? 1 : E.res_nt);
        ++i;
        if (G.kind == 2 && v > 1 && !bits_slab_fits(G.p.L, G.p.H, v, G.slab_seg[v].nord, E.shadow)) v = 1;
        if (G.kind == 1 && !resident_nt_valid(G.p.L, v)) v = 256;
        P.launches.push_back(group_launch(E, g, v, v));
    }
    P.concurrent = t.concurrent && P.launches.size() > 1 &&
                   std::none_of(P.launches.begin(), P.launches.end(), [](const ResLaunch &r) { return r.slabs > 1; });
    return P;
}

Plan default_plan(const Engine &E)
{
    Plan P;
    if (E.all_one) {
        P.launches.push_back(all_launch(E, E.res_nt));
        return P;
    }
    for (int g = 0; g < (int)E.groups.size(); ++g)
        if (E.groups[g].resident) P.launches.push_back(group_launch(E, g, E.res_nt, E.bit_slabs));
    return P;
}

// Work list of the lattices of configs [c0, c1): configs largest first,
// each config's replicas together, configs dealt to the 8 XCDs by least
// accumulated work, then interleaved so that block 8k + x is XCD x's k-th
// lattice (-1 pads the shorter XCD lists).  slabs > 1: one entry
// BIT_SLAB_MAX * lattice + s per slab s, a lattice's slabs together on one XCD.
// bundle: entries of up to res_bundle(L) lattices of one config (first lattice
// plus (count - 1) << BUNDLE_SHIFT), for the 1024-thread launch over all sizes.
std::vector<int32_t> make_order(const Engine &E, int c0, int c1, int slabs = 1, bool bundle = false)
{
    std::vector<int> cs;
    for (int c = c0; c < c1; ++c) cs.push_back(c);
    std::stable_sort(cs.begin(), cs.end(), [&](int a, int b) { return E.confs[a].N > E.confs[b].N; });
    std::vector<std::vector<int32_t>> bins(8);
    std::vector<int64_t> work(8, 0);
    for (int c : cs) {
        const int x = (int)(std::min_element(work.begin(), work.end()) - work.begin());
        const ConfDesc &cd = E.confs[c];
        const int G = bundle ? res_bundle(cd.L) : 1;
        for (int r = 0; r < cd.nrep; r += G) {
            if (slabs > 1)
                for (int q = 0; q < slabs; ++q) bins[x].push_back((cd.lat0 + r) * BIT_SLAB_MAX + q);
            else
                bins[x].push_back((cd.lat0 + r) | ((std::min(G, cd.nrep - r) - 1) << BUNDLE_SHIFT));
        }
        work[x] += (int64_t)cd.N * cd.nrep;
    }
    size_t len = 0;
    for (auto &b : bins) len = std::max(len, b.size());
    std::vector<int32_t> out(8 * len, -1);
    for (size_t k = 0; k < len; ++k)
        for (int x = 0; x < 8; ++x)
            if (k < bins[x].size()) out[8 * k + x] = bins[x][k];
    return out;
}

int ensure_side(Engine &E, size_t n)
{
    if (!E.fork_ev) HIPCHK(hipEventCreateWithFlags(&E.fork_ev, hipEventDisableTiming));
    while (E.side.size() < n) {
        hipStream_t q = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        E.side.push_back(q);
        hipEvent_t ev = nullptr;
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        E.join_ev.push_back(ev);
    }
    return 0;
}

void drop_graphs(Engine &E)
{
    for (auto &gm : E.graph)
        for (auto &g : gm)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

// One eager cycle under plan P; per-launch times (sequential plans) and the
// resident phase's time, in ms.
int tune_cycle(Engine &E, int measure, const Plan &P, std::vector<float> &launch_ms, float &phase_ms)
{
    int rc;
    if (P.concurrent && (rc = ensure_side(E, P.launches.size()))) return rc;
    Probe probe;
    hipEvent_t start = nullptr;
    HIPCHK(hipEventCreate(&start));
    HIPCHK(hipEventCreate(&probe.phase_end));
    if (!P.concurrent)
        for (size_t i = 0; i < P.launches.size(); ++i) {
            std::pair<hipEvent_t, hipEvent_t> ev;
            HIPCHK(hipEventCreate(&ev.first));
            HIPCHK(hipEventCreate(&ev.second));
            probe.per.push_back(ev);
        }
    // a lead-in cycle of the same plan right before the timed one, so that
    // the timed launches start on a busy GPU as they do in a run
    rc = enqueue_cycle(E, measure, &P);
    if (!rc && hipEventRecord(start, E.stream) != hipSuccess) rc = fail(J1J2_EDEVICE, "hipEventRecord failed");
    if (!rc) rc = enqueue_cycle(E, measure, &P, &probe);
    if (!rc && hipStreamSynchronize(E.stream) != hipSuccess) rc = fail(J1J2_EDEVICE, "hipStreamSynchronize failed");
    launch_ms.assign(probe.per.size(), 0.0f);
    phase_ms = 0.0f;
    if (!rc) {
        for (size_t i = 0; i < probe.per.size(); ++i)
            (void)hipEventElapsedTime(&launch_ms[i], probe.per[i].first, probe.per[i].second);
        (void)hipEventElapsedTime(&phase_ms, start, probe.phase_end);
    }
    for (auto &ev : probe.per) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    (void)hipEventDestroy(probe.phase_end);
    (void)hipEventDestroy(start);
    return rc;
}

std::string tune_key(const Engine &E)
{
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, E.device);
    // flip counts double a resident block's LDS (fewer blocks per CU), so a
    // choice timed without them does not carry over
    std::string k = std::string(prop.gcnArchName) + (E.shadow ? " shadow" : "") + (E.st.flips ? " flips" : "");
    // the configs x replicas split fixes the work lists' shapes (their XCD
    // balance and padding), so it is part of the shape a choice was timed on
    for (const Group &G : E.groups)
        k += " " + std::to_string(G.p.L) + "x" + std::to_string(G.p.H) + ":" + std::to_string(G.p.nconf) + "x" +
             std::to_string(G.p.nreplica) + (G.resident ? "r" : "g");
    return k;
}

// Times the candidates on cycles of the call (counted in `done`) when all of
// them fit: at most TUNE_CYCLES x TUNE_REPS x (4 + 5) = 54.
constexpr int TUNE_REPS = 3;
constexpr int TUNE_CYCLES = 2;        // cycles per timing: a lead-in and the timed one
constexpr float TUNE_TIE = 1.005f;

int tune(Engine &E, int measure, int ncycles, int &done)
{
    std::vector<int> R;
    for (int g = 0; g < (int)E.groups.size(); ++g)
        if (E.groups[g].resident) R.push_back(g);
    if (R.empty()) {
        E.tuned = true;
        return 0;
    }
    if (E.tune_key.empty()) E.tune_key = tune_key(E);
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        auto it = tune_cache().find(E.tune_key);
        if (it != tune_cache().end()) {
            E.plan = plan_from_choice(E, it->second);
            E.tuned = true;
            drop_graphs(E);
            return ensure_side(E, E.plan.concurrent ? E.plan.launches.size() : 0);
        }
    }
    std::vector<std::vector<int32_t>> cand(R.size());
    size_t maxc = 0;
    for (size_t i = 0; i < R.size(); ++i) {
        const Group &G = E.groups[R[i]];
        if (G.kind == 2) {                                    // bit layout: whole lattices or z-slabs
            cand[i].push_back(1);
            for (int32_t S : {2, 4})
                if (bits_slab_fits(G.p.L, G.p.H, S, G.slab_seg[S].nord, E.shadow)) cand[i].push_back(S);
        } else
            for (int32_t nt : {256, 512, 576, 1024})
                if (resident_nt_valid(E.groups[R[i]].p.L, nt)) cand[i].push_back(nt);
        maxc = std::max(maxc, cand[i].size());
    }
    const bool multi = R.size() > 1;
    static constexpr int32_t ALL_NT[] = {256, 512, 576, 1024};     // one launch over all lattices
    constexpr int n_all = (int)(sizeof(ALL_NT) / sizeof(ALL_NT[0]));
    // cycles the candidates below may take (an upper bound: the call must hold them all)
    const int need = TUNE_CYCLES * (TUNE_REPS * (int)maxc + (multi ? TUNE_REPS * (1 + (E.all_one ? n_all : 0)) : 0));
    if (ncycles - done < need) return 0;             // too short a call; tune on a later one
    int rc;
    const bool log = std::getenv("J1J2_TUNE_LOG") != nullptr;   // each timed candidate to stderr
    // 1. threads per block of each group, its launch timed alone
    std::vector<std::vector<float>> best(R.size());
    for (size_t i = 0; i < R.size(); ++i) best[i].assign(cand[i].size(), 1e30f);
    for (int rep = 0; rep < TUNE_REPS; ++rep)
        for (size_t r = 0; r < maxc; ++r) {
            Plan P;
            std::vector<size_t> pick(R.size());
            for (size_t i = 0; i < R.size(); ++i) {
                pick[i] = std::min(r, cand[i].size() - 1);
                P.launches.push_back(group_launch(E, R[i], cand[i][pick[i]], cand[i][pick[i]]));
            }
            std::vector<float> ms;
            float phase = 0.0f;
            if ((rc = tune_cycle(E, measure, P, ms, phase))) return rc;
            done += TUNE_CYCLES;
            for (size_t i = 0; i < R.size(); ++i) best[i][pick[i]] = std::min(best[i][pick[i]], ms[i]);
            if (log)
                for (size_t i = 0; i < R.size(); ++i)
                    std::fprintf(stderr, "j1j2 tune: rep %d launch %s: %.4f ms\n", rep,
                                 P.launches[i].nt ? ("L" + std::to_string(P.launches[i].L) + " nt" +
                                                     std::to_string(P.launches[i].nt)).c_str()
                                                  : ("L" + std::to_string(P.launches[i].L) + " s" +
                                                     std::to_string(P.launches[i].slabs)).c_str(),
                                 ms[i]);
        }
    Plan per;
    TuneChoice per_choice;
    float t_seq = 0.0f;
    // a candidate within TUNE_TIE of the fastest counts as a tie and the
    // earlier one (fewer threads) is kept, so that the choice is stable
    for (size_t i = 0; i < R.size(); ++i) {
        const float tmin = *std::min_element(best[i].begin(), best[i].end());
        size_t b = 0;
        while (best[i][b] > tmin * TUNE_TIE) ++b;
        per.launches.push_back(group_launch(E, R[i], cand[i][b], cand[i][b]));
        per_choice.pick.push_back(cand[i][b]);
        t_seq += best[i][b];
    }
    Plan choice = per;
    TuneChoice tc = per_choice;
    float t_best = t_seq;
    if (multi) {
        // 2. the same launches on parallel branches, and one launch over all
        std::vector<Plan> alt;
        std::vector<TuneChoice> alt_choice;
        Plan conc = per;
        conc.concurrent = true;
        // slab launches need all their blocks resident at once: not beside others
        if (std::none_of(per.launches.begin(), per.launches.end(), [](const ResLaunch &r) { return r.slabs > 1; })) {
            alt.push_back(conc);
            alt_choice.push_back(per_choice);
            alt_choice.back().concurrent = true;
        }
        if (E.all_one)
            for (int32_t nt : ALL_NT) {
                Plan one;
                one.launches.push_back(all_launch(E, nt));
                alt.push_back(one);
                TuneChoice c;
                c.all_nt = nt;
                alt_choice.push_back(c);
            }
        for (size_t a = 0; a < alt.size(); ++a) {
            const Plan &P = alt[a];
            float tmin = 1e30f;
            for (int rep = 0; rep < TUNE_REPS; ++rep) {
                std::vector<float> ms;
                float phase = 0.0f;
                if ((rc = tune_cycle(E, measure, P, ms, phase))) return rc;
                done += TUNE_CYCLES;
                tmin = std::min(tmin, phase);
                if (log) std::fprintf(stderr, "j1j2 tune: rep %d %s: %.4f ms\n", rep, P.describe().c_str(), phase);
            }
            if (tmin * TUNE_TIE < t_best) {
                t_best = tmin;
                choice = P;
                tc = alt_choice[a];
            }
        }
    }
    E.plan = choice;
    E.tuned = true;
    drop_graphs(E);
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        tune_cache()[E.tune_key] = tc;
    }
    return ensure_side(E, choice.concurrent ? choice.launches.size() : 0);
}

// FP32 rule: |y| = |b2 h| must stay far inside the int range of rint(y)
// (threshold32 has no clamp); 2^24 leaves every threshold exactly 0 or inf.
int check_temperatures(const Engine &E)
{
    if (E.common.precision == J1J2_FP32_BATCHED)
        for (int c = 0; c < (int)E.confs.size(); ++c)
            for (double t : E.temps[c])
                if (2.0 * LOG2E / t * E.hbound[c] >= 0x1p24)
                    return fail(J1J2_EINVAL, "temperature too low for the FP32 form (|beta dE| >= 2^24 ln 2)");
    return 0;
}

int run_cycles(Engine &E, int phase, int ncycles, int begin)
{
    if (phase != J1J2_PHASE_EQUIL && phase != J1J2_PHASE_MEAS) return fail(J1J2_EINVAL, "unknown phase");
    const int measure = phase == J1J2_PHASE_MEAS ? 1 : 0;
    int rc;
    if ((rc = check_temperatures(E))) return rc;
    if (begin) {
        HIPCHK(launch_begin_phase(E.st, E.stream));
        if (measure)
            for (int c = 0; c < (int)E.confs.size(); ++c)
                if ((rc = reset_measurement(E, c))) return rc;
    }
    int done = 0;
    if (!E.tuned && E.autotune && !E.timing && (rc = tune(E, measure, ncycles, done))) return rc;
    const bool graph = E.use_graph && !E.timing;
    if (graph) {
        for (int lg = GRAPH_SIZES - 1; lg >= 0; --lg) {
            const int n = 1 << lg;
            if (ncycles - done < n) continue;
            if ((rc = build_graph(E, measure, lg))) return rc;
            do {                                  // 16 as often as it fits, then each smaller size at most once
                HIPCHK(hipGraphLaunch(E.graph[measure][lg], E.stream));
                done += n;
            } while (lg == GRAPH_SIZES - 1 && ncycles - done >= n);
        }
    }
    for (; done < ncycles; ++done)
        if ((rc = enqueue_cycle(E, measure))) return rc;
    E.cycles_done += (uint64_t)ncycles;
    if (E.st.fuse_locnem && ncycles > 0) {
        // accumulate the state the last measuring cycle left behind
        HIPCHK(launch_locnem(E.st, E.max_n, 1, E.stream));
        HIPCHK(hipMemsetAsync(E.st.locnem_pending, 0, sizeof(int32_t) * E.confs.size(), E.stream));
    }
    HIPCHK(hipStreamSynchronize(E.stream));
    if (E.st.slab_err) {
        int32_t err = 0;
        HIPCHK(hipMemcpy(&err, E.st.slab_err, sizeof(err), hipMemcpyDeviceToHost));
        if (err)
            return fail(J1J2_EDEVICE, "bit-layout slabs: a wait for a neighbouring slab ran out (slabs not resident together)");
    }
    if (E.timing && (rc = collect_timers(E))) return rc;
    if (ncycles > 0) E.observed = true;
    return 0;
}

// ---------------------------------------------------------------- no-alloc cycles
// reset_measurement for every config, by fill kernels on stream q
int reset_measurement_fill(const Engine &E, hipStream_t q)
{
    const DeviceState &st = E.st;
    const int64_t nconf = (int64_t)E.confs.size();
    uint64_t dmax;
    const double big = 1.7976931348623157e308;
    std::memcpy(&dmax, &big, sizeof(dmax));
    HIPCHK(launch_fill_u64(st.sums, E.nslots * NSUMS, 0, q));
    HIPCHK(launch_fill_u64(st.best, E.nslots, dmax, q));
    HIPCHK(launch_fill_u64(st.final_en, E.nslots, 0, q));
    HIPCHK(launch_fill_u64(st.hist, E.nhist, 0, q));
    HIPCHK(launch_fill_u64(st.hmin, E.nslots, 0, q));
    HIPCHK(launch_fill_u64(st.hmax, E.nslots, 0, q));
    HIPCHK(launch_fill_u64(st.hbin, E.nslots, 0, q));
    HIPCHK(launch_fill_u8(st.vcold, E.nslots, 0, q));
    HIPCHK(launch_fill_u8(st.vhot, E.nslots, 0, q));
    HIPCHK(launch_fill_u64(st.nmeasure, nconf, 0, q));
    HIPCHK(launch_fill_u8(st.locspin, E.nsites * (int64_t)sizeof(int32_t), 0, q));
    HIPCHK(launch_fill_u8(st.locnem, E.nsites * (int64_t)sizeof(int32_t), 0, q));
    HIPCHK(launch_fill_u64(st.locsamples, nconf, 0, q));
    HIPCHK(launch_fill_u64(st.series_len, nconf, 0, q));
    return 0;
}

// One cycle, eagerly on stream q with device state st: the plan's resident
// launches one after another (a cycle's result does not depend on the plan),
// the global-memory passes, the swap step.  No events, no side streams.
int enqueue_cycle_noalloc(const Engine &E, const DeviceState &st, hipStream_t q, bool measure)
{
    int rc;
    if ((rc = check_plan(E, E.plan))) return rc;
    for (const ResLaunch &r : E.plan.launches)
        HIPCHK(r.kind == 2 ? launch_sweep_bits(st, r.ord0, r.nord, r.L, r.H, r.slabs, E.shadow, q)
                           : launch_sweep_resident(st, r.ord0, r.nord, r.max_n, r.kernel_L(), r.H, r.nt, E.shadow,
                                                   q));
    for (const Group &G : E.groups) {
        if (G.resident) continue;
        for (int k = 0; k < E.common.ptswapever; ++k)
            for (int c = 0; c < 4; ++c)
                HIPCHK(E.common.precision == J1J2_FP64_PLAIN
                           ? launch_sweep_plain64(st, G.lat0, G.nlat, G.p.L, G.p.H, k, c, q)
                           : launch_sweep_global(st, G.lat0, G.nlat, G.p.L, G.p.H, k, c, E.shadow, q));
        HIPCHK(launch_observables(st, G.lat0, G.nlat, q));
    }
    HIPCHK(launch_pt_step(st, measure ? 1 : 0, q));
    if (measure && E.locnem && !st.fuse_locnem) HIPCHK(launch_locnem(st, E.max_n, 0, q));
    return 0;
}

int run_cycles_noalloc(Engine &E, int phase, int ncycles, int begin, hipStream_t q)
{
    if (phase != J1J2_PHASE_EQUIL && phase != J1J2_PHASE_MEAS) return fail(J1J2_EINVAL, "unknown phase");
    if (E.st.nlat > NOALLOC_MAX_LATTICES)
        return fail(J1J2_EINVAL, "j1j2_run_cycles_noalloc: the engine has " + std::to_string(E.st.nlat) +
                                     " lattices; its static scratch holds at most " +
                                     std::to_string(NOALLOC_MAX_LATTICES));
    int64_t *scr = noalloc_scratch(E.device);
    if (!scr) return fail(J1J2_ESTATE, "j1j2_run_cycles_noalloc: scratch address not set up for this device");
    const bool measure = phase == J1J2_PHASE_MEAS;
    int rc;
    if ((rc = check_temperatures(E))) return rc;
    DeviceState st = E.st;
    st.obs = scr;
    if (begin) {
        HIPCHK(launch_begin_phase(st, q));
        if (measure && (rc = reset_measurement_fill(E, q))) return rc;
    }
    for (int i = 0; i < ncycles; ++i)
        if ((rc = enqueue_cycle_noalloc(E, st, q, measure))) return rc;
    E.cycles_done += (uint64_t)ncycles;
    if (st.fuse_locnem && ncycles > 0) {
        HIPCHK(launch_locnem(st, E.max_n, 1, q));
        HIPCHK(launch_fill_u8(st.locnem_pending, (int64_t)E.confs.size() * (int64_t)sizeof(int32_t), 0, q));
    }
    if (ncycles > 0) E.observed = true;
    return 0;
}

// Observables of the current state without advancing it (NMESS <= 0 branch
// of Pt_measurement_loop, J1J2_6tv.f90:1414-1424).
int refresh_values(Engine &E)
{
    int rc;
    if ((rc = enqueue_cycle(E, 2))) return rc;
    HIPCHK(hipStreamSynchronize(E.stream));
    E.observed = true;
    return 0;
}

Engine *as_engine(void *h) { return static_cast<Engine *>(h); }

int check_conf(Engine *E, int32_t conf)
{
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (conf < 0 || conf >= (int32_t)E->confs.size()) return fail(J1J2_EINVAL, "config index out of range");
    return 0;
}

void to_internal(int L, int H, const int8_t *ref, uint8_t *out)
{
    for (int z = 0; z < H; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) out[internal_index(L, H, x, y, z)] = ref[ref_site(L, x, y, z)] > 0 ? 1 : 0;
}

template <typename S, typename D>
void from_internal(int L, int H, const S *in, D *ref, D (*conv)(S))
{
    for (int z = 0; z < H; ++z)
        for (int y = 0; y < L; ++y)
            for (int x = 0; x < L; ++x) ref[ref_site(L, x, y, z)] = conv(in[internal_index(L, H, x, y, z)]);
}

// Bytes of a lattice's spins in HBM: one per site, one bit per site (packed
// 1), or the bit-layout sweep's two dwords per half-row unit (packed 2, see
// unit_word_bit).
int64_t spin_layout_bytes(const LatDesc &d)
{
    if (d.packed == 2) return 8 * (int64_t)d.N / (d.L / 2);
    return d.packed ? d.N / 8 : d.N;
}

// Internal-order spin bytes (0/1) -> the lattice's packed form.  Unit words
// also carry each dword's copies of the sites just outside it.
void pack_spins(const LatDesc &d, const uint8_t *in, std::vector<uint8_t> &out)
{
    out.assign((size_t)spin_layout_bytes(d), 0);
    if (d.packed == 1) {                              // site i at bit i & 7 of byte i >> 3
        for (int64_t i = 0; i < d.N; ++i) out[i >> 3] |= (uint8_t)(in[i] << (i & 7));
        return;
    }
    const int HW = d.L / 2, N0 = 4 * ((HW / 4 + 1) / 2), N1 = HW - N0;
    const int64_t nu = (int64_t)d.N / HW;
    std::vector<uint32_t> w((size_t)(2 * nu));
    for (int64_t u = 0; u < nu; ++u) {
        const uint8_t *s = in + u * HW;
        uint32_t a = (uint32_t)s[HW - 1] | ((uint32_t)s[N0] << (N0 + 1));
        uint32_t b = (uint32_t)s[N0 - 1] | ((uint32_t)s[0] << (N1 + 1));
        for (int m = 0; m < N0; ++m) a |= (uint32_t)s[m] << (m + 1);
        for (int m = N0; m < HW; ++m) b |= (uint32_t)s[m] << (m - N0 + 1);
        w[2 * u] = a;
        w[2 * u + 1] = b;
    }
    std::memcpy(out.data(), w.data(), out.size());
}

// The lattice's packed form -> internal-order spin bytes (0/1).
void unpack_spins(const LatDesc &d, const uint8_t *in, uint8_t *out)
{
    if (d.packed == 1) {
        for (int64_t i = 0; i < d.N; ++i) out[i] = (in[i >> 3] >> (i & 7)) & 1;
        return;
    }
    for (int64_t i = 0; i < d.N; ++i) {
        int64_t dw;
        int bit;
        unit_word_bit(d.L, i, dw, bit);
        uint32_t v;
        std::memcpy(&v, in + 4 * dw, 4);
        out[i] = (uint8_t)((v >> bit) & 1u);
    }
}

int create(const j1j2_params *groups, int32_t ngroups, int32_t device, void **handle)
{
    if (!handle) return fail(J1J2_EINVAL, "null handle pointer");
    *handle = nullptr;
    if (!groups || ngroups < 1) return fail(J1J2_EINVAL, "need at least one parameter group");
    int rc;
    for (int g = 0; g < ngroups; ++g)
        if ((rc = validate_group(groups[g]))) return rc;
    const j1j2_params &c0 = groups[0];
    for (int g = 1; g < ngroups; ++g) {
        const j1j2_params &p = groups[g];
        if (p.ptswapever != c0.ptswapever || p.neq != c0.neq || p.nmess != c0.nmess || p.bins != c0.bins ||
            p.sampler != c0.sampler || p.precision != c0.precision || p.flags != c0.flags ||
            p.ptmeasindex != c0.ptmeasindex || p.seed != c0.seed)
            return fail(J1J2_EINVAL, "groups must share PTSWAPEVER, PTMEASINDEX, NEQ, NMESS, BINS, sampler, "
                                     "precision, flags and seed");
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(J1J2_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    HIPCHK(noalloc_setup(device));    // address of the no-alloc scratch on this device, once

    Engine *E = new Engine();
    std::unique_ptr<Engine> guard(E);
    E->device = device;
    E->common = c0;
    HIPCHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    E->shadow = (c0.flags & J1J2_FLAG_SHADOW_FP64) && c0.precision == J1J2_FP32_BATCHED;
    E->locnem = (c0.flags & J1J2_FLAG_WRITELOCNEM) != 0;
    E->use_graph = !(c0.flags & J1J2_FLAG_NO_GRAPH);
    const bool flips = (c0.flags & J1J2_FLAG_WRITENUMFLIPS) != 0;
    const size_t lds_limit = resident_lds_limit();

    int64_t nlat = 0;
    int64_t series_cap = c0.sampler == J1J2_SAMPLER_LOCAL ? std::max(c0.nmess, 1) : 0;
    E->all_resident = true;
    E->all_byte = true;
    for (int g = 0; g < ngroups; ++g) {
        const j1j2_params &p = groups[g];
        Group G;
        G.p = p;
        G.N = p.L * p.L * p.H;
        G.conf0 = (int32_t)E->confs.size();
        G.lat0 = (int32_t)nlat;
        G.nlat = p.nconf * p.nreplica;
        // flip counts (WRITENUMFLIPS) keep a byte per site in LDS beside the
        // byte layout (counts per launch <= PTSWAPEVER); the bit layout has no
        // room for them, so such lattices take the global-memory sweep
        const bool may = p.precision == J1J2_FP32_BATCHED && !(p.flags & J1J2_FLAG_FORCE_GLOBAL);
        // (cubic L = 48 keeps its counts in the sites' own bytes instead)
        const bool own_counts = byte_counts(p.L) && fast_counts_ok(p.L, p.H == p.L,
                                                                   (p.flags & J1J2_FLAG_SHADOW_FP64) != 0, p.ptswapever);
        const size_t res_need = (size_t)G.N * (flips && !own_counts ? 2 : 1);
        if (may && resident_supported(p.L) && res_need <= lds_limit && (!flips || p.ptswapever <= 255)) G.kind = 1;
        else if (may && !flips && bits_supported(p.L) && bits_lds_need(p.L, p.H) <= lds_limit + RESIDENT_SCRATCH_LDS)
            G.kind = 2;
        G.resident = G.kind != 0;
        E->all_resident = E->all_resident && G.resident;
        E->all_byte = E->all_byte && G.kind == 1;
        E->max_n = std::max(E->max_n, G.N);
        std::vector<double> t(p.nreplica);
        temperature_ladder(p.ladder, p.nreplica, p.tmin, p.tmax, p.dt, p.tladdermid, p.tladderfrac, t.data());
        for (int i = 0; i < p.nconf; ++i) {
            ConfDesc cd{};
            cd.L = p.L; cd.H = p.H; cd.N = G.N; cd.nrep = p.nreplica;
            cd.lat0 = (int32_t)nlat;
            cd.slot_base = (int32_t)E->nslots;
            cd.pair_base = (int32_t)E->npairs;
            cd.iconf = p.iconf0 + i;
            cd.site_off = E->nsites;
            cd.hist_off = E->nhist;
            cd.series_off = E->nseries;
            cd.flip_off = E->nflips;
            ConfConst cc{};
            cc.J1h = p.J1 + p.deltaJ1; cc.J1v = p.J1 - p.deltaJ1; cc.J2 = p.J2; cc.J3 = p.J3;
            cc.j2x2 = (float)(2.0 * p.J2); cc.j3x2 = (float)(2.0 * p.J3);
            for (int r = 0; r < p.nreplica; ++r) {
                LatDesc d{};
                d.L = p.L; d.H = p.H; d.N = G.N; d.conf = (int32_t)E->confs.size(); d.rep = r;
                d.slot_base = cd.slot_base; d.spin_off = E->spin_bytes; d.site_off = cd.site_off;
                // the byte-layout resident sweep keeps one bit per site, the
                // bit-layout sweep its unit words
                d.packed = G.kind == 1 ? 1 : (G.kind == 2 ? 2 : 0);
                E->lats.push_back(d);
                E->spin_bytes += (spin_layout_bytes(d) + 255) / 256 * 256;
                ++nlat;
            }
            E->nslots += p.nreplica;
            E->npairs += std::max(p.nreplica - 1, 0);
            E->nsites += ((int64_t)G.N + 63) / 64 * 64;
            E->nhist += (int64_t)p.nreplica * p.bins;
            E->nseries += (int64_t)p.nreplica * 5 * series_cap;
            if (flips) E->nflips += (int64_t)p.nreplica * G.N;
            E->confs.push_back(cd);
            E->cconst.push_back(cc);
            E->temps.push_back(t);
            E->hbound.push_back(12.0 * std::max(std::fabs(cc.J1h), std::fabs(cc.J1v)) + 12.0 * std::fabs(cc.J2) +
                                6.0 * std::fabs(cc.J3) + 1.0);
        }
        E->groups.push_back(G);
    }
    E->path = c0.precision == J1J2_FP64_PLAIN ? 0 : (E->all_byte ? 1 : (E->all_resident ? 3 : 2));
    // the launch over all sizes keeps flip counts beside each lattice
    E->all_one = E->all_byte && (!flips || 2 * (size_t)E->max_n <= lds_limit);
    E->uniform_L = c0.L;
    E->uniform_H = c0.H;
    E->all_cubic = true;
    for (const Group &G : E->groups) {
        if (G.p.L != c0.L) E->uniform_L = 0;
        if (G.p.H != c0.H) E->uniform_H = 0;
        if (G.p.H != G.p.L) E->all_cubic = false;
    }
    if (const char *nt = std::getenv("J1J2_RES_THREADS")) {
        // a fixed block size for every resident launch, no autotuning
        const int v = std::atoi(nt);
        E->res_nt = (v == 512 || v == 576 || v == 1024) ? v : 256;
        E->autotune = false;
    }
    if (const char *sl = std::getenv("J1J2_BIT_SLABS")) {
        // a fixed slab count for the bit-layout groups (where valid), no autotuning
        E->bit_slabs = std::atoi(sl);
        E->autotune = false;
    }
    if (const char *at = std::getenv("J1J2_AUTOTUNE")) E->autotune = E->autotune && std::atoi(at) != 0;
    size_t byte_lds = 0;                           // the byte-layout launches' largest lattice (and counts)
    for (const Group &G : E->groups)
        if (G.kind == 1) byte_lds = std::max(byte_lds, std::min((size_t)G.N * (flips ? 2 : 1), lds_limit));
    if (byte_lds) HIPCHK(setup_resident(byte_lds));
    size_t bits_lds = 0;
    for (const Group &G : E->groups)
        if (G.kind == 2) bits_lds = std::max(bits_lds, bits_lds_need(G.p.L, G.p.H));
    if (bits_lds) HIPCHK(setup_bits(bits_lds));

    DeviceState &st = E->st;
    const int64_t nconf = (int64_t)E->confs.size();
    LatDesc *lats; ConfDesc *confs; ConfConst *cconst;
    float *kd, *b2; double *p64, *beta; uint8_t *q;
    if ((rc = E->alloc(lats, nlat)) || (rc = E->alloc(confs, nconf)) || (rc = E->alloc(cconst, nconf)) ||
        (rc = E->alloc(st.spins, E->spin_bytes)) || (rc = E->alloc(kd, 5 * E->nsites)) ||
        (rc = E->alloc(p64, E->nsites)) || (rc = E->alloc(q, 5 * E->nsites)) ||
        (rc = E->alloc(st.obs, nlat * OBS_N)) || (rc = E->alloc(st.rep_at_temp, E->nslots)) ||
        (rc = E->alloc(st.temp_of_rep, E->nslots)) || (rc = E->alloc(beta, E->nslots)) ||
        (rc = E->alloc(b2, E->nslots)) || (rc = E->alloc(st.clk, 1)) ||
        (rc = E->alloc(st.sums, E->nslots * NSUMS)) || (rc = E->alloc(st.best, E->nslots)) ||
        (rc = E->alloc(st.final_en, E->nslots)) || (rc = E->alloc(st.hist, E->nhist)) ||
        (rc = E->alloc(st.hmin, E->nslots)) || (rc = E->alloc(st.hmax, E->nslots)) ||
        (rc = E->alloc(st.hbin, E->nslots)) || (rc = E->alloc(st.att, std::max<int64_t>(E->npairs, 1))) ||
        (rc = E->alloc(st.acc, std::max<int64_t>(E->npairs, 1))) || (rc = E->alloc(st.vcold, E->nslots)) ||
        (rc = E->alloc(st.vhot, E->nslots)) || (rc = E->alloc(st.nmeasure, nconf)) ||
        (rc = E->alloc(st.etot, E->nslots)) || (rc = E->alloc(st.vals, E->nslots * 5)) ||
        (rc = E->alloc(st.locspin, E->nsites)) || (rc = E->alloc(st.locnem, E->nsites)) ||
        (rc = E->alloc(st.locsamples, nconf)) || (rc = E->alloc(st.series, E->nseries)) ||
        (rc = E->alloc(st.series_len, nconf)) || (rc = E->alloc(st.shadow, 2 * nconf)) ||
        (rc = E->alloc(st.locnem_pending, nconf)))
        return rc;
    st.flips = nullptr;
    st.flips_pend = nullptr;
    if (flips && ((rc = E->alloc(st.flips, E->nflips)) || (rc = E->alloc(st.flips_pend, E->nflips)))) return rc;
    st.slab_x = st.slab_obs = nullptr;
    st.slab_err = nullptr;
    st.slab_spin_max = 1 << 21;                        // about seconds of polling
    if (const char *sp = std::getenv("J1J2_SLAB_SPIN_MAX")) st.slab_spin_max = std::atoi(sp);
    if (bits_lds &&
        ((rc = E->alloc(st.slab_x, (int64_t)bits_slab_x_words((int32_t)nlat))) ||
         (rc = E->alloc(st.slab_obs, nlat * BIT_SLAB_MAX * 16)) || (rc = E->alloc(st.slab_err, 1))))
        return rc;
    st.lats = lats; st.confs = confs; st.cconst = cconst;
    st.nlat = (int32_t)nlat; st.nconf = (int32_t)nconf; st.nsites = E->nsites;
    st.uni_nrep = c0.nreplica;
    st.uni_n = E->groups[0].N;
    for (const Group &G : E->groups)
        if (G.p.nreplica != c0.nreplica || G.N != E->groups[0].N) st.uni_nrep = st.uni_n = 0;
    st.kd = kd; st.p64 = p64; st.qb = q; st.beta = beta; st.b2 = b2;
    st.series_cap = series_cap;
    st.ptswapever = c0.ptswapever; st.ptmeasindex = c0.ptmeasindex; st.bins = c0.bins;
    st.local = c0.sampler == J1J2_SAMPLER_LOCAL ? 1 : 0;
    st.seed = (uint32_t)c0.seed;
    st.fuse_locnem = (E->locnem && E->all_resident) ? 1 : 0;
    st.locsnap = nullptr;
    if (st.fuse_locnem && std::any_of(E->groups.begin(), E->groups.end(), [](const Group &G) { return G.kind == 1; }) &&
        (rc = E->alloc(st.locsnap, E->nsites / 32)))
        return rc;
    {
        // work lists of the resident launches: per resident group, and over all lattices
        std::vector<int32_t> order;
        E->group_seg.assign(E->groups.size(), OrderSeg{});
        auto add = [&](int c0, int c1) {
            const std::vector<int32_t> seg = make_order(*E, c0, c1);
            OrderSeg o;
            o.ord0 = (int32_t)order.size();
            o.nord = (int32_t)seg.size();
            order.insert(order.end(), seg.begin(), seg.end());
            return o;
        };
        auto add_slabs = [&](int c0, int c1, int S) {
            const std::vector<int32_t> seg = make_order(*E, c0, c1, S);
            OrderSeg o;
            o.ord0 = (int32_t)order.size();
            o.nord = (int32_t)seg.size();
            order.insert(order.end(), seg.begin(), seg.end());
            return o;
        };
        for (int g = 0; g < (int)E->groups.size(); ++g) {
            Group &G = E->groups[g];
            if (G.resident) E->group_seg[g] = add(G.conf0, G.conf0 + G.p.nconf);
            if (G.kind == 2)
                for (int S : {2, 4})
                    if (bits_slab_valid(G.p.L, G.p.H, S)) G.slab_seg[S] = add_slabs(G.conf0, G.conf0 + G.p.nconf, S);
        }
        if (E->all_one) E->all_seg = add(0, (int)nconf);
        // the same in bundles for the 1024-thread launch (FP32 without the shadow
        // rule), when some size bundles and every lattice index fits an entry
        bool bundles = false;
        for (const Group &G : E->groups) bundles = bundles || res_bundle(G.p.L) > 1;
        if (E->all_one && !E->shadow && bundles && nlat <= BUNDLE_LAT_MASK) {
            const std::vector<int32_t> seg = make_order(*E, 0, (int)nconf, 1, true);
            E->all_seg_b.ord0 = (int32_t)order.size();
            E->all_seg_b.nord = (int32_t)seg.size();
            order.insert(order.end(), seg.begin(), seg.end());
            for (const Group &G : E->groups) E->all_max_b = std::max(E->all_max_b, res_bundle(G.p.L) * G.N);
            if ((size_t)E->all_max_b * (flips ? 2 : 1) > lds_limit) E->all_seg_b = OrderSeg{};   // no room
        }
        // a forced slab count (2 or 4) only where it applies to every bit-layout group
        if (E->bit_slabs != 2 && E->bit_slabs != 4) E->bit_slabs = 1;
        for (const Group &G : E->groups)
            if (G.kind == 2 && E->bit_slabs > 1 &&
                !bits_slab_fits(G.p.L, G.p.H, E->bit_slabs, G.slab_seg[E->bit_slabs].nord, E->shadow))
                E->bit_slabs = 1;
        int32_t *dorder;
        E->order_len = (int32_t)order.size();
        if ((rc = E->alloc(dorder, (int64_t)order.size()))) return rc;
        if (!order.empty() && (rc = h2d(*E, dorder, order))) return rc;
        st.order = dorder;
    }
    E->plan = default_plan(*E);
    if ((rc = h2d(*E, lats, E->lats)) || (rc = h2d(*E, confs, E->confs)) || (rc = h2d(*E, cconst, E->cconst)))
        return rc;
    std::vector<double> zero_p((size_t)E->max_n, 0.0);
    for (int c = 0; c < (int)nconf; ++c) {
        if ((rc = set_ladder(*E, c, E->temps[c].data()))) return rc;
        if ((rc = upload_site_planes(*E, c, zero_p.data()))) return rc;
        if ((rc = reset_config(*E, c))) return rc;
    }
    HIPCHK(hipStreamSynchronize(E->stream));
    *handle = guard.release();
    return 0;
}

}  // namespace

// ======================================================================
// exported C-ABI
// ======================================================================
extern "C" {

int j1j2_last_error(char *buf, int32_t len)
{
    if (!buf || len <= 0) return J1J2_EINVAL;
    std::strncpy(buf, g_err.c_str(), (size_t)len - 1);
    buf[len - 1] = '\0';
    return 0;
}

int j1j2_ladder(const j1j2_params *p, double *temps)
{
    if (!p || !temps) return fail(J1J2_EINVAL, "null argument");
    int rc;
    if ((rc = validate_group(*p))) return rc;
    temperature_ladder(p->ladder, p->nreplica, p->tmin, p->tmax, p->dt, p->tladdermid, p->tladderfrac, temps);
    return 0;
}

int j1j2_make_disorder(int32_t L, int32_t H, double xi, int32_t seed, int32_t map_to_uniform, double phi,
                       double *out)
{
    if (!out || L < 1 || H < 1) return fail(J1J2_EINVAL, "invalid disorder request");
    return corr_field(L, L, H, xi, seed, map_to_uniform != 0, phi, out);
}

int j1j2_make_disorder_batch(int32_t L, int32_t H, double xi, int32_t seed0, int32_t nconf, int32_t map_to_uniform,
                             double phi, int32_t nthreads, double *out)
{
    if (!out || L < 1 || H < 1 || nconf < 0) return fail(J1J2_EINVAL, "invalid disorder request");
    const int64_t N = (int64_t)L * L * H;
    int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, nconf));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k; (k = next.fetch_add(1)) < nconf;)
            (void)corr_field(L, L, H, xi, seed0 + k, map_to_uniform != 0, phi, out + (int64_t)k * N);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    return 0;
}

int j1j2_initial_spins(int32_t L, int32_t H, int32_t start, int32_t straxis, int32_t seed, int64_t skip,
                       int8_t *spins)
{
    if (!spins || L < 1 || H < 1 || start < 0 || start > 2 || straxis < 0 || straxis > 1 || skip < 0)
        return fail(J1J2_EINVAL, "invalid initial-state request");
    initial_spins(L, H, start, straxis, seed, skip, spins);
    return 0;
}

int j1j2_create(const j1j2_params *p, int32_t device, void **handle) { return create(p, 1, device, handle); }

int j1j2_create_batch(const j1j2_params *groups, int32_t ngroups, int32_t device, void **handle)
{
    return create(groups, ngroups, device, handle);
}

void j1j2_destroy(void *handle) { delete as_engine(handle); }

int j1j2_num_configs(void *handle, int32_t *n)
{
    Engine *E = as_engine(handle);
    if (!E || !n) return fail(J1J2_EINVAL, "null argument");
    *n = (int32_t)E->confs.size();
    return 0;
}

int j1j2_config_shape(void *handle, int32_t conf, int32_t *L, int32_t *H, int32_t *nrep)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    if (L) *L = cd.L;
    if (H) *H = cd.H;
    if (nrep) *nrep = cd.nrep;
    return 0;
}

int j1j2_upload_disorder(void *handle, int32_t conf, const double *p, int64_t n)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    if (!p || n != cd.N) return fail(J1J2_EINVAL, "disorder array must hold L*L*H values");
    double pmax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(std::fabs(p[i]) < 64.0)) return fail(J1J2_EINVAL, "disorder values must be finite with |p| < 64");
        pmax = std::max(pmax, std::fabs(p[i]));
    }
    // |h| <= |C| + 2 sum|K| + 8|J2| + 4|J3| with |C| <= sum|K| + 4|J2| + 2|J3|, |K| <= |J1| + 2 pmax
    const ConfConst &cc = E->cconst[conf];
    const double kmax = std::max(std::fabs(cc.J1h), std::fabs(cc.J1v)) + 2.0 * pmax;
    E->hbound[conf] = 12.0 * kmax + 12.0 * std::fabs(cc.J2) + 6.0 * std::fabs(cc.J3) + 1.0;
    HIPCHK(hipSetDevice(E->device));
    E->observed = false;
    return upload_site_planes(*E, conf, p);
}

int j1j2_set_spins(void *handle, int32_t conf, const int8_t *spins, int32_t nrep)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    if (!spins || nrep != cd.nrep) return fail(J1J2_EINVAL, "spins must hold NREPLICA*L*L*H values");
    for (int64_t i = 0; i < (int64_t)nrep * cd.N; ++i)
        if (spins[i] != 1 && spins[i] != -1) return fail(J1J2_EINVAL, "spins must be +1 or -1");
    HIPCHK(hipSetDevice(E->device));
    std::vector<uint8_t> buf(cd.N), packed;
    for (int r = 0; r < nrep; ++r) {
        const LatDesc &d = E->lats[cd.lat0 + r];
        to_internal(cd.L, cd.H, spins + (int64_t)r * cd.N, buf.data());
        const uint8_t *src = buf.data();
        if (d.packed) {
            pack_spins(d, buf.data(), packed);
            src = packed.data();
        }
        HIPCHK(hipMemcpyAsync(E->st.spins + d.spin_off, src, spin_layout_bytes(d), hipMemcpyHostToDevice,
                              E->stream));
        HIPCHK(hipStreamSynchronize(E->stream));
    }
    if ((rc = reset_config(*E, conf))) return rc;
    HIPCHK(hipStreamSynchronize(E->stream));
    E->observed = false;
    return 0;
}

int j1j2_set_temperatures(void *handle, int32_t conf, const double *temps, int32_t nrep)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    if (!temps || nrep != E->confs[conf].nrep) return fail(J1J2_EINVAL, "need NREPLICA temperatures");
    HIPCHK(hipSetDevice(E->device));
    return set_ladder(*E, conf, temps);
}

int j1j2_run_cycles(void *handle, int32_t phase, int32_t ncycles, int32_t begin)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (ncycles < 0) return fail(J1J2_EINVAL, "ncycles must be >= 0");
    HIPCHK(hipSetDevice(E->device));
    return run_cycles(*E, phase, ncycles, begin);
}

int j1j2_run_cycles_noalloc(void *handle, int32_t phase, int32_t ncycles, int32_t begin, void *stream)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (ncycles < 0) return fail(J1J2_EINVAL, "ncycles must be >= 0");
    HIPCHK(hipSetDevice(E->device));
    return run_cycles_noalloc(*E, phase, ncycles, begin, stream ? static_cast<hipStream_t>(stream) : E->stream);
}

int j1j2_run(void *handle, int32_t phase, int32_t ncycles)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (ncycles < 0) ncycles = phase == J1J2_PHASE_MEAS ? E->common.nmess : E->common.neq;
    HIPCHK(hipSetDevice(E->device));
    int rc = run_cycles(*E, phase, ncycles, 1);
    if (rc == 0 && phase == J1J2_PHASE_MEAS && ncycles == 0 && E->locnem) {
        // a measuring phase of no cycles still takes one local-nematic sample
        // of the current state (NMESS <= 0 branches, J1J2_6tv.f90:1414-1424
        // for PT, 2157-2160 for LOCAL)
        HIPCHK(launch_locnem(E->st, E->max_n, 0, E->stream));
        HIPCHK(hipStreamSynchronize(E->stream));
    }
    return rc;
}

int j1j2_get_spins(void *handle, int32_t conf, int8_t *spins, int32_t nrep)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    if (!spins || nrep != cd.nrep) return fail(J1J2_EINVAL, "spins must hold NREPLICA*L*L*H values");
    HIPCHK(hipSetDevice(E->device));
    std::vector<uint8_t> buf(cd.N), packed;
    for (int r = 0; r < nrep; ++r) {
        const LatDesc &d = E->lats[cd.lat0 + r];
        packed.assign((size_t)spin_layout_bytes(d), 0);
        HIPCHK(hipMemcpyAsync(d.packed ? packed.data() : buf.data(), E->st.spins + d.spin_off,
                              spin_layout_bytes(d), hipMemcpyDeviceToHost, E->stream));
        HIPCHK(hipStreamSynchronize(E->stream));
        if (d.packed) unpack_spins(d, packed.data(), buf.data());
        from_internal<uint8_t, int8_t>(cd.L, cd.H, buf.data(), spins + (int64_t)r * cd.N,
                                       [](uint8_t b) -> int8_t { return b ? 1 : -1; });
    }
    return 0;
}

int j1j2_get_maps(void *handle, int32_t conf, int32_t *rep_at_temp, int32_t *temp_of_rep)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    HIPCHK(hipSetDevice(E->device));
    if (rep_at_temp)
        HIPCHK(hipMemcpyAsync(rep_at_temp, E->st.rep_at_temp + cd.slot_base, sizeof(int32_t) * cd.nrep,
                              hipMemcpyDeviceToHost, E->stream));
    if (temp_of_rep)
        HIPCHK(hipMemcpyAsync(temp_of_rep, E->st.temp_of_rep + cd.slot_base, sizeof(int32_t) * cd.nrep,
                              hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return 0;
}

int j1j2_get_observables(void *handle, int32_t conf, double *out)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    if (!out) return fail(J1J2_EINVAL, "null output");
    HIPCHK(hipSetDevice(E->device));
    if (!E->observed && (rc = refresh_values(*E))) return rc;
    const ConfDesc &cd = E->confs[conf];
    std::vector<double> v;
    if ((rc = d2h(*E, v, E->st.vals + (int64_t)cd.slot_base * 5, (int64_t)cd.nrep * 5))) return rc;
    HIPCHK(hipStreamSynchronize(E->stream));
    for (int r = 0; r < cd.nrep; ++r)
        for (int q = 0; q < 5; ++q) out[(int64_t)q * cd.nrep + r] = v[(size_t)r * 5 + q];
    return 0;
}

int j1j2_get_sums(void *handle, int32_t conf, double *sums, int64_t *nmeasure, double *best, double *final_en)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    HIPCHK(hipSetDevice(E->device));
    if (sums)
        HIPCHK(hipMemcpyAsync(sums, E->st.sums + (int64_t)cd.slot_base * NSUMS, sizeof(double) * cd.nrep * NSUMS,
                              hipMemcpyDeviceToHost, E->stream));
    if (nmeasure)
        HIPCHK(hipMemcpyAsync(nmeasure, E->st.nmeasure + conf, sizeof(int64_t), hipMemcpyDeviceToHost, E->stream));
    if (best)
        HIPCHK(hipMemcpyAsync(best, E->st.best + cd.slot_base, sizeof(double) * cd.nrep, hipMemcpyDeviceToHost,
                              E->stream));
    if (final_en)
        HIPCHK(hipMemcpyAsync(final_en, E->st.final_en + cd.slot_base, sizeof(double) * cd.nrep,
                              hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return 0;
}

int j1j2_get_series(void *handle, int32_t conf, double *series, int64_t cap, int64_t *len)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    if (!E->st.local) return fail(J1J2_ESTATE, "series are kept by the LOCAL sampler only");
    const ConfDesc &cd = E->confs[conf];
    HIPCHK(hipSetDevice(E->device));
    int64_t n = 0;
    HIPCHK(hipMemcpyAsync(&n, E->st.series_len + conf, sizeof(int64_t), hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    n = std::min(n, E->st.series_cap);
    if (len) *len = n;
    if (!series) return 0;
    if (cap < n) return fail(J1J2_EINVAL, "series buffer too small");
    std::vector<double> buf;
    if ((rc = d2h(*E, buf, E->st.series + cd.series_off, (int64_t)cd.nrep * 5 * E->st.series_cap))) return rc;
    HIPCHK(hipStreamSynchronize(E->stream));
    for (int r = 0; r < cd.nrep; ++r)
        for (int q = 0; q < 5; ++q)
            std::memcpy(series + ((int64_t)r * 5 + q) * n, buf.data() + ((int64_t)r * 5 + q) * E->st.series_cap,
                        sizeof(double) * n);
    return 0;
}

int j1j2_fetch(void *handle, int32_t conf, double *basefields, double *hist, double *hist_min, double *hist_bin,
               int32_t *swap_att, int32_t *swap_acc, int32_t *visitboth, double *locspin, double *locnem,
               int32_t *locnem_samples)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    const ConfDesc &cd = E->confs[conf];
    const int nrep = cd.nrep, sb = cd.slot_base, bins = E->common.bins;
    HIPCHK(hipSetDevice(E->device));
    if (!E->observed && (rc = refresh_values(*E))) return rc;
    std::vector<double> sums, best, fin, hmin, hbin, vals, series;
    std::vector<int64_t> hcnt, att, acc, nm, ls;
    std::vector<uint8_t> vc, vh;
    std::vector<int32_t> lsp, lnm;
    if ((rc = d2h(*E, sums, E->st.sums + (int64_t)sb * NSUMS, (int64_t)nrep * NSUMS)) ||
        (rc = d2h(*E, best, E->st.best + sb, nrep)) || (rc = d2h(*E, fin, E->st.final_en + sb, nrep)) ||
        (rc = d2h(*E, hmin, E->st.hmin + sb, nrep)) || (rc = d2h(*E, hbin, E->st.hbin + sb, nrep)) ||
        (rc = d2h(*E, vals, E->st.vals + (int64_t)sb * 5, (int64_t)nrep * 5)) ||
        (rc = d2h(*E, hcnt, E->st.hist + cd.hist_off, (int64_t)nrep * bins)) ||
        (rc = d2h(*E, att, E->st.att + cd.pair_base, std::max(nrep - 1, 1))) ||
        (rc = d2h(*E, acc, E->st.acc + cd.pair_base, std::max(nrep - 1, 1))) ||
        (rc = d2h(*E, nm, E->st.nmeasure + conf, 1)) || (rc = d2h(*E, ls, E->st.locsamples + conf, 1)) ||
        (rc = d2h(*E, vc, E->st.vcold + sb, nrep)) || (rc = d2h(*E, vh, E->st.vhot + sb, nrep)) ||
        (rc = d2h(*E, lsp, E->st.locspin + cd.site_off, cd.N)) ||
        (rc = d2h(*E, lnm, E->st.locnem + cd.site_off, cd.N)))
        return rc;
    std::vector<int32_t> rat(nrep);
    HIPCHK(hipMemcpyAsync(rat.data(), E->st.rep_at_temp + sb, sizeof(int32_t) * nrep, hipMemcpyDeviceToHost,
                          E->stream));
    if (E->st.local && (rc = d2h(*E, series, E->st.series + cd.series_off, (int64_t)nrep * 5 * E->st.series_cap)))
        return rc;
    HIPCHK(hipStreamSynchronize(E->stream));
    const int64_t nmeas = nm[0];
    std::vector<double> hout((size_t)nrep * bins, 0.0);
    for (int s = 0; s < nrep; ++s) {
        double o[49];
        const double beta = 1.0 / E->temps[conf][s];
        if (E->st.local && nmeas > 0) {
            int64_t len = std::min<int64_t>(nmeas, E->st.series_cap);
            const double *sr = series.data() + (int64_t)s * 5 * E->st.series_cap;
            std::vector<double> h(bins);
            double hm = 0, hb = 0;
            process_local_series(sr, sr + E->st.series_cap, sr + 2 * E->st.series_cap, sr + 3 * E->st.series_cap,
                                 sr + 4 * E->st.series_cap, len, beta, cd.N, bins, o, h.data(), &hm, &hb);
            for (int b = 0; b < bins; ++b) hout[(size_t)b * nrep + s] = h[b];
            hmin[s] = hm;
            hbin[s] = hb;
        } else if (nmeas > 0) {
            derive_basefields(&sums[(size_t)s * NSUMS], nmeas, beta, cd.N, best[s], fin[s], o);
            for (int b = 0; b < bins; ++b)
                hout[(size_t)b * nrep + s] = (double)hcnt[(size_t)s * bins + b] / (double)nmeas;
        } else {
            // NMESS <= 0: instantaneous values of the replica at this slot (J1J2_6tv.f90:1545-1572)
            const double *v = &vals[(size_t)rat[s] * 5];
            const double am = v[0], ae = v[1], ax = v[2], ay = v[3], an = v[4];
            const double ax2 = ax * ax, ay2 = ay * ay, str2 = ax2 + ay2;
            const double S[NSUMS] = {am, std::fabs(am), am * am, (am * am) * (am * am), ae, ae * ae,
                                     std::fabs(ax), std::fabs(ay), ax2, ax2 * ax2, ay2, ay2 * ay2, str2,
                                     str2 * str2, std::sqrt(str2), std::fabs(ay2 - ax2), an, std::fabs(an),
                                     an * an, (an * an) * (an * an)};
            derive_basefields(S, 1, beta, cd.N, ae, ae, o);
        }
        if (basefields)
            for (int f = 0; f < 49; ++f) basefields[(int64_t)f * nrep + s] = o[f];
    }
    // confhist(slot, bin) is column-major in the reference: slot fastest
    if (hist) std::memcpy(hist, hout.data(), sizeof(double) * hout.size());
    if (hist_min) std::memcpy(hist_min, hmin.data(), sizeof(double) * nrep);
    if (hist_bin) std::memcpy(hist_bin, hbin.data(), sizeof(double) * nrep);
    for (int s = 0; s + 1 < nrep; ++s) {
        if (swap_att) swap_att[s] = (int32_t)att[s];
        if (swap_acc) swap_acc[s] = (int32_t)acc[s];
    }
    if (visitboth) {
        int vb = 0;
        for (int r = 0; r < nrep; ++r) vb += (vc[r] && vh[r]) ? 1 : 0;
        *visitboth = vb;
    }
    if (locspin) from_internal<int32_t, double>(cd.L, cd.H, lsp.data(), locspin, [](int32_t v) { return (double)v; });
    if (locnem) from_internal<int32_t, double>(cd.L, cd.H, lnm.data(), locnem, [](int32_t v) { return (double)v; });
    if (locnem_samples) *locnem_samples = (int32_t)ls[0];
    return 0;
}

int j1j2_get_flips(void *handle, int32_t conf, int32_t *flips)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    if (!E->st.flips) return fail(J1J2_ESTATE, "flip counts need J1J2_FLAG_WRITENUMFLIPS");
    const ConfDesc &cd = E->confs[conf];
    HIPCHK(hipSetDevice(E->device));
    std::vector<int32_t> buf;
    std::vector<uint8_t> pend;
    const int64_t n = (int64_t)cd.nrep * cd.N;
    if ((rc = d2h(*E, buf, E->st.flips + cd.flip_off, n)) || (rc = d2h(*E, pend, E->st.flips_pend + cd.flip_off, n)))
        return rc;
    HIPCHK(hipStreamSynchronize(E->stream));
    for (int64_t i = 0; i < n; ++i) buf[i] += pend[i];
    for (int s = 0; s < cd.nrep; ++s)
        from_internal<int32_t, int32_t>(cd.L, cd.H, buf.data() + (int64_t)s * cd.N, flips + (int64_t)s * cd.N,
                                        [](int32_t v) { return v; });
    return 0;
}

int j1j2_get_mismatch(void *handle, int32_t conf, int64_t *proposals, int64_t *mismatches)
{
    Engine *E = as_engine(handle);
    int rc;
    if ((rc = check_conf(E, conf))) return rc;
    HIPCHK(hipSetDevice(E->device));
    unsigned long long v[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(v, E->st.shadow + 2 * conf, sizeof(v), hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    if (proposals) *proposals = (int64_t)v[0];
    if (mismatches) *mismatches = (int64_t)v[1];
    return 0;
}

int j1j2_set_timing(void *handle, int32_t enable)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    E->timing = enable != 0;
    for (auto &T : E->timers) {
        T.total_ms = 0.0;
        T.launches = 0;
    }
    return 0;
}

int j1j2_get_timing(void *handle, int32_t kind, double *total_ms, int64_t *launches)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (kind < 0 || kind > 3) return fail(J1J2_EINVAL, "timing kind must be 0..3");
    if (total_ms) *total_ms = E->timers[kind].total_ms;
    if (launches) *launches = E->timers[kind].launches;
    return 0;
}

int j1j2_sweep_path(void *handle, int32_t *path, double *bytes)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (path) *path = E->path;
    // algorithmic HBM bytes per site update: spin read + write (1 + 1 B) and the
    // site's disorder value (FP32 4 B or FP64 8 B)
    if (bytes) *bytes = E->path == 0 ? 10.0 : 6.0;
    return 0;
}

int j1j2_launch_plan(void *handle, char *buf, int32_t len)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    if (!buf || len <= 0) return fail(J1J2_EINVAL, "need a buffer");
    std::string d;
    if (E->plan.launches.empty()) d = "no resident launches";
    else d = E->plan.describe() + (E->tuned ? " (autotuned)" : " (default)");
    std::snprintf(buf, (size_t)len, "%s", d.c_str());
    return 0;
}

int j1j2_synchronize(void *handle)
{
    Engine *E = as_engine(handle);
    if (!E) return fail(J1J2_EINVAL, "null engine handle");
    HIPCHK(hipSetDevice(E->device));
    HIPCHK(hipStreamSynchronize(E->stream));
    if (E->st.slab_err) {
        int32_t err = 0;
        HIPCHK(hipMemcpy(&err, E->st.slab_err, sizeof(err), hipMemcpyDeviceToHost));
        if (err)
            return fail(J1J2_EDEVICE, "bit-layout slabs: a wait for a neighbouring slab ran out (slabs not resident together)");
    }
    return 0;
}

}  // extern "C"
