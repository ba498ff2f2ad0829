// kernels.h -- launch interface between the engine's host code and its HIP
// kernels (kernels.hip).  All launches are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace j1j2 {

// Per-config coupling constants needed on the device.
struct ConfConst {
    double J1h, J1v, J2, J3;
    float j2x2, j3x2;          // (float)(2*J2), (float)(2*J3)
    int64_t qsum;              // sum over the sites of llrint(p * 2^32)
};

// Offset that makes the fixed-point disorder q = llrint(p * 2^32), |q| < 2^38,
// non-negative for its five byte planes.
constexpr int64_t QOFF = int64_t(1) << 38;

struct DeviceState {
    // geometry
    const LatDesc *lats;       // nlat
    const ConfDesc *confs;     // nconf
    const ConfConst *cconst;   // nconf
    int32_t nlat, nconf;
    int32_t uni_nrep, uni_n;   // every config has uni_nrep replicas of uni_n sites (0: they differ)
    int64_t nsites;            // sum over configs of N (per-site planes)
    const int32_t *order;      // resident launches: lattice of each block (-1: none), per launch segment
    // lattice state
    uint8_t *spins;            // per lattice (LatDesc::spin_off): N bytes, or N/8 when packed; internal order
    const float *kd;           // 5 planes x nsites: kc, then 2^14 * (2Kx+, 2Kx-, 2Ky+, 2Ky-)
    const double *p64;         // nsites
    const uint8_t *qb;         // 5 planes x nsites: byte b of llrint(p * 2^32) + QOFF (|p| < 64)
    int64_t *obs;              // nlat x OBS_N
    // ladder
    int32_t *rep_at_temp;      // nslots (per config blocks of nrep)
    int32_t *temp_of_rep;      // nslots
    const double *beta;        // nslots
    const float *b2;           // nslots, (float)(2*beta*log2 e)
    DeviceClock *clk;
    // accumulators
    double *sums;              // nslots x NSUMS
    double *best, *final_en;   // nslots
    int64_t *hist;             // sum nrep*bins
    double *hmin, *hmax, *hbin;
    int64_t *att, *acc;        // npairs
    uint8_t *vcold, *vhot;     // nslots (per replica)
    int64_t *nmeasure;         // nconf
    double *etot;              // nslots (per replica)
    double *vals;              // nslots x 5 (per replica m, e, strx, stry, nbar)
    int32_t *locspin, *locnem; // nsites
    int64_t *locsamples;       // nconf
    int32_t *locnem_pending;   // nconf: the last cycle measured; its state is still to be accumulated
    int32_t fuse_locnem;       // 1: the resident sweep accumulates pending local-nematic sums
    uint32_t *locsnap;         // nsites / 32 words, or null: per byte-layout config (at site_off / 32) the
                               // packed spins of the replica a measuring PT step left at slot PTMEASINDEX
    double *series;            // LOCAL series
    int64_t *series_len;       // nconf
    int64_t series_cap;
    int32_t *flips;            // per slot per site (optional)
    uint8_t *flips_pend;       // per slot per site, beside flips: counts of the register-count sweep not
                               // yet added to flips (each < 128; the total is flips + flips_pend)
    unsigned long long *shadow;// nconf x 2
    // bit-layout lattices split into z-slabs (one block each): per lattice
    // BIT_SLAB_MAX slots of exchange words -- 4 planes x 2 boundary layers x
    // L/2 unit words (a 14-bit version above bit 50) and 16 words of partial
    // observables (a 32-bit version above 32-bit halves) -- and an error flag
    // set when a bounded wait for a neighbour's words runs out
    unsigned long long *slab_x, *slab_obs;
    int32_t *slab_err;
    int32_t slab_spin_max;     // polls before a wait gives up (< 0: every wait gives up at once)
    // run constants
    int32_t ptswapever, ptmeasindex, bins, local;
    uint32_t seed;
};

// LDS-resident FP32 sweep of `ptswapever` sweeps + observables, one block of
// nt threads per lattice: block b takes lattice order[ord0 + b] (nord blocks,
// a multiple of 8; -1 entries are idle).  Requires resident_supported(L) for
// every lattice and resident_nt_valid(L, nt).  L: the common lattice size of
// the segment, or 0 for mixed sizes; H: the common height when L > 0 (0 if
// the heights differ), and for mixed sizes -1 if every lattice has H = L (0
// otherwise).  The result depends neither on nt nor on the order.  With flip
// counts (st.flips set) each block keeps a byte per site after its lattice
// in LDS (2 max_n bytes; PTSWAPEVER <= 255) and adds them to the slot's
// counters at the end of the launch; the kernels that keep them in registers
// (256-thread blocks of cubic L = 32, PTSWAPEVER <= RC_MAX_SWAPEVER) need
// the lattice's LDS alone and add them to st.flips_pend.
hipError_t launch_sweep_resident(const DeviceState &st, int32_t ord0, int32_t nord,
                                 int32_t max_n, int32_t L, int32_t H, int32_t nt, bool shadow, hipStream_t s);
bool resident_supported(int32_t L);
// Bundles of the 1024-thread FP32 launch over lattices of several sizes: up to
// res_bundle(L) lattices of one config per block, side by side in LDS (one
// group of 1024 / res_bundle(L) threads each).  A work-list entry of that
// launch holds the first lattice in its low BUNDLE_SHIFT bits and the count
// minus one above them.
constexpr int BUNDLE_SHIFT = 24;
constexpr int32_t BUNDLE_LAT_MASK = (1 << BUNDLE_SHIFT) - 1;
constexpr int res_bundle(int32_t L) { return L == 16 ? 8 : (L == 24 || L == 32) ? 2 : 1; }
// The resident sweeps that keep flip counts in registers add them to a byte
// per site (flips_pend), which must not carry: PTSWAPEVER at most this.
constexpr int32_t RC_MAX_SWAPEVER = 127;
// Cubic L = 48 lattices with flip counts (a byte of counts beside the lattice
// would not fit one CU's LDS) keep them in bits 3..7 of the sites' own bytes,
// cleared at the end of each launch: PTSWAPEVER at most CB_MAX_SWAPEVER.
constexpr int32_t CB_MAX_SWAPEVER = 31;
constexpr bool byte_counts(int32_t L) { return L == 48; }
// The resident kernels' flip counts need no LDS beside the lattice for these
// (cubic FP32 lattices, PTSWAPEVER within the kernel's limit).
inline bool fast_counts_ok(int32_t L, bool cubic, bool shadow, int32_t ptswapever)
{
    return cubic && !shadow && ptswapever <= (byte_counts(L) ? CB_MAX_SWAPEVER : RC_MAX_SWAPEVER);
}
// Bytes of LDS the resident kernels add after the lattice (reduction scratch).
constexpr size_t RESIDENT_SCRATCH_LDS = sizeof(int64_t) * (1024 / 64) * 8;
// Bit-layout resident sweep (kernels_bits.hip): L % 8 == 0, 48 < L <= 96, one
// bit per site in LDS, block size fixed by L; same results as the byte layout.
bool bits_supported(int32_t L);
size_t bits_lds_need(int32_t L, int32_t H);
hipError_t setup_bits(size_t max_lds);
hipError_t launch_sweep_bits(const DeviceState &st, int32_t ord0, int32_t nord, int32_t L, int32_t H, int32_t slabs,
                             bool shadow, hipStream_t s);
// Slabs: S > 1 splits every lattice of the launch into S z-slabs of H / S
// (even) layers, one block each, order entries 4 * lattice + slab, the S
// blocks of a lattice co-resident (bits_slab_fits).  The result does not
// depend on S.
constexpr int BIT_SLAB_MAX = 4;
bool bits_slab_valid(int32_t L, int32_t H, int32_t slabs);
bool bits_slab_fits(int32_t L, int32_t H, int32_t slabs, int32_t nblocks, bool shadow);
size_t bits_slab_x_words(int32_t nlat);     // slab_x words for nlat lattices (L <= 96)
// nt in {256, 512, 576, 1024} for every L (and L = 0, mixed).
bool resident_nt_valid(int32_t L, int32_t nt);
// The resident kernel with NT threads per block for lattices of size L (0: a
// launch of several sizes), FP64 shadow or not, H = L for every lattice
// (cubic) or not, for an engine that counts flips or not; defined in
// kernels_res<NT>.hip.
using ResKernel = void (*)(DeviceState, int32_t, int32_t);
template <int NT>
ResKernel resident_kernels(int32_t L, bool shadow, bool cubic, bool flips);
size_t resident_lds_limit();
// Raise the dynamic-LDS limit of the resident kernel (once, before any launch or capture).
hipError_t setup_resident(size_t max_lds);
// Global-memory FP32 sweep, one colour pass of sweep k of the current cycle
// for the lattices [lat_begin, lat_begin + nlat) (all of one geometry).
hipError_t launch_sweep_global(const DeviceState &st, int32_t lat_begin, int32_t nlat,
                               int32_t L, int32_t H, int32_t k, int32_t colour, bool shadow,
                               hipStream_t s);
// FP64 plain sweep in the reference's field order, one colour pass.
hipError_t launch_sweep_plain64(const DeviceState &st, int32_t lat_begin, int32_t nlat,
                                int32_t L, int32_t H, int32_t k, int32_t colour, hipStream_t s);
// Integer observables of every lattice from global memory.
hipError_t launch_observables(const DeviceState &st, int32_t lat_begin, int32_t nlat, hipStream_t s);
// Energies, swaps and per-slot accumulation, one block per config.
// mode bit 0: accumulate measurements; bit 1: energies and values only (no swaps).
hipError_t launch_pt_step(const DeviceState &st, int32_t mode, hipStream_t s);
// locspin/locnem accumulation for the replica at slot PTMEASINDEX (of every
// config, or with only_pending of the configs whose accumulation is pending).
hipError_t launch_locnem(const DeviceState &st, int32_t max_n, int32_t only_pending, hipStream_t s);
// A phase begins at the current cycle: clk->phase_start = clk->ticks (the
// swap alternation restarts with the odd pairs).
hipError_t launch_begin_phase(const DeviceState &st, hipStream_t s);

// No-alloc cycles (noalloc.hip): the per-lattice partial observables of a
// cycle in a static device array of NOALLOC_MAX_LATTICES x OBS_N int64 per GPU.
// noalloc_setup looks its address up once per device (with that device
// current); noalloc_scratch returns it (nullptr before the set-up).
constexpr int32_t NOALLOC_MAX_LATTICES = 16384;
hipError_t noalloc_setup(int device);
int64_t *noalloc_scratch(int device);
// n 64-bit words (resp. bytes) at p set to v, by a kernel (capturable).
hipError_t launch_fill_u64(void *p, int64_t n, uint64_t v, hipStream_t s);
hipError_t launch_fill_u8(void *p, int64_t n, uint8_t v, hipStream_t s);

}  // namespace j1j2
