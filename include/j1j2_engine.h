/*
 * j1j2_engine.h -- C-ABI of the MI355X engine for the Metropolis and
 * parallel-tempering core of sim/J1J2_6tv.f90.
 *
 * The reference has no plugin or FFI surface: its hot path is a set of
 * internal subroutines of `program J1J2_6` that share program-scope globals.
 * The boundary is therefore the reference's call sites, and each entry point
 * below names the call site it replaces:
 *
 *   PT SERIAL, per disorder config (J1J2_6tv.f90:340-344)
 *     call Initialize_pt_state        -> j1j2_set_spins (also resets the maps and accumulators)
 *     call Pt_equilibration_loop      -> j1j2_run(h, J1J2_PHASE_EQUIL, -1)
 *     call Pt_measurement_loop        -> j1j2_run(h, J1J2_PHASE_MEAS, -1)
 *     call Snapshot_pt_target_replica -> j1j2_get_spins + j1j2_get_maps
 *     call Collect_data packing       -> j1j2_fetch (J1J2_6tv.f90:2937-3000)
 *   LOCAL, per temperature (J1J2_6tv.f90:347-361)
 *     call Equilibration_loop / Measurement_loop / Process_data
 *                                     -> j1j2_run(EQUIL), j1j2_run(MEAS), j1j2_fetch
 *   Setup_phi_table (J1J2_6tv.f90:590-633)   -> j1j2_upload_disorder (field made on the host)
 *   make_corr_rand_3d (corr_rand_3d.f90:191-386) -> j1j2_make_disorder (host helper)
 *   Setup_temperature_ladder (J1J2_6tv.f90:928-962) -> j1j2_ladder (host helper)
 *   Validate_run_options (J1J2_6tv.f90:841-925) -> status codes of j1j2_create
 *
 * Conventions
 *   - Every function returns 0 on success and a negative J1J2_E* code on
 *     failure; j1j2_last_error() gives the message.  Nothing aborts the
 *     caller: a Fortran host maps a non-zero status to print + stop 1, as the
 *     reference does on invalid options.
 *   - Arrays are plain pointers in the reference's index order: sites x
 *     fastest (site = x + L*y + L*L*z, J1J2_6tv.f90:539), spins as int8 +1/-1,
 *     per-slot arrays slot fastest, replicas numbered 0..NREPLICA-1.
 *   - The caller owns every host buffer; the engine owns its device memory.
 *   - Calls are blocking and not re-entrant on one handle; one host thread
 *     per handle.  Configs are numbered 0..nconf_total-1 in the order of the
 *     groups passed to j1j2_create_batch.
 */
#ifndef J1J2_ENGINE_H
#define J1J2_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define J1J2_OK            0
#define J1J2_EINVAL      (-1)   /* invalid parameter or argument                */
#define J1J2_ENOMEM      (-2)   /* device allocation failed                     */
#define J1J2_EDEVICE     (-3)   /* HIP runtime error                            */
#define J1J2_ESTATE      (-4)   /* call not valid in the engine's current state */

/* sampler (SAMPLERMODE, J1J2_6tv.f90:64) */
#define J1J2_SAMPLER_PT     0
#define J1J2_SAMPLER_LOCAL  1
/* arithmetic form of the sweep */
#define J1J2_FP64_PLAIN     0   /* FP64 local field in the reference's order   */
#define J1J2_FP32_BATCHED   1   /* FP32 field and dE, FP64 accumulation        */
/* ladder (PTLADDERMODE, J1J2_6tv.f90:66; LOCAL uses TMIN + (k-1) DT) */
#define J1J2_LADDER_LINEAR  0
#define J1J2_LADDER_SPLIT   1
#define J1J2_LADDER_LOCAL   2
/* phases */
#define J1J2_PHASE_EQUIL    0   /* Pt_equilibration_loop / Equilibration_loop  */
#define J1J2_PHASE_MEAS     1   /* Pt_measurement_loop / Measurement_loop      */
/* flags */
#define J1J2_FLAG_WRITELOCNEM    1    /* J1J2_6tv.f90:74  */
#define J1J2_FLAG_WRITEHIST      2    /* J1J2_6tv.f90:75  */
#define J1J2_FLAG_WRITENUMFLIPS  4    /* J1J2_6tv.f90:77  */
#define J1J2_FLAG_SHADOW_FP64    8    /* count FP32-vs-FP64 flip-decision mismatches */
#define J1J2_FLAG_NO_GRAPH      16    /* launch every cycle eagerly (no hipGraph)    */
#define J1J2_FLAG_FORCE_GLOBAL  32    /* FP32: use the global-memory sweep only      */

/* Run parameters of one group of disorder configs with equal geometry.
 * Field names follow the reference's parameter block (J1J2_6tv.f90:43-96). */
typedef struct j1j2_params {
    int32_t L, H;             /* lattice L x L x H; both even                   */
    int32_t nreplica;         /* NREPLICA = NTEMP (PT) or number of chains (LOCAL) */
    int32_t ptswapever;       /* PTSWAPEVER (sweeps per cycle); 1 for LOCAL      */
    int32_t ptmeasindex;      /* PTMEASINDEX, 1-based                            */
    int32_t neq, nmess;       /* NEQ, NMESS: cycles (PT) or sweeps (LOCAL)      */
    int32_t nconf;            /* configs in this group                           */
    int32_t iconf0;           /* reference iconf of the group's first config     */
    int32_t bins;             /* BINS                                            */
    int32_t sampler;          /* J1J2_SAMPLER_*                                  */
    int32_t precision;        /* J1J2_FP64_PLAIN or J1J2_FP32_BATCHED           */
    int32_t ladder;           /* J1J2_LADDER_*                                   */
    int32_t flags;            /* J1J2_FLAG_*                                     */
    double  J1, deltaJ1, J2, J3;
    double  tmin, tmax, dt, tladdermid, tladderfrac;
    uint64_t seed;            /* device RNG seed (low 32 bits are the Philox key) */
} j1j2_params;

/* Temperature ladder of a parameter set, restating Setup_temperature_ladder
 * (J1J2_6tv.f90:928-962).  temps has p->nreplica entries. */
int j1j2_ladder(const j1j2_params *p, double *temps);

/* Correlated disorder of make_corr_rand_3d (corr_rand_3d.f90:191-386) for
 * seed IRINIT+iconf-1 (J1J2_6tv.f90:611): out has L*L*H values, x fastest. */
int j1j2_make_disorder(int32_t L, int32_t H, double xi, int32_t seed, int32_t map_to_uniform,
                       double phi, double *out);

/* The same field for configs iconf = 1 .. nconf of a run at once (seeds
 * seed0 + k, k = iconf - 1; seed0 = IRINIT as in Setup_phi_table,
 * J1J2_6tv.f90:611), on nthreads host threads (0: all cores).  out holds
 * nconf * L*L*H values, config-major; each field equals j1j2_make_disorder's. */
int j1j2_make_disorder_batch(int32_t L, int32_t H, double xi, int32_t seed0, int32_t nconf, int32_t map_to_uniform,
                             double phi, int32_t nthreads, double *out);

/* Initialize_spin_array (J1J2_6tv.f90:1179-1209) for one replica: start 0 = HOT
 * (KISS05 seeded with `seed`, then `skip` draws discarded, then one draw per
 * site: IRINIT+(iconf-1)*NREPLICA+rep and skip 0 for Initialize_pt_state,
 * J1J2_6tv.f90:1238; IRINIT+iconf-1 and skip 3*L*L*H for the LOCAL start after
 * Setup_vacancies, J1J2_6tv.f90:316-322), 1 = FER, 2 = STR; straxis 0 = X,
 * 1 = Y.  spins: L*L*H int8, x fastest. */
int j1j2_initial_spins(int32_t L, int32_t H, int32_t start, int32_t straxis, int32_t seed, int64_t skip,
                       int8_t *spins);

/* Engine lifetime.  j1j2_create = one group.  j1j2_create_batch takes ngroups
 * parameter sets (uneven L / NREPLICA batches) whose ptswapever, ptmeasindex,
 * neq, nmess, bins, sampler, precision and flags must agree. */
int  j1j2_create(const j1j2_params *p, int32_t device, void **handle);
int  j1j2_create_batch(const j1j2_params *groups, int32_t ngroups, int32_t device, void **handle);
void j1j2_destroy(void *handle);
int  j1j2_last_error(char *buf, int32_t len);

/* Geometry queries. */
int  j1j2_num_configs(void *handle, int32_t *nconf_total);
int  j1j2_config_shape(void *handle, int32_t conf, int32_t *L, int32_t *H, int32_t *nreplica);

/* Setup_phi_table (J1J2_6tv.f90:590-633): the field p(0:L3-1) of config conf. */
int  j1j2_upload_disorder(void *handle, int32_t conf, const double *p, int64_t n);
/* Initialize_pt_state (J1J2_6tv.f90:1218-1247): spins of all replicas of one
 * config (nreplica*N int8, replica-major), maps reset to the identity and every
 * per-config accumulator zeroed. */
int  j1j2_set_spins(void *handle, int32_t conf, const int8_t *spins, int32_t nreplica);
/* Overwrite the temperatures of one config (LOCAL carried protocol,
 * J1J2_6tv.f90:347-352).  temps: nreplica values. */
int  j1j2_set_temperatures(void *handle, int32_t conf, const double *temps, int32_t nreplica);

/* Run ncycles cycles of a phase for every config (ncycles < 0: NEQ or NMESS).
 * PT cycle = PTSWAPEVER sweeps of every replica at its slot's temperature,
 * then one pass of adjacent swaps alternating odd/even pairs (J1J2_6tv.f90:
 * 1282-1338); MEAS cycles also accumulate the per-slot moments (J1J2_6tv.f90:
 * 1437-1524).  The swap alternation restarts at each call with begin != 0.
 * j1j2_run of a measuring phase with no cycles (NMESS = 0) takes one
 * local-nematic sample of the current state, as the reference's NMESS <= 0
 * branches do (J1J2_6tv.f90:1414-1424, 2157-2160). */
int  j1j2_run(void *handle, int32_t phase, int32_t ncycles);
int  j1j2_run_cycles(void *handle, int32_t phase, int32_t ncycles, int32_t begin);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory once an engine exists: code under their own HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Enqueues the same
 * cycles as j1j2_run_cycles (bit-identical results) on `stream` (a
 * hipStream_t; NULL = the engine's own stream) and returns without waiting:
 * no device or host allocation, no graph build, no autotuning, no events, no
 * synchronisation, accumulators cleared by fill kernels.  The partial
 * observables a cycle's sweep hands to its swap step live in one static
 * device array per GPU of J1J2_NOALLOC_MAX_LATTICES x 8 int64 (1 MiB): an
 * engine with more lattices (configs x replicas) is rejected with J1J2_EINVAL.
 * All calls on one GPU share that array, so calls that overlap on one GPU
 * (other streams, other engines) race.  Wait with j1j2_synchronize (or on
 * `stream`, before any other call on the handle); j1j2_synchronize also
 * reports a bounded z-slab wait that ran out. */
#define J1J2_NOALLOC_MAX_LATTICES 16384
int  j1j2_run_cycles_noalloc(void *handle, int32_t phase, int32_t ncycles, int32_t begin, void *stream);

/* Results of one config after MEAS, packed as Collect_data does
 * (J1J2_6tv.f90:2937-2999).  Any pointer may be NULL.
 *   basefields : 49*nreplica, field-major, slot fastest (BASEFIELDS layout)
 *   hist       : nreplica*bins, slot fastest per bin (confhist(slot,bin))
 *   hist_min, hist_bin : nreplica (hist_en_min_t, hist_bin_size_t)
 *   swap_att, swap_acc : nreplica-1 (cfg_ptswap_attempts / _accepts)
 *   visitboth  : 1 (cfg_ptvisitboth)
 *   locspin, locnem : N sums over MEAS cycles (locspin_accum, locnem_accum)
 *   locnem_samples  : 1                                                   */
int  j1j2_fetch(void *handle, int32_t conf, double *basefields, double *hist,
                double *hist_min, double *hist_bin, int32_t *swap_att, int32_t *swap_acc,
                int32_t *visitboth, double *locspin, double *locnem, int32_t *locnem_samples);

/* Current state. */
int  j1j2_get_spins(void *handle, int32_t conf, int8_t *spins, int32_t nreplica);
int  j1j2_get_maps(void *handle, int32_t conf, int32_t *rep_at_temp, int32_t *temp_of_rep);
/* Per replica (by replica index) m, e, strx, stry, nbar after the last cycle:
 * out is 5*nreplica, quantity-major. */
int  j1j2_get_observables(void *handle, int32_t conf, double *out);
/* Raw per-slot moment sums (20*nreplica, slot-major), nmeasure, best, final. */
int  j1j2_get_sums(void *handle, int32_t conf, double *sums, int64_t *nmeasure,
                   double *best, double *final_en);
/* LOCAL: per-sweep series m, e, strx, stry, nbar of each chain (5*nmess per
 * chain, chain-major then quantity-major). */
int  j1j2_get_series(void *handle, int32_t conf, double *series, int64_t cap, int64_t *len);
/* confnumflips(slot, site) (WRITENUMFLIPS): nreplica*N int32, slot-major. */
int  j1j2_get_flips(void *handle, int32_t conf, int32_t *flips);
/* FP32_BATCHED with J1J2_FLAG_SHADOW_FP64: proposals and decisions that the
 * FP64 rule would have taken differently, per config. */
int  j1j2_get_mismatch(void *handle, int32_t conf, int64_t *proposals, int64_t *mismatches);

/* Measurement support.  Kernel timing brackets every launch of a kernel kind
 * with HIP events on the engine's stream. kind 0 = sweep, 1 = pt step,
 * 2 = observables, 3 = local nematic.  Returns total ms and launch count. */
int  j1j2_set_timing(void *handle, int32_t enable);
int  j1j2_get_timing(void *handle, int32_t kind, double *total_ms, int64_t *launches);
/* Which sweep path the engine picked (0 = FP64 plain, 1 = FP32 LDS-resident
 * byte layout for every lattice, 2 = FP32 global memory for at least one
 * lattice, 3 = FP32 LDS-resident with the bit layout for some lattices) and the
 * algorithmic bytes per site update it is priced at. */
int  j1j2_sweep_path(void *handle, int32_t *path, double *bytes_per_update);
/* The resident sweep's launch plan as text: per launch the lattice size, the
 * threads per block and the lattice count, whether the launches run on
 * parallel graph branches, and whether the autotuner picked it.  The engine
 * tunes on the first call of j1j2_run / j1j2_run_cycles long enough for the
 * candidates (at most 36 cycles, counted in the call's cycles); the result of
 * a run does not depend on the plan.  No counterpart in the reference. */
int  j1j2_launch_plan(void *handle, char *buf, int32_t len);
int  j1j2_synchronize(void *handle);

#ifdef __cplusplus
}
#endif
#endif
