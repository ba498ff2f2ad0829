/*
 * j1j2_oracle.h -- CPU restatement of the Metropolis + parallel-tempering
 * core of the J1-J2 random-field Ising code (sim/J1J2_6tv.f90) and of its
 * correlated-disorder generator (sim/corr_rand_3d.f90).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the engine library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker (or as the timed CPU
 * baseline), never as the thing measured or shipped.
 *
 * Each function names the reference lines it restates.  Where the engine
 * deliberately differs from the reference (checkerboard order instead of the
 * typewriter SEQ order, a counter-based RNG instead of the shared KISS stream
 * inside the sweep) the oracle implements the ENGINE's rule, because the
 * oracle is what the GPU results must equal bit for bit; agreement with the
 * reference itself is statistical and is checked separately.
 *
 * Pins (tests/test_oracle.py exercises each):
 *   kiss, corr field -> the reference's own phi_sites.dat for three (L, H,
 *                   seed), pow-2 and padded grids (tests/golden/ref_*), which
 *                   only the exact KISS seeding and stream, polar Gaussian,
 *                   filter, normalisation and erf map reproduce
 *   initial spins -> the reference's cfen of HOT and STR starts run with
 *                   NEQ = NMESS = 0
 *   philox       -> published Random123 known-answer vectors
 *   site words   -> every site of a colour gets exactly one call's word
 *                   half, each half of each call used once, high and low
 *                   parts from different calls (test_oracle.py)
 *   exp64, threshold32 -> exp / 2^y in double over dense grids
 *   fields       -> brute-force dE = E(after) - E(before); closed forms of
 *                   the stripe and ferro states
 *   sweep rule   -> beta = 0 flips every site; T -> 0 accepts only dE <= 0;
 *                   exact enumeration of 2x2x4 equilibrium averages
 *   swap rule    -> equal temperatures accept every attempt
 *   derived      -> closed forms on synthetic series
 *   LOCAL histogram -> the reference's own LOCAL cfhs outputs (a dynamic
 *                   run occupies bins 25..75, a frozen run only bin 50)
 */
#ifndef J1J2_ORACLE_H
#define J1J2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- KISS05 (J.f90:3614-3695, C.f90:48-160) ---------------- */
typedef struct { int32_t x, y, z, w; int32_t iset; double gset; } ora_kiss;
void   ora_kiss_init(ora_kiss *k, int32_t iinit);      /* kissinit  */
double ora_kiss_uniform(ora_kiss *k);                  /* rkiss05   */
double ora_kiss_gauss(ora_kiss *k);                    /* gkiss05   */
void   ora_kiss_fill(int32_t seed, int64_t n, double *out);   /* n draws after kissinit(seed) */

/* ---------------- Philox4x32-10 (Salmon et al., SC'11) ---------------- */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---------------- deterministic exponentials ---------------- */
/* 2^32 * 2^y in FP32 (fma Horner, no clamp; 0 below y = -183, inf above 96). */
float  ora_accept_threshold_f32(float y);
/* e^x in FP64: Cody-Waite reduction + degree-13 Taylor Horner, fma. */
double ora_exp64(double x);

/* ---------------- lattice and disorder ---------------- */
typedef struct {
    int32_t L, H, N;
    double  J1h, J1v, J2, J3;
    double  *p;        /* N, site = x + L*y + L*L*z (J.f90:539)        */
    /* FP32 site constants (engine FP32 rule) */
    float   *kc, *k2xp, *k2xm, *k2yp, *k2ym;
    float   j2x2, j3x2;
    int64_t *q;        /* llrint(p * 2^32): fixed-point disorder for energies */
} ora_lattice;

int  ora_lattice_init(ora_lattice *lat, int32_t L, int32_t H, double J1, double deltaJ1,
                      double J2, double J3, const double *p);
void ora_lattice_free(ora_lattice *lat);

/* Compute_local_field, J.f90:2498-2541 (occu = jh = jv = 1), FP64, same order. */
double ora_field_f64(const ora_lattice *lat, const int8_t *s, int32_t site, double *n_out);
/* Engine FP32 local field (order fixed by the engine). */
float  ora_field_f32(const ora_lattice *lat, const int8_t *s, int32_t site);

/* Colour of a site and its index within the colour (engine order). */
int32_t ora_colour(int32_t L, int32_t site);
int32_t ora_colour_index(int32_t L, int32_t site);

/* The sites of Philox call m of colour c (engine stream; see the .c file):
 * their colour indices and 32-bit words, up to 8; returns how many. */
int ora_call_words(int32_t L, uint32_t seed, uint32_t iconf, uint32_t replica, uint64_t sweep, int32_t c,
                   int32_t m, int32_t qi_out[8], uint32_t r_out[8]);

/* One 4-colour sweep of one replica.  precision: 0 = FP64 (reference field
 * order), 1 = FP32 engine rule.  beta is FP64; the FP32 form uses
 * b2 = (float)(2*beta*log2(e)).  Counts into *flips (may be NULL) and, when
 * shadow != NULL, counts decisions where the FP32 and FP64 rules disagree on
 * the same state and draw: shadow[0] += proposals, shadow[1] += mismatches. */
void ora_sweep(const ora_lattice *lat, int8_t *s, int precision, double beta,
               uint32_t seed, uint32_t iconf, uint32_t replica, uint64_t sweep,
               int32_t *flipcounts, int64_t *shadow);

/* Integer observables of a spin state:
 * out[0]=M=sum s, [1]=SX=sum s(-1)^x, [2]=SY=sum s(-1)^y, [3]=Bx, [4]=By,
 * [5]=Bd (+x+y and -x+y bonds), [6]=Bz, [7]=sum q_i l_i (fixed point). */
void   ora_observables(const ora_lattice *lat, const int8_t *s, int64_t out[8]);
/* Total energy from the integer observables (engine formula). */
double ora_energy_total(const ora_lattice *lat, const int64_t obs[8]);
/* Measurement (J.f90:2807-2842) recomputed literally in FP64, for pins. */
void   ora_measurement_ref(const ora_lattice *lat, const int8_t *s, double out[4]);
/* Spatial_mean_local_nematic, J.f90:1028-1048 (literal). */
double ora_nbar_ref(const ora_lattice *lat, const int8_t *s);

/* ---------------- initial states (J.f90:1179-1247) ---------------- */
/* start: 0 = HOT, 1 = FER, 2 = STR; straxis: 0 = X, 1 = Y. For HOT the
 * KISS stream is seeded with 'seed' (the caller passes the reference's
 * IRINIT + (iconf-1)*NREPLICA + rep). */
void ora_initial_spins(int32_t L, int32_t H, int32_t start, int32_t straxis,
                       int32_t seed, int8_t *s);

/* ---------------- temperature ladder (J.f90:928-962) ---------------- */
/* mode: 0 = LINEAR (PT), 1 = SPLIT (PT), 2 = LOCAL (TMIN + (k-1) DT). */
int ora_ladder(int32_t mode, int32_t ntemp, double tmin, double tmax, double dt,
               double tmid, double tfrac, double *temps);

/* ---------------- PT engine restatement (J.f90:1250-1685) ---------------- */
#define ORA_NSUMS 20
typedef struct {
    ora_lattice lat;
    int32_t nrep, ptswapever, ptmeasindex, bins, precision, local;
    uint32_t seed, iconf;
    int8_t  *spins;        /* nrep * N, replica-major            */
    int32_t *rep_at_temp, *temp_of_rep;
    double  *beta;         /* per slot                            */
    uint64_t cycle, sweep;
    int32_t  swap_phase;
    int64_t *att, *acc;    /* nrep-1                              */
    /* per-slot accumulators */
    double  *sums;         /* nrep * ORA_NSUMS, slot-major        */
    double  *best, *final_en;
    int64_t *hist;         /* nrep * bins                         */
    double  *hmin, *hmax, *hbin;
    int64_t  nmeasure;
    uint8_t *vcold, *vhot;
    int64_t *locspin, *locnem; int64_t locsamples;
    int32_t *flips;        /* nrep * N (per slot), may be NULL    */
    int64_t shadow[2];
    int      do_shadow;
    /* per replica current observables */
    int64_t *obs;          /* nrep * 8 */
    double  *etot;         /* nrep     */
    /* LOCAL series: nrep * 5 * cap */
    double  *series; int64_t series_len, series_cap;
} ora_pt;

int  ora_pt_init(ora_pt *pt, int32_t L, int32_t H, double J1, double deltaJ1, double J2, double J3,
                 const double *p, int32_t nrep, const double *temps, int32_t ptswapever,
                 int32_t ptmeasindex, int32_t bins, int32_t precision, int32_t local,
                 uint32_t seed, uint32_t iconf, const int8_t *spins0, int64_t series_cap,
                 int32_t count_flips, int32_t shadow);
void ora_pt_free(ora_pt *pt);
void ora_pt_begin_phase(ora_pt *pt);   /* swap_phase = 1 (J.f90:1266, 1425) */
void ora_pt_reset_measurement(ora_pt *pt);   /* start of a measuring phase */
void ora_pt_cycle(ora_pt *pt, int measure);
/* Derived per-slot fields, 49 per slot, layout of J.f90:2937-2985 (one config),
 * out[f*nrep + slot]. */
void ora_pt_basefields(const ora_pt *pt, double *out);
/* Histogram normalised by nmeasure (J.f90:1678-1680), out[slot*bins + b]. */
void ora_pt_hist(const ora_pt *pt, double *out);

/* Derived observables from per-slot sums (J.f90:1560-1684). sums: ORA_NSUMS
 * values in the accumulation order documented in the .c file. */
void ora_derive(const double *sums, int64_t nmeasure, double beta, int32_t N,
                double best, double final_en, double out49[49]);
/* LOCAL Process_data (J.f90:2219-2474) from per-sweep series. */
void ora_process_local(const double *mag, const double *en, const double *strx,
                       const double *stry, const double *nbar, int64_t nmess,
                       double beta, int32_t N, int32_t bins, double out49[49],
                       double *hist, double *hmin, double *hbin);

/* ---------------- exact enumeration ---------------- */
/* Boltzmann averages over all 2^N states (N <= 24) with the FP64 energy:
 * out = {<e>, <e^2>, <|m|>, <m^2>, <nbar>, <strx^2>, <stry^2>, Z-normalised
 * log partition function}. */
int ora_enumerate(const ora_lattice *lat, double beta, double out[8]);

/* ---------------- correlated disorder (C.f90:191-386) ---------------- */
/* out: n1*n2*n3, i fastest.  Own radix-2 complex FFT (not Ooura's). */
int ora_corr_field(int32_t n1, int32_t n2, int32_t n3, double xi, int32_t seed,
                   int32_t do_uniform, double max_value, double *out);
double ora_erfcc(double x);

#ifdef __cplusplus
}
#endif
#endif
