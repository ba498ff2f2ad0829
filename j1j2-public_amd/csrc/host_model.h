// host_model.h -- host-side pieces of the engine that the reference computes on
// the CPU once per config or once per run: the KISS05 stream used for HOT
// starts and disorder noise, the correlated random field, the temperature
// ladder and the derived per-slot observables.
#pragma once

#include <stdint.h>

namespace j1j2 {

// KISS05 (J1J2_6tv.f90:3579-3695) and its polar Gaussian (corr_rand_3d.f90:48-85).
struct Kiss05 {
    uint32_t x, y, z, w;
    int iset;
    double gset;
    explicit Kiss05(int32_t seed);       // kissinit + one discarded draw
    double uniform();                    // rkiss05
    double gauss();                      // gkiss05
};

// make_corr_rand_3d (corr_rand_3d.f90:191-386): out[i + n1*(j + n2*k)].
int corr_field(int32_t n1, int32_t n2, int32_t n3, double xi, int32_t seed, bool map_uniform,
               double max_value, double *out);

// Initialize_spin_array (J1J2_6tv.f90:1179-1209); start 0 HOT, 1 FER, 2 STR.
void initial_spins(int32_t L, int32_t H, int32_t start, int32_t straxis, int32_t seed, int64_t skip, int8_t *s);

// Setup_temperature_ladder (J1J2_6tv.f90:928-962).
void temperature_ladder(int32_t mode, int32_t ntemp, double tmin, double tmax, double dt,
                        double tmid, double tfrac, double *t);

// Derived per-slot observables of Pt_measurement_loop (J1J2_6tv.f90:1560-1684),
// 49 values in the BASEFIELDS order of Collect_data (J1J2_6tv.f90:2937-2985).
void derive_basefields(const double *sums20, int64_t nmeasure, double beta, int32_t N,
                       double best, double final_en, double out49[49]);

// Process_data (J1J2_6tv.f90:2219-2474) from per-sweep series.
void process_local_series(const double *mag, const double *en, const double *strx,
                          const double *stry, const double *nbar, int64_t nmess, double beta,
                          int32_t N, int32_t bins, double out49[49], double *hist, double *hmin,
                          double *hbin);

}  // namespace j1j2
