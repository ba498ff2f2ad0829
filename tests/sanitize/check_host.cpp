// check_host.cpp -- built with AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_sanitize.py): exercises the engine's host-side code
// (j1j2-public_amd/csrc/host_model.cpp: KISS05, correlated disorder, starts,
// ladders, derived observables, LOCAL processing) against the CPU oracle, and
// the oracle's PT, LOCAL, shadow, flip-count and enumeration paths, so that
// every buffer they touch is checked.  Prints "sanitize check: ok" and exits 0
// when every comparison holds; any sanitizer report aborts the run.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "host_model.h"
#include "j1j2_oracle.h"

namespace {

int failures = 0;

void expect(bool ok, const char *what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        ++failures;
    }
}

double max_abs_diff(const double *a, const double *b, size_t n)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) m = std::fmax(m, std::fabs(a[i] - b[i]));
    return m;
}

void check_kiss()
{
    for (int seed : {1, 7, 12345, 2147483000}) {
        j1j2::Kiss05 k(seed);
        ora_kiss o;
        ora_kiss_init(&o, seed);
        bool same = true;
        for (int i = 0; i < 5000; ++i) same = same && k.uniform() == ora_kiss_uniform(&o);
        for (int i = 0; i < 1000; ++i) same = same && k.gauss() == ora_kiss_gauss(&o);
        expect(same, "KISS05 uniform and gaussian streams");
    }
}

void check_fields()
{
    const int shapes[][3] = {{8, 8, 8}, {6, 6, 4}, {12, 12, 10}, {10, 10, 10}, {16, 16, 16}, {2, 2, 4}, {24, 24, 6}};
    for (const auto &sh : shapes) {
        const size_t n = (size_t)sh[0] * sh[1] * sh[2];
        for (int uniform : {0, 1}) {
            std::vector<double> a(n), b(n);
            expect(j1j2::corr_field(sh[0], sh[1], sh[2], 3.0, 11, uniform != 0, 2.0, a.data()) == 0, "corr_field rc");
            expect(ora_corr_field(sh[0], sh[1], sh[2], 3.0, 11, uniform, 2.0, b.data()) == 0, "ora_corr_field rc");
            expect(max_abs_diff(a.data(), b.data(), n) < 1e-9, "correlated field: host model vs oracle");
        }
    }
}

void check_starts_and_ladders()
{
    const int L = 6, H = 4, N = L * L * H;
    std::vector<int8_t> a(N), b(N);
    for (int start : {0, 1, 2})
        for (int axis : {0, 1}) {
            j1j2::initial_spins(L, H, start, axis, 9, 0, a.data());
            ora_initial_spins(L, H, start, axis, 9, b.data());
            expect(a == b, "initial spins: host model vs oracle");
        }
    j1j2::initial_spins(L, H, 0, 0, 9, 3 * N, a.data());     // after the vacancy draws
    expect(std::all_of(a.begin(), a.end(), [](int8_t v) { return v == 1 || v == -1; }), "HOT start with skip");
    for (int mode : {0, 1, 2}) {
        const int nt = 9;
        std::vector<double> t1(nt), t2(nt);
        j1j2::temperature_ladder(mode, nt, 8.0, 11.5, 0.5, 9.5, 0.7, t1.data());
        expect(ora_ladder(mode, nt, 8.0, 11.5, 0.5, 9.5, 0.7, t2.data()) == 0, "ora_ladder rc");
        expect(max_abs_diff(t1.data(), t2.data(), nt) == 0.0, "temperature ladder: host model vs oracle");
    }
}

void check_derived()
{
    std::srand(5);
    auto rnd = [] { return (double)std::rand() / RAND_MAX; };
    for (int trial = 0; trial < 50; ++trial) {
        double S[ORA_NSUMS];
        for (double &v : S) v = rnd() * 100.0;
        double a[49], b[49];
        const int64_t nm = 1 + trial * 37;
        j1j2::derive_basefields(S, nm, 0.11, 512, -1.5, -1.2, a);
        ora_derive(S, nm, 0.11, 512, -1.5, -1.2, b);
        bool same = true;
        for (int f = 0; f < 49; ++f) same = same && (a[f] == b[f] || (std::isnan(a[f]) && std::isnan(b[f])));
        expect(same, "derived observables: host model vs oracle");
    }
    for (int degenerate : {0, 1}) {
        const int64_t n = 300;
        std::vector<double> m(n), e(n), x(n), y(n), nb(n);
        for (int64_t i = 0; i < n; ++i) {
            m[i] = degenerate ? 0.25 : rnd() - 0.5;
            e[i] = degenerate ? -1.0 : -3.0 + rnd();
            x[i] = degenerate ? 0.0 : rnd() - 0.5;
            y[i] = degenerate ? 0.0 : rnd() - 0.5;
            nb[i] = degenerate ? 0.0 : rnd() * 4.0 - 2.0;
        }
        double a[49], b[49], ha[100], hb[100], mina, minb, bina, binb;
        j1j2::process_local_series(m.data(), e.data(), x.data(), y.data(), nb.data(), n, 0.25, 512, 100, a, ha,
                                   &mina, &bina);
        ora_process_local(m.data(), e.data(), x.data(), y.data(), nb.data(), n, 0.25, 512, 100, b, hb, &minb, &binb);
        expect(max_abs_diff(a, b, 49) == 0.0 && max_abs_diff(ha, hb, 100) == 0.0 && mina == minb && bina == binb,
               "LOCAL processing: host model vs oracle");
    }
}

void check_oracle_runs()
{
    const int L = 4, H = 4, N = L * L * H;
    std::vector<double> p(N);
    expect(ora_corr_field(L, L, H, 3.0, 3, 1, 2.0, p.data()) == 0, "field");
    for (int precision : {0, 1})
        for (int local : {0, 1}) {
            const int nrep = local ? 2 : 3;
            std::vector<double> t(nrep);
            for (int i = 0; i < nrep; ++i) t[i] = 8.0 + i;
            std::vector<int8_t> s0((size_t)nrep * N);
            for (int r = 0; r < nrep; ++r) ora_initial_spins(L, H, 0, 0, 100 + r, s0.data() + (size_t)r * N);
            ora_pt pt;
            expect(ora_pt_init(&pt, L, H, 1.0, 0.1, -1.0, 1.0, p.data(), nrep, t.data(), local ? 1 : 5, 1, 20,
                               precision, local, 77, 4, s0.data(), local ? 40 : 0, 1, precision) == 0,
                   "ora_pt_init");
            ora_pt_begin_phase(&pt);
            for (int i = 0; i < 7; ++i) ora_pt_cycle(&pt, 0);
            ora_pt_reset_measurement(&pt);
            ora_pt_begin_phase(&pt);
            for (int i = 0; i < 13; ++i) ora_pt_cycle(&pt, 1);
            std::vector<double> bf((size_t)49 * nrep), h((size_t)nrep * 20);
            ora_pt_basefields(&pt, bf.data());
            ora_pt_hist(&pt, h.data());
            expect(pt.nmeasure == 13 && pt.locsamples == 13, "measurement counts");
            bool finite = true;
            for (double v : bf) finite = finite && std::isfinite(v);
            expect(finite, "finite basefields");
            ora_pt_free(&pt);
        }
    ora_lattice lat;
    const int Le = 2, He = 4;
    std::vector<double> pe(Le * Le * He, 0.3);
    expect(ora_lattice_init(&lat, Le, He, 1.0, 0.0, -1.0, 1.0, pe.data()) == 0, "lattice");
    double out[8];
    expect(ora_enumerate(&lat, 0.25, out) == 0 && std::isfinite(out[0]), "exact enumeration");
    std::vector<int8_t> s(Le * Le * He, 1);
    double f64[1];
    for (int i = 0; i < Le * Le * He; ++i) {
        (void)ora_field_f64(&lat, s.data(), i, f64);
        (void)ora_field_f32(&lat, s.data(), i);
    }
    ora_lattice_free(&lat);
}

}  // namespace

int main()
{
    check_kiss();
    check_fields();
    check_starts_and_ladders();
    check_derived();
    check_oracle_runs();
    if (failures) {
        std::printf("sanitize check: %d failures\n", failures);
        return 1;
    }
    std::printf("sanitize check: ok\n");
    return 0;
}
