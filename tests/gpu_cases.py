"""Shared set-up for the GPU parity tests: one engine run and the oracle run of
the same configs on the same inputs."""
import sys

import numpy as np

import j1j2_engine as J
import oracle as O

SEED = 20261203


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def config_inputs(L, H, R, iconf, start, irinit=1, straxis=0):
    """Disorder and starts as the reference makes them for config iconf
    (J1J2_6tv.f90:611, 1238)."""
    p = J.make_disorder(L, H, 3.0, irinit + iconf - 1, True, 2.0)
    if start == "hot":
        s = np.stack([J.initial_spins(L, H, J.START_HOT, 0, irinit + (iconf - 1) * R + r + 1)
                      for r in range(R)])
    elif start == "str":
        s = np.stack([J.initial_spins(L, H, J.START_STR, straxis, 0) for _ in range(R)])
    else:
        s = np.stack([J.initial_spins(L, H, J.START_FER, 0, 0) for _ in range(R)])
    return p, s


def build_engine(groups, starts):
    """groups: list of Params; starts: per config 'hot'/'str' list or one string."""
    eng = J.Engine(groups)
    inputs = []
    c = 0
    for g in groups:
        for i in range(g.nconf):
            st = starts if isinstance(starts, str) else starts[c]
            p, s = config_inputs(g.L, g.H, g.nreplica, g.iconf0 + i, st)
            eng.upload_disorder(c, p)
            eng.set_spins(c, s.reshape(-1))
            inputs.append((g, g.iconf0 + i, p, s))
            c += 1
    return eng, inputs


def oracle_for(g, iconf, p, s, shadow=False, count_flips=False):
    temps = J.ladder(g)
    return O.OraclePT(g.L, g.H, p, temps, s, ptswapever=g.ptswapever, ptmeasindex=g.ptmeasindex,
                      bins=g.bins, precision=g.precision, local=g.sampler == J.SAMPLER_LOCAL,
                      seed=g.seed & 0xFFFFFFFF, iconf=iconf, J1=g.J1, deltaJ1=g.deltaJ1, J2=g.J2,
                      J3=g.J3, series_cap=max(g.nmess, 1) if g.sampler == J.SAMPLER_LOCAL else 0,
                      shadow=shadow, count_flips=count_flips)


def compare_state(eng, conf, ora, temps, measured=True):
    """Bit-exact comparison of everything the engine reports for one config."""
    R = ora.nrep
    assert np.array_equal(eng.spins(conf), ora.spins), f"spins differ (conf {conf})"
    rat, tor = eng.maps(conf)
    assert np.array_equal(rat, ora.rep_at_temp) and np.array_equal(tor, ora.temp_of_rep)
    obs = eng.observables(conf)
    ref_vals = np.stack([ora.obs[:, 0] / ora.N, ora.etot / ora.N, ora.obs[:, 1] / ora.N,
                         ora.obs[:, 2] / ora.N, 2 * (ora.obs[:, 3] - ora.obs[:, 4]) / ora.N])
    assert np.array_equal(obs, ref_vals)
    if not measured:
        return
    out = eng.fetch(conf)
    if R > 1:
        assert np.array_equal(out["swap_att"], ora.attempts[: R - 1])
        assert np.array_equal(out["swap_acc"], ora.accepts[: R - 1])
        assert out["visitboth"] == ora.visitboth
    if eng.groups[0].flags & J.FLAG_WRITELOCNEM:
        assert np.array_equal(out["locspin"], ora.locspin.astype(np.float64))
        assert np.array_equal(out["locnem"], ora.locnem.astype(np.float64))
        assert out["locnem_samples"] == ora.s.locsamples
    if ora.s.local:
        ser = eng.series(conf, 0)
        assert np.array_equal(ser, ora.series)
        for sl in range(R):
            o49, h, hmin, hbin = O.process_local(*ora.series[sl], 1.0 / temps[sl], ora.N, ora.bins)
            assert np.array_equal(out["basefields"][:, sl], o49)
            assert np.array_equal(out["hist"][:, sl], h)
            assert out["hist_min"][sl] == hmin and out["hist_bin"][sl] == hbin
        return out
    sums, nm, best, fin = eng.sums(conf)
    assert nm == ora.s.nmeasure
    assert np.array_equal(sums, ora.sums), np.max(np.abs(sums - ora.sums))
    assert np.array_equal(best, ora.best) and np.array_equal(fin, ora.final_en)
    assert np.array_equal(out["hist"], ora.hist().T)
    assert np.array_equal(out["basefields"], ora.basefields())
    assert np.array_equal(out["hist_min"], ora.hist_min) and np.array_equal(out["hist_bin"], ora.hist_bin)
    return out
