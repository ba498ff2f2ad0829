"""Benchmark: site updates/s of batched parallel tempering at L=32 on one MI355X.

Workload (one step = one PT cycle over the whole batch, inputs resident in HBM):
  L = H = 32 (32768 Ising spins), 16 replicas on a linear ladder T = 8.0-11.5
  x 64 disorder configs (PHI = 2, XI = 3, seeds IRINIT + iconf - 1), STR-X start,
  PTSWAPEVER = 5 sweeps + one swap pass + the measurement accumulation
  (moments, histogram, local nematic sums) -- the Pt_measurement_loop cycle.
  Site updates per step = 32768 * 16 * 64 * 5 = 167,772,160.

Printed as one JSON line: the FP32 batched form's rate (value), the plain FP64
form's rate on the same workload, the sweep kernel's roofline figures and the
reference's CPU rate on a bounded sample of the same workload.

Before the timed region: the autotuner's cycles, --warmup equilibration
cycles, further equilibration cycles until the GPU has been busy for
CLOCK_WARM_S (its engine clock ramps up over the first ~0.1 s of sustained
load: the sweep kernel runs ~10% slower before), and 31 measuring cycles
that build all five measuring hipGraphs (16, 8, 4, 2 and 1 cycles), so the
timed steps are steady-state replays.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "j1j2-public_amd"))

L, NREP, NCONF, PTSWAPEVER = 32, 16, 64, 5
HBM_PEAK = 8.0e12                  # MI355X HBM3E, bytes/s
TUNE_CYCLES = 60                   # the autotuner times its candidates on at most 54 cycles of one call
GRAPH_WARM = 31                    # one launch each of the 16-, 8-, 4-, 2- and 1-cycle graphs
CLOCK_WARM_S = 0.5                 # untimed equilibration until the GPU has run this long
FP64_MIN_STEPS = 20
SEED = 20261203


def build_batch(J, precision, flags=0, nconf=NCONF):
    g = J.make_params(L, L, nreplica=NREP, nconf=nconf, iconf0=1, ptswapever=PTSWAPEVER,
                      precision=precision, flags=flags | J.FLAG_WRITELOCNEM, tmin=8.0, tmax=11.5,
                      seed=SEED)
    eng = J.Engine([g])
    stripes = J.initial_spins(L, L, J.START_STR, 0, 0)
    import numpy as np
    spins = np.tile(stripes, NREP)
    for c in range(nconf):
        eng.upload_disorder(c, J.make_disorder(L, L, 3.0, 1 + c, True, 2.0))
        eng.set_spins(c, spins)
    return eng


def warm(J, eng, warmup):
    """Untimed lead-in: the autotuner's cycles, `warmup` equilibration cycles,
    more equilibration cycles until the GPU has been busy for CLOCK_WARM_S
    (steady engine clock), then GRAPH_WARM measuring cycles, which capture
    and instantiate the measuring hipGraphs (16, 8, 4, 2 and 1 cycles: a
    call's cycles replay as 16-cycle graphs and one graph of each smaller
    size for the remainder) before the timed region."""
    t0 = time.perf_counter()
    eng.run(J.PHASE_EQUIL, TUNE_CYCLES)
    eng.run(J.PHASE_EQUIL, warmup)
    while time.perf_counter() - t0 < CLOCK_WARM_S:
        eng.run(J.PHASE_EQUIL, 32)                 # blocking
    eng.run_cycles(J.PHASE_MEAS, GRAPH_WARM, begin=True)
    eng.synchronize()


def time_form(J, precision, steps, warmup, kernel_timing, nconf=NCONF, flags=0):
    """Timed region: `steps` MEAS cycles replayed as hipGraphs (both graphs
    built during the warm-up).  With kernel_timing, a second pass of the same
    length brackets every sweep launch with HIP events on the engine's stream
    (launched eagerly) for the sweep kernel's average duration."""
    eng = build_batch(J, precision, flags=flags, nconf=nconf)
    warm(J, eng, warmup)
    t0 = time.perf_counter()
    eng.run_cycles(J.PHASE_MEAS, steps, begin=False)
    eng.synchronize()
    dt = time.perf_counter() - t0
    sweep_ms, launches = 0.0, 0
    if kernel_timing:
        eng.set_timing(True)
        eng.run_cycles(J.PHASE_MEAS, steps, begin=False)
        sweep_ms, launches = eng.timing(0)
        eng.set_timing(False)
    path, bytes_per = eng.sweep_path()
    plan = eng.launch_plan()
    fetch = eng.fetch(0)
    eng.close()
    su = float(L * L * L) * NREP * nconf * PTSWAPEVER * steps
    return dict(seconds=dt, su=su, rate=su / dt, sweep_ms=sweep_ms, launches=launches, path=path,
                bytes_per=bytes_per, e_slot0=float(fetch["basefields"][9][0]), plan=plan)


def mismatch_rate(J, cycles):
    eng = build_batch(J, J.FP32_BATCHED, J.FLAG_SHADOW_FP64)
    eng.run(J.PHASE_EQUIL, cycles)
    tot = mis = 0
    for c in range(NCONF):
        a, b = eng.mismatch(c)
        tot += a
        mis += b
    eng.close()
    return mis / max(tot, 1), tot


PMC_SUMMARY = os.path.join("profiles", "k_sweep_resident_l32_16x64_pmc.md")


def traffic_from_profiles(nt):
    """HBM bytes per measuring sweep launch of the nt-thread variant from the
    committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), or None when the summary has no figure for
    that block size."""
    path = os.path.join(REPO, PMC_SUMMARY)
    if nt is None or not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch\[%d\]:\s*([0-9.eE+]+)" % nt, open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def physical_cores():
    """Distinct (physical id, core id) pairs of /proc/cpuinfo."""
    cores, phys = set(), None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":", 1)[1].strip()
            elif line.startswith("core id"):
                cores.add((phys, line.split(":", 1)[1].strip()))
    except OSError:
        pass
    return len(cores) or (os.cpu_count() or 1)


# host cores this one-GPU job may use (the GPU host's share for one GPU)
CPU_SHARE = 16


def reference_rate(name, L_, nrep, cycles, ranks):
    """Run the unmodified reference built from source (oracle/_ref/<name>:
    PT, L_^3 sites, nrep replicas, PTSWAPEVER=5, `cycles` equilibration
    cycles, one disorder config per MPI rank) on `ranks` host cores; su/s
    from its own 'timing PT_EQUIL' lines, or None if it cannot run here."""
    exe = os.path.join(REPO, "oracle", "_ref", name)
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    tmp = tempfile.mkdtemp(prefix="j1j2cpu")
    env = dict(os.environ, LD_LIBRARY_PATH="/opt/conda/lib:" + os.environ.get("LD_LIBRARY_PATH", ""))
    try:
        out = subprocess.run([mpiexec, "-n", str(ranks), exe], cwd=tmp, env=env, timeout=300,
                             capture_output=True, text=True)
        secs = {}
        for m in re.finditer(r"timing PT_EQUIL rank\s*(\d+) iconf\s*(\d+) seconds\s*([0-9.]+)", out.stdout):
            secs.setdefault(int(m.group(1)), []).append(float(m.group(3)))
        if not secs:
            return None
        per_cfg = float(L_ ** 3) * nrep * cycles * PTSWAPEVER
        # ranks run concurrently: each rank's rate over its first config
        rate = sum(per_cfg / v[0] for v in secs.values() if v[0] > 0)
        return dict(value=rate, unit="site updates/s", cores=len(secs), kind="reference",
                    sample=f"J1J2_6tv.f90 built from source (amdflang -O2, MPICH), L={L_}, "
                           f"NREPLICA={nrep}, {cycles} PT cycles x 5 sweeps, one config per rank, "
                           f"{len(secs)} rank(s) on {cpu_model()}; rate = sum over ranks of each "
                           f"rank's PT_EQUIL timing line")
    except (subprocess.SubprocessError, OSError):
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline():
    """The unmodified reference at L=32, NREPLICA=16, 100 cycles, one config
    per MPI rank, P = min(NCONF, physical cores, the job's host-core share)
    ranks (else the oracle, single thread)."""
    ranks = min(NCONF, physical_cores(), CPU_SHARE, os.cpu_count() or 1)
    ref = reference_rate("j1j2_cpu_l32", L, NREP, 100, ranks)
    if ref:
        ref["rank_rule"] = (f"P = min(NCONF = {NCONF}, physical cores = {physical_cores()}, "
                            f"host-core share of a one-GPU job = {CPU_SHARE}) = {ranks}")
        return ref
    # fall back to the oracle's restatement, single thread
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import oracle as O
    import j1j2_engine as J
    p = J.make_disorder(L, L, 3.0, 1, True, 2.0)
    s0 = np.tile(J.initial_spins(L, L, J.START_STR, 0, 0), NREP)
    ora = O.OraclePT(L, L, p, O.ladder(0, NREP, 8.0, 11.5), s0, precision=1)
    ora.begin_phase()
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 10.0:
        ora.cycle(False)
        n += 1
    dt = time.perf_counter() - t0
    return dict(value=float(L ** 3) * NREP * PTSWAPEVER * n / dt, unit="site updates/s", cores=1,
                kind="port", sample=f"oracle FP32 rule, one config, {n} PT cycles, single thread")


# ---------------------------------------------------------------- other workloads
# Reported beside the headline line (python bench.py --workload c4|c5|c2); the
# default run measures the L=32 batch above only.
def build_groups(J, groups, precision, flags=0, seed=SEED):
    """groups: list of (L, nrep, nconf, start) -> one engine over all of them."""
    import numpy as np
    params = [J.make_params(Lg, Lg, nreplica=R, nconf=C, iconf0=1 + sum(g[2] for g in groups[:k]),
                            ptswapever=PTSWAPEVER, precision=precision, flags=flags, tmin=8.0, tmax=11.5,
                            seed=seed)
              for k, (Lg, R, C, _) in enumerate(groups)]
    eng = J.Engine(params)
    conf = 0
    for (Lg, R, C, start), g in zip(groups, params):
        for c in range(C):
            iconf = g.iconf0 + c
            eng.upload_disorder(conf, J.make_disorder(Lg, Lg, 3.0, iconf, True, 2.0))
            st = start if start != "mixed" else ("hot" if c % 2 == 0 else "str")
            if st == "hot":
                spins = np.concatenate([J.initial_spins(Lg, Lg, J.START_HOT, 0, (iconf - 1) * R + r + 1)
                                        for r in range(R)])
            else:
                spins = np.tile(J.initial_spins(Lg, Lg, J.START_STR, 0, 0), R)
            eng.set_spins(conf, spins)
            conf += 1
    return eng


def time_groups(J, groups, precision, steps, warmup, flags=0):
    eng = build_groups(J, groups, precision, flags)
    warm(J, eng, warmup)
    t0 = time.perf_counter()
    eng.run_cycles(J.PHASE_MEAS, steps, begin=False)
    eng.synchronize()
    dt = time.perf_counter() - t0
    eng.set_timing(True)
    eng.run_cycles(J.PHASE_MEAS, steps, begin=False)
    sweep_ms, launches = eng.timing(0)
    eng.set_timing(False)
    path, bytes_per = eng.sweep_path()
    plan = eng.launch_plan()
    eng.close()
    su = sum(float(Lg ** 3) * R * C for Lg, R, C, _ in groups) * PTSWAPEVER * steps
    return dict(seconds=dt, su=su, rate=su / dt, sweep_ms=sweep_ms, launches=launches, path=path,
                bytes_per=bytes_per, plan=plan)


def group_mismatch(J, groups, cycles):
    eng = build_groups(J, groups, J.FP32_BATCHED, J.FLAG_SHADOW_FP64)
    eng.run(J.PHASE_EQUIL, cycles)
    tot = mis = 0
    for c in range(eng.nconf):
        a, b = eng.mismatch(c)
        tot += a
        mis += b
    eng.close()
    return mis / max(tot, 1), tot


WORKLOADS = {
    "c4": ("uneven batch: L in {16,24,32,48} x 16 configs each, replicas {8,12,16,24}, "
           "HOT/STR starts 50/50, T=8.0-11.5, FP32",
           [(16, 8, 16, "mixed"), (24, 12, 16, "mixed"), (32, 16, 16, "mixed"), (48, 24, 16, "mixed")]),
    "c5": ("large lattice: L=96, 24 replicas x 8 configs, STR start, T=8.0-11.5, FP32",
           [(96, 24, 8, "str")]),
    "c2": ("PT, L=16, 16 replicas x 1 config, STR start, T=8.0-11.5",
           [(16, 16, 1, "str")]),
}


def other_workload(J, name, steps, warmup, with_fp64=True, with_cpu=True, with_flips=False):
    desc, groups = WORKLOADS[name]
    fp32 = time_groups(J, groups, J.FP32_BATCHED, steps, warmup)
    fl = time_groups(J, groups, J.FP32_BATCHED, steps, warmup, J.FLAG_WRITENUMFLIPS) if with_flips else None
    n64 = max(steps // 10, FP64_MIN_STEPS)
    fp64 = time_groups(J, groups, J.FP64_PLAIN, n64, max(warmup // 10, 2)) if with_fp64 else None
    su_launch = fp32["su"] / max(fp32["launches"], 1)
    avg_s = fp32["sweep_ms"] / 1e3 / max(fp32["launches"], 1)
    rate, n = group_mismatch(J, groups, 4)
    return {
        "metric": "site updates/s (all replicas)", "value": fp32["rate"], "unit": "site updates/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * fp32["seconds"] / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": {"workload": desc, "sweep_path": J.PATH_NAMES[fp32["path"]],
                                        "launch_plan": fp32["plan"]},
        "sweep_kernel": {"avg_launch_ms": avg_s * 1e3, "launches_per_step": fp32["launches"] / steps,
                         "algorithmic_GBps": su_launch * fp32["bytes_per"] / avg_s / 1e9,
                         "frac_of_hbm_peak": su_launch * fp32["bytes_per"] / avg_s / HBM_PEAK},
        "fp64_plain": ({"value": fp64["rate"], "steps": n64, "fp32_over_fp64": fp32["rate"] / fp64["rate"]}
                       if fp64 else None),
        "flip_mismatch_rate": {"fp32_vs_fp64": rate, "proposals": n},
        "with_flip_counts": ({"value": fl["rate"], "sweep_path": J.PATH_NAMES[fl["path"]], "launch_plan": fl["plan"],
                              "over_value": fl["rate"] / fp32["rate"]} if fl else None),
        # L = 16, one config: the reference runs it on one core (configs are
        # its unit of parallelism), 1000 cycles (~10 s)
        "cpu_baseline": (reference_rate("j1j2_cpu_l16", 16, 16, 1000, 1)
                         if name == "c2" and with_cpu else None),
    }


def copy_probe(nbytes, reps=5):
    """Five plain device-to-device copies of nbytes (the runtime's copy
    kernel): the known traffic that calibrates the FETCH_SIZE / WRITE_SIZE
    counters."""
    import ctypes
    J.lib()                                        # the runtime the engine uses
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    for buf in (src, dst):
        if hip.hipMalloc(ctypes.byref(buf), n) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, n) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(reps):
            if hip.hipMemcpy(dst, src, n, 3) != 0:     # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=0, help="cycles for the FP64 form (default steps/10)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--workload", default="c3", choices=["c3"] + sorted(WORKLOADS),
                    help="c3 = the headline L=32 batch; the others are reported beside it")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run 5 plain device copies of this many bytes (PMC counter calibration)")
    ap.add_argument("--block-threads", type=int, default=0,
                    help="fix the resident sweep's threads per block (no autotuning), e.g. for counter runs")
    ap.add_argument("--flips", action="store_true",
                    help="also time the L=32 batch with per-site flip counts (WRITENUMFLIPS)")
    ap.add_argument("--bit-slabs", type=int, default=0,
                    help="fix the z-slabs per bit-layout lattice (no autotuning), e.g. for counter runs")
    args = ap.parse_args()
    if args.block_threads:
        os.environ["J1J2_RES_THREADS"] = str(args.block_threads)
    if args.bit_slabs:
        os.environ["J1J2_BIT_SLABS"] = str(args.bit_slabs)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; the path would shard disorder configs across GPUs "
              "(one process per GPU, each its own engine over its share of the configs, no exchange "
              "during the run, results gathered once at the end).", file=sys.stderr)
        sys.exit(2)
    global J
    import j1j2_engine as J

    if args.copy_probe:
        copy_probe(args.copy_probe)

    if args.workload != "c3":
        out = other_workload(J, args.workload, args.steps, args.warmup, not args.no_fp64, not args.no_cpu,
                             args.flips)
        cpu = out["cpu_baseline"]
        if cpu and out["fp64_plain"]:
            out["fp64_plain"]["over_cpu_reference"] = out["fp64_plain"]["value"] / cpu["value"]
        print(json.dumps(out), flush=True)
        return

    fp32 = time_form(J, J.FP32_BATCHED, args.steps, args.warmup, kernel_timing=True)
    res = {
        "metric": "site updates/s (all replicas) at L=32, FP32 vs FP64 form; % HBM peak",
        "value": fp32["rate"],
        "unit": "site updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * fp32["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched PT, L=32, 16 replicas (T=8.0-11.5) x 64 disorder configs "
                               "(PHI=2, XI=3), 5 sweeps + swap + measurement per step",
                   "L": L, "H": L, "nreplica": NREP, "nconf": NCONF, "ptswapever": PTSWAPEVER,
                   "sweep_path": J.PATH_NAMES[fp32["path"]], "launch_plan": fp32["plan"]},
    }
    su_launch = fp32["su"] / max(fp32["launches"], 1)
    avg_s = fp32["sweep_ms"] / 1e3 / max(fp32["launches"], 1)
    achieved = su_launch * fp32["bytes_per"] / avg_s
    m_nt = re.search(r"nt(\d+)", fp32["plan"])
    timed_nt = int(m_nt.group(1)) if m_nt else None
    traffic = traffic_from_profiles(timed_nt)
    res["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK, "traffic": traffic,
                       "kernel": "k_sweep_resident", "avg_launch_ms": avg_s * 1e3,
                       "site_updates_per_launch": su_launch,
                       "algorithmic_bytes_per_update": fp32["bytes_per"],
                       # achieved/frac are the algorithmic-byte equivalent (6 B per site
                       # update over the launch time); the bytes that cross HBM are the
                       # counters' traffic, read on the variant named here
                       "achieved_basis": "algorithmic bytes (6 B per site update) / average launch time",
                       "traffic_source": PMC_SUMMARY,
                       "traffic_block_threads": timed_nt if traffic else None,
                       "timed_block_threads": timed_nt,
                       "measured_hbm_GBps": (traffic / avg_s / 1e9) if traffic else None,
                       "limiter": "VALU issue (the lattice stays in LDS for the launch's 5 sweeps)"}
    if not args.no_fp64:
        n64 = args.fp64_steps or max(args.steps // 10, FP64_MIN_STEPS)
        fp64 = time_form(J, J.FP64_PLAIN, n64, max(args.warmup // 10, 2), kernel_timing=False)
        res["fp64_plain"] = {"value": fp64["rate"], "steps": n64, "ms_per_step": 1e3 * fp64["seconds"] / n64,
                             "fp32_over_fp64": fp32["rate"] / fp64["rate"],
                             "reading": "same batching as the FP32 form: one engine over the 64 configs x "
                                        "16 replicas, FP64 field in the reference's order, global-memory "
                                        "colour passes"}
        # the same FP64 form run the reference's way, one disorder config at a
        # time (one engine of 16 replicas)
        one = time_form(J, J.FP64_PLAIN, 4 * n64, max(args.warmup // 10, 2), kernel_timing=False, nconf=1)
        res["fp64_plain"]["one_config_at_a_time"] = {"value": one["rate"], "steps": 4 * n64,
                                                     "fp32_over_fp64": fp32["rate"] / one["rate"],
                                                     "reading": "configs one at a time as the reference "
                                                                "loops over them (an engine of 16 replicas)"}
        rate, n = mismatch_rate(J, 4)
        res["flip_mismatch_rate"] = {"fp32_vs_fp64": rate, "proposals": n}
    if args.flips:
        fl = time_form(J, J.FP32_BATCHED, args.steps, args.warmup, kernel_timing=False, flags=J.FLAG_WRITENUMFLIPS)
        res["with_flip_counts"] = {"value": fl["rate"], "sweep_path": J.PATH_NAMES[fl["path"]],
                                   "launch_plan": fl["plan"], "over_value": fl["rate"] / fp32["rate"]}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
