"""The Fortran host (j1j2-public_amd/fortran): the reference's driver over the
engine's C-ABI through iso_c_binding.

CPU: the host is built and linked against the in-tree library; invalid
options end the run with the reference's message and exit status 1
(Validate_run_options, J1J2_6tv.f90:841-925) before any device call.

GPU: (1) where the reference's output is fixed by its disorder and starts
alone -- NEQ = NMESS = 0, and the frozen PT run (FER start, phi = 0,
T = 0.01-0.02: no spin flips, every swap is taken) -- the host must rewrite
the reference's own files (tests/golden/ref_*: cfen, cfma, cfnm, cfst, cfhs,
cfxy, cfnd, cfpt, cfph, cfnu, phi_sites.dat and the locnem files) line for
line.  (2) With dynamics, every file has the layout of the reference's file
from the same parameters: the same lines, the same words, each number in the
same format.  (3) With dynamics, the host's files carry exactly what the
engine returns through the Python binding (PT, batched across engine
launches), and what the oracle gives for the LOCAL protocols (INDEP reseeds
every temperature, CARRY carries one chain down the ladder).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from gpu_cases import say
from local_protocol import read_locnem, start_spins, temperatures
from refdata import find, read_table

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "j1j2-public_amd", "j1j2_host")

# Parameter blocks of the reference runs kept under tests/golden/ (the
# variants built by oracle/ref_build.sh).
REF_RUNS = {
    "hot_init_l8": "L=8, H=8, SAMPLERMODE='PT', NREPLICA=4, TMIN=8.0D0, TMAX=11.5D0, START='HOT', "
                   "NEQ=0, NMESS=0, NCONF=3, IRINIT=1, WRITEBINDTIME=.false.",
    "hot_init_l6h4": "L=6, H=4, SAMPLERMODE='PT', NREPLICA=2, TMIN=8.0D0, TMAX=11.5D0, START='HOT', "
                     "NEQ=0, NMESS=0, NCONF=2, IRINIT=7, WRITEBINDTIME=.false., WRITENUMFLIPS=.true.",
    "str_init_l12": "L=12, H=10, SAMPLERMODE='PT', NREPLICA=2, TMIN=8.0D0, TMAX=11.5D0, START='STR', "
                    "STRAXIS='Y', NEQ=0, NMESS=0, NCONF=1, IRINIT=11, WRITEBINDTIME=.false.",
    "pt_frozen": "L=8, H=8, SAMPLERMODE='PT', NREPLICA=5, TMIN=0.01D0, TMAX=0.02D0, START='FER', phi=0.0D0, "
                 "NEQ=11, NMESS=7, NCONF=1, IRINIT=1, WRITEBINDTIME=.false.",
    "local_indep_init": "L=6, H=4, SAMPLERMODE='LOCAL', TEMPPROTOCOL='INDEP', TMIN=4.0D0, TMAX=8.0D0, DT=2.0D0, "
                        "START='HOT', NEQ=0, NMESS=0, NCONF=2, IRINIT=5, WRITEBINDTIME=.false.",
    "local_carry_init": "L=6, H=4, SAMPLERMODE='LOCAL', TEMPPROTOCOL='CARRY', TMIN=4.0D0, TMAX=8.0D0, DT=2.0D0, "
                        "START='HOT', NEQ=0, NMESS=0, NCONF=2, IRINIT=5, WRITEBINDTIME=.false.",
}

# Reference runs with dynamics (different random streams: compared by layout)
DYN_RUNS = {
    "pt_dyn_l6h4": "L=6, H=4, SAMPLERMODE='PT', NREPLICA=3, TMIN=8.0D0, TMAX=11.5D0, START='HOT', NEQ=100, "
                   "NMESS=300, NCONF=2, IRINIT=7, WRITEBINDTIME=.false., WRITENUMFLIPS=.true.",
    "local_dyn_l6h4": "L=6, H=4, SAMPLERMODE='LOCAL', TEMPPROTOCOL='INDEP', TMIN=4.0D0, TMAX=8.0D0, DT=2.0D0, "
                      "START='HOT', NEQ=100, NMESS=200, NCONF=2, IRINIT=5, WRITEBINDTIME=.false., "
                      "WRITENUMFLIPS=.true.",
    "pt_l8_stat": "L=8, H=8, SAMPLERMODE='PT', NREPLICA=4, TMIN=8.0D0, TMAX=11.5D0, START='STR', STRAXIS='X', "
                  "NEQ=2000, NMESS=40000, NCONF=1, IRINIT=1, WRITEBINDTIME=.false., WRITELOCNEM=.false.",
    "pt_l16_c2": "L=16, H=16, SAMPLERMODE='PT', NREPLICA=16, TMIN=8.0D0, TMAX=11.5D0, START='STR', STRAXIS='X', "
                 "NEQ=2000, NMESS=2000, NCONF=1, IRINIT=1, WRITEBINDTIME=.false., WRITELOCNEM=.false., "
                 "WRITEPHITABLE=.false.",
}


def output_files(gd):
    return sorted(n for n in os.listdir(gd)
                  if n.startswith("cf") or n.startswith("locnem") or n == "phi_sites.dat")

NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([EeDd][+-]?\d+)?$")


def run_host(tmp_path, body, timeout=300):
    nml = tmp_path / "run.nml"
    nml.write_text("&j1j2\n " + body + "\n/\n")
    return subprocess.run([HOST, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)


def test_host_is_built_against_the_engine():
    assert os.access(HOST, os.X_OK), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", HOST], capture_output=True, text=True, check=True).stdout
    lib = [ln for ln in out.splitlines() if "libj1j2_engine.so" in ln]
    assert lib and os.path.join(REPO, "j1j2-public_amd") in lib[0]


@pytest.mark.parametrize("body, message", [
    ("SAMPLERMODE='XY'", "SAMPLERMODE must be LOCAL or PT. Got XY"),
    ("TEMPPROTOCOL='WARM'", "TEMPPROTOCOL must be CARRY or INDEP. Got WARM"),
    ("SAMPLERMODE='PT', PTLADDERMODE='CUBIC'", "PTLADDERMODE must be LINEAR or SPLIT. Got CUBIC"),
    ("TMIN=10.0, TMAX=9.0", "TMAX must be >= TMIN"),
    ("DT=0.0", "DT must be positive"),
    ("PP=0.1", "dilution"),
    ("UPDATEORDER='RND'", "UPDATEORDER=RND is not supported"),
    ("PTPARALLELMODE='GROUPED'", "PTPARALLELMODE=GROUPED is not supported"),
    ("SAMPLERMODE='PT', NREPLICA=4, PTSWAPEVER=0", "PTSWAPEVER must be positive"),
    ("SAMPLERMODE='PT', NREPLICA=4, PTMEASINDEX=5", "PTMEASINDEX out of range"),
    ("SAMPLERMODE='PT', NREPLICA=4, PTLADDERMODE='SPLIT', TLADDERMID=20.0",
     "TLADDERMID must lie strictly inside"),
])
def test_invalid_options_stop_with_status_1(tmp_path, body, message):
    r = run_host(tmp_path, body + ", WRITEBINDTIME=.false.", timeout=60)
    assert r.returncode == 1
    assert "ERROR" in r.stdout and message in r.stdout, r.stdout


def test_missing_namelist_file(tmp_path):
    r = subprocess.run([HOST, str(tmp_path / "absent.nml")], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "cannot open namelist file" in r.stdout


def compare_files(ours, ref, rtol, atol=0.0):
    """Same lines; text tokens equal; numeric tokens equal to atol + rtol times
    the larger magnitude -- the reference prints with e12.6 / es24.16."""
    a = open(ours).read().splitlines()
    b = open(ref).read().splitlines()
    assert len(a) == len(b), (ours, len(a), len(b))
    worst = 0.0
    for i, (la, lb) in enumerate(zip(a, b)):
        ta, tb = la.split(), lb.split()
        assert len(ta) == len(tb), (ours, i, la, lb)
        for x, y in zip(ta, tb):
            if NUM.match(x) and NUM.match(y):
                fx, fy = float(x.replace("D", "E")), float(y.replace("D", "E"))
                scale = max(abs(fx), abs(fy))
                if scale > 0:
                    worst = max(worst, abs(fx - fy) / scale)
                assert abs(fx - fy) <= atol + rtol * scale, (ours, i, la, lb)
            else:
                assert x == y, (ours, i, la, lb)
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("variant", sorted(REF_RUNS))
def test_host_rewrites_reference_output(tmp_path, golden_dir, variant):
    say(f"case Fortran host vs reference files: {variant}")
    r = run_host(tmp_path, REF_RUNS[variant])
    assert r.returncode == 0, r.stdout + r.stderr
    gd = os.path.join(golden_dir, "ref_" + variant)
    for name in output_files(gd):
        ours = tmp_path / name
        assert ours.exists(), name
        # e12.6 fields (mantissa 0.1 .. 1): one unit in the sixth digit.
        # es24.16 energies: the engine sums the field term in fixed point,
        # q_i = nint(p_i 2^32), so e per site differs from the FP64 sum by at
        # most 4 * 2^-33 (|l_i| <= 4 bonds per site); its std column likewise.
        # locnem files: spins and local nematic values (integers) and
        # stripe magnitudes (integer sums / N): exact to the printed digits.
        if name.startswith("cfen"):
            rtol, atol = 1e-14, 4 * 2.0 ** -33
        elif name.startswith("locnem"):
            rtol, atol = 1e-14, 1e-15
        else:
            rtol, atol = 1.1e-5, 0.0
        worst = compare_files(str(ours), os.path.join(gd, name), rtol, atol)
        say(f"  {name}: worst relative difference {worst:.2e}")


def token_format(tok):
    """How a word of an output file is written: 'I' for an integer, 'E<n>'
    for a real with an exponent and n mantissa digits (e12.6: 6, E18.10: 10,
    es24.16: 17), 'R' for a list-directed real, else the word itself."""
    if not NUM.match(tok):
        return tok
    t = tok.lstrip("+-")
    mant, _, exp = re.split(r"([EeDd])", t, maxsplit=1) if re.search(r"[EeDd]", t) else (t, "", "")
    if not exp and "." not in mant:
        return "I"
    if not exp:
        return "R"
    digits = mant.replace(".", "")
    if mant.startswith("0.") or mant.startswith("."):
        digits = mant.split(".", 1)[1]
    return f"E{len(digits)}"


def compare_layout(ours, ref):
    a = open(ours).read().splitlines()
    b = open(ref).read().splitlines()
    assert len(a) == len(b), (ours, len(a), len(b))
    for i, (la, lb) in enumerate(zip(a, b)):
        ta, tb = la.split(), lb.split()
        assert [token_format(x) for x in ta] == [token_format(y) for y in tb], (ours, i, la, lb)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", sorted(DYN_RUNS))
def test_host_output_layout_under_dynamics(tmp_path, golden_dir, variant):
    """Every file the reference wrote in a run with dynamics, written by the
    host for the same parameters: same line count, the same words, every
    number in the same format (the values differ: the random streams do)."""
    say(f"case Fortran host output layout under dynamics: {variant}")
    r = run_host(tmp_path, DYN_RUNS[variant])
    assert r.returncode == 0, r.stdout + r.stderr
    gd = os.path.join(golden_dir, "ref_" + variant)
    names = output_files(gd)
    assert names
    for name in names:
        assert (tmp_path / name).exists(), name
        compare_layout(str(tmp_path / name), os.path.join(gd, name))
    say(f"  {len(names)} files: {' '.join(names)}")


@pytest.mark.gpu
def test_host_pt_matches_python_binding(tmp_path):
    """PT, L=8, 4 replicas, 3 configs run by the host in batches of 2 and 1;
    the same configs in one Python-driven engine."""
    import j1j2_engine as J
    say("case Fortran host PT dynamics vs Python binding")
    body = ("L=8, H=8, SAMPLERMODE='PT', NREPLICA=4, TMIN=8.0D0, TMAX=11.5D0, START='HOT', NEQ=30, "
            "NMESS=60, NCONF=3, IRINIT=1, DEVSEED=99, BATCH=2, WRITEBINDTIME=.false., WRITENUMFLIPS=.true.")
    r = run_host(tmp_path, body)
    assert r.returncode == 0, r.stdout + r.stderr
    L, R, C = 8, 4, 3
    g = J.make_params(L, L, nreplica=R, neq=30, nmess=60, nconf=C, iconf0=1, seed=99,
                      flags=J.FLAG_WRITELOCNEM | J.FLAG_WRITEHIST | J.FLAG_WRITENUMFLIPS)
    eng = J.Engine([g])
    for c in range(C):
        iconf = 1 + c
        eng.upload_disorder(c, J.make_disorder(L, L, 3.0, iconf, True, 2.0))
        eng.set_spins(c, np.concatenate([J.initial_spins(L, L, J.START_HOT, 0, (iconf - 1) * R + rep + 1)
                                         for rep in range(1, R + 1)]))
    eng.run(J.PHASE_EQUIL, 30)
    eng.run(J.PHASE_MEAS, 60)
    outs = [eng.fetch(c) for c in range(C)]
    flips = np.stack([eng.flips(c) for c in range(C)])      # conf, slot, site
    eng.close()
    bf = np.stack([o["basefields"] for o in outs])          # conf, field, slot
    en = read_table(str(tmp_path / "cfen00080008.dat"))
    np.testing.assert_allclose(en[:, 0], J.ladder(g), rtol=1e-15)
    np.testing.assert_allclose(en[:, 1], bf[:, 9, :].mean(0), rtol=1e-14, atol=1e-15)
    xy = np.loadtxt(str(tmp_path / "cfxy00080008.dat"), skiprows=2)   # its header wraps to 2 lines
    assert xy.shape[0] == C * R
    for k, row in enumerate(xy):
        c, slot = divmod(k, R)
        assert int(row[0]) == c + 1
        np.testing.assert_allclose(row[2:8], bf[c, [14, 15, 33, 34, 35, 36], slot], rtol=6e-6, atol=1e-12)
    pt = open(str(tmp_path / "cfpt00080008.dat")).read().splitlines()
    att = sum(o["swap_att"] for o in outs)
    acc = sum(o["swap_acc"] for o in outs)
    for k in range(R - 1):
        f = pt[1 + k].split()
        assert int(f[3]) == att[k] and int(f[4]) == acc[k]
    nu = np.loadtxt(str(tmp_path / "cfnu00080008.dat"), skiprows=1)
    assert nu.shape[0] == C * R * L ** 3
    np.testing.assert_array_equal(nu[:, 3].reshape(C, R, L ** 3), flips)
    assert os.path.exists(tmp_path / "locnem0001.dat")
    # the phases ran in chunks of a tenth (3 cycles) with the reference's
    # progress reports (Write_progress) and timing lines (Write_timing)
    prog = dict(ln.split(" ", 1) for ln in open(str(tmp_path / "progress.txt")).read().splitlines())
    assert prog["sampler"] == "PT" and prog["phase"] == "PT_MEAS" and prog["iconf"] == "3"
    assert prog["nconf"] == "3" and prog["current"] == "60" and prog["total"] == "60"
    steps = re.findall(r"^progress PT_EQUIL iconf1 step(\d+) of(\d+)$", r.stdout, re.M)
    assert [int(s) for s, _ in steps] == list(range(0, 31, 3)) and {t for _, t in steps} == {"30"}
    timing = re.findall(r"^timing (\S+) rank0 iconf(\d+) seconds +([0-9.]+) configs (\d+)-(\d+)$", r.stdout, re.M)
    assert [(t[0], t[1], t[3], t[4]) for t in timing] == [
        (lab, a, a, b) for a, b in (("1", "2"), ("3", "3"))
        for lab in ("PT_EQUIL", "PT_MEAS", "COLLECT_DATA", "CONFIG_TOTAL")]


@pytest.mark.gpu
@pytest.mark.parametrize("protocol", ["INDEP", "CARRY"])
def test_host_local_protocol_matches_oracle(tmp_path, protocol):
    """LOCAL, HOT start, L = 8, T = 8, 6, 4 (the ladder walked downwards), two
    configs, 20 + 40 sweeps per temperature: INDEP restarts every temperature
    from its own seed, CARRY carries one chain per config down the ladder.
    The oracle runs each config's chain the same way (Philox key (DEVSEED,
    iconf), the sweep clock running on across temperatures); the host's cfen,
    cfxy and locnem files must hold its results."""
    import j1j2_engine as J
    say(f"case Fortran host LOCAL {protocol} vs oracle")
    L, N, NEQ, NMESS, IRINIT, DEVSEED, C = 8, 512, 20, 40, 5, 9, 2
    body = (f"L=8, H=8, SAMPLERMODE='LOCAL', TEMPPROTOCOL='{protocol}', TMIN=4.0D0, TMAX=8.0D0, DT=2.0D0, "
            f"START='HOT', NEQ={NEQ}, NMESS={NMESS}, NCONF={C}, IRINIT={IRINIT}, DEVSEED={DEVSEED}, "
            "WRITEBINDTIME=.false.")
    r = run_host(tmp_path, body)
    assert r.returncode == 0, r.stdout + r.stderr
    temps = temperatures(4.0, 8.0, 2.0, "HOT")
    nt = temps.size
    bf = np.zeros((C, nt, 49))
    loc = []
    for c in range(C):
        iconf = 1 + c
        p = J.make_disorder(L, L, 3.0, IRINIT + iconf - 1, True, 2.0)
        s0 = start_spins(L, L, "HOT", protocol, IRINIT, iconf, nt, 1)
        ora = O.OraclePT(L, L, p, [temps[0]], s0, ptswapever=1, local=True, seed=DEVSEED, iconf=iconf,
                         series_cap=NMESS)
        for it in range(nt):
            if protocol == "INDEP" and it > 0:
                ora.set_spins(start_spins(L, L, "HOT", protocol, IRINIT, iconf, nt, it + 1))
            ora.set_temperatures([temps[it]])
            ora.run(NEQ, False)
            ora.reset_measurement()
            ora.run(NMESS, True)
            bf[c, it] = O.process_local(*ora.series[0], 1.0 / temps[it], N, 100)[0]
            if c == 0:
                loc.append((ora.locspin / ora.s.locsamples, ora.locnem / ora.s.locsamples))
    en = read_table(str(tmp_path / "cfen00080008.dat"))
    np.testing.assert_array_equal(en[:, 0], temps)
    np.testing.assert_allclose(en[:, 1], bf[:, :, 9].mean(0), rtol=1e-14, atol=1e-15)
    xy = np.loadtxt(str(tmp_path / "cfxy00080008.dat"), skiprows=2)
    for k, row in enumerate(xy):
        c, it = divmod(k, nt)
        assert int(row[0]) == c + 1 and row[1] == temps[it]
        np.testing.assert_allclose(row[2:8], bf[c, it, [14, 15, 33, 34, 35, 36]], rtol=6e-6, atol=1e-12)
    for it in range(nt):
        head, spin, nem = read_locnem(str(tmp_path / f"locnem{it + 1:04d}.dat"))
        assert head["Temperature"] == [temps[it]]
        np.testing.assert_allclose(spin, loc[it][0], rtol=1e-10, atol=1e-11)
        np.testing.assert_allclose(nem, loc[it][1], rtol=1e-10, atol=1e-11)
