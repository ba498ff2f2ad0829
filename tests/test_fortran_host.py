"""The Fortran host (j1j2-public_amd/fortran): the reference's driver over the
engine's C-ABI through iso_c_binding.

CPU: the host is built and linked against the in-tree library; invalid
options end the run with the reference's message and exit status 1
(Validate_run_options, J1J2_6tv.f90:841-925) before any device call.

GPU: (1) with NEQ = NMESS = 0 the reference's output is fixed by its disorder
and HOT/STR starts alone, so the host must rewrite the reference's own files
(tests/golden/ref_*: cfen, cfma, cfnm, cfst, cfhs, cfxy, cfnd, cfpt and
phi_sites.dat) line for line.  (2) With dynamics, the host's files must carry
exactly what the engine returns through the Python binding for the same
parameters (batching across engine launches must not change a result, since
the device RNG is keyed by iconf).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from gpu_cases import say
from refdata import find, read_table

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "j1j2-public_amd", "j1j2_host")

# Parameter blocks of the reference runs kept under tests/golden/ (the
# variants built by oracle/ref_build.sh).
REF_RUNS = {
    "hot_init_l8": "L=8, H=8, SAMPLERMODE='PT', NREPLICA=4, TMIN=8.0D0, TMAX=11.5D0, START='HOT', "
                   "NEQ=0, NMESS=0, NCONF=3, IRINIT=1, WRITEBINDTIME=.false.",
    "hot_init_l6h4": "L=6, H=4, SAMPLERMODE='PT', NREPLICA=2, TMIN=8.0D0, TMAX=11.5D0, START='HOT', "
                     "NEQ=0, NMESS=0, NCONF=2, IRINIT=7, WRITEBINDTIME=.false.",
    "str_init_l12": "L=12, H=10, SAMPLERMODE='PT', NREPLICA=2, TMIN=8.0D0, TMAX=11.5D0, START='STR', "
                    "STRAXIS='Y', NEQ=0, NMESS=0, NCONF=1, IRINIT=11, WRITEBINDTIME=.false.",
}

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
    for name in sorted(os.listdir(gd)):
        if not (name.startswith("cf") or name == "phi_sites.dat"):
            continue
        ours = tmp_path / name
        assert ours.exists(), name
        # e12.6 fields (mantissa 0.1 .. 1): one unit in the sixth digit.
        # es24.16 energies: the engine sums the field term in fixed point,
        # q_i = nint(p_i 2^32), so e per site differs from the FP64 sum by at
        # most 4 * 2^-33 (|l_i| <= 4 bonds per site); its std column likewise.
        if name.startswith("cfen"):
            rtol, atol = 1e-14, 4 * 2.0 ** -33
        else:
            rtol, atol = 1.1e-5, 0.0
        worst = compare_files(str(ours), os.path.join(gd, name), rtol, atol)
        say(f"  {name}: worst relative difference {worst:.2e}")


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
def test_host_local_carry_matches_python_binding(tmp_path):
    """LOCAL, CARRY protocol, HOT start: the host walks T = 11, 10, 9 on one
    chain per config (2 configs); the Python binding repeats the sequence."""
    say("case Fortran host LOCAL CARRY vs Python binding")
    L = 8
    N = L * L * L
    body = ("L=8, H=8, SAMPLERMODE='LOCAL', TEMPPROTOCOL='CARRY', TMIN=9.0D0, TMAX=11.0D0, DT=1.0D0, "
            "START='HOT', NEQ=20, NMESS=40, NCONF=2, IRINIT=3, DEVSEED=5, WRITEBINDTIME=.false.")
    r = run_host(tmp_path, body)
    assert r.returncode == 0, r.stdout + r.stderr
    import j1j2_engine as J
    g = J.make_params(L, L, nreplica=1, ptswapever=1, sampler=J.SAMPLER_LOCAL, ladder=J.LADDER_LOCAL,
                      tmin=11.0, tmax=11.0, dt=1.0, neq=20, nmess=40, nconf=2, iconf0=1, seed=5, flags=1 | 2)
    eng = J.Engine([g])
    for c in range(2):
        iconf = 1 + c
        eng.upload_disorder(c, J.make_disorder(L, L, 3.0, 3 + iconf - 1, True, 2.0))
        eng.set_spins(c, J.initial_spins(L, L, J.START_HOT, 0, 3 + iconf - 1, skip=3 * N))
    temps = [11.0, 10.0, 9.0]                   # HOT start: the ladder walked downwards
    bf = np.zeros((2, 3, 49))
    for it, T in enumerate(temps):
        for c in range(2):
            eng.set_temperatures(c, np.array([T]))
        eng.run(J.PHASE_EQUIL, 20)
        eng.run(J.PHASE_MEAS, 40)
        for c in range(2):
            bf[c, it] = eng.fetch(c)["basefields"][:, 0]
    eng.close()
    en = read_table(str(tmp_path / "cfen00080008.dat"))
    np.testing.assert_allclose(en[:, 0], temps, rtol=1e-15)
    np.testing.assert_allclose(en[:, 1], bf[:, :, 9].mean(0), rtol=1e-14, atol=1e-15)
    ma = read_table(str(tmp_path / "cfma00080008.dat"))
    np.testing.assert_allclose(ma[:, 1], bf[:, :, 0].mean(0), rtol=6e-6)
    assert os.path.exists(tmp_path / "locnem0001.dat")
