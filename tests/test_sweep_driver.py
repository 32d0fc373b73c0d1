"""The batched sweep driver (cuopo_sweep) against the sweep script
src/cuOPO.sh: the loop nest and folder names (CPU, --dry-run) and, on the
GPU, each point's files against a single cuOPO run of the same point and
against the FP64 oracle (CW and NS, two and three equations)."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT
from tests.helpers import rel_l2

CLI = os.path.join(ROOT, "cuopo_amd", "cli")

# the arrays of src/cuOPO.sh:29,62-88 (two crystals with their own LP/TCR/GP)
SCRIPT_SPEC = """# sweep of src/cuOPO.sh
CRYSTAL = sPPLT PPLN
REG = cw
EQS = 2
PPLN.LP = 1550
PPLN.TCR = 32.88
PPLN.GP = 35.01
sPPLT.LP = 532
sPPLT.TCR = 37.4305
sPPLT.GP = 7.97
NN = 2 3 4 5        # pumping level
RE = 98
D = 0
GD = 100
TD = 0
UPM = 1
MD = 0.5 0.6 0.7 0.8
FM = 4
LCR = 40
"""


def exe(name):
    path = os.path.join(CLI, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (make -C cuopo_amd/cli)")
    return path


def run(args, cwd):
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=600)


def script_points(crystals, arrays):
    """The script's loop nest (cuOPO.sh:91-110: LCR, LP, UPM, NN, MD, FM, D, RE,
    GD, TD, innermost last) and its FOLDER (cuOPO.sh:148-153), restated."""
    out = []
    for cr in crystals:
        a = arrays[cr]
        for lc in a["LCR"]:
            for l in range(len(a["LP"])):
                for u in a["UPM"]:
                    for n in a["NN"]:
                        for m in a["MD"]:
                            for f in a["FM"]:
                                for d in a["D"]:
                                    for r in a["RE"]:
                                        for g in a["GD"]:
                                            for t in a["TD"]:
                                                L = a["LP"][l]
                                                folder = (f"MgO_{cr}_beta_{m}_N_{n}_GDDcomp_{g}_TOD_comp{t}_LP_{L}nm"
                                                          f"_Lcr_{lc}")
                                                out.append((folder, [L, a["TCR"][l], a["GP"][l], n, r, d, g, t, u,
                                                                     m, f, lc]))
    return out


def test_dry_run_matches_the_script_loops(tmp_path):
    spec = tmp_path / "spec.txt"
    spec.write_text(SCRIPT_SPEC)
    r = run([exe("cuopo_sweep"), str(spec), "--dry-run", "--seed=100"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert "32 points" in lines[0]
    common = dict(NN=["2", "3", "4", "5"], RE=["98"], D=["0"], GD=["100"], TD=["0"], UPM=["1"],
                  MD=["0.5", "0.6", "0.7", "0.8"], FM=["4"], LCR=["40"])
    arrays = {"sPPLT": dict(common, LP=["532"], TCR=["37.4305"], GP=["7.97"]),
              "PPLN": dict(common, LP=["1550"], TCR=["32.88"], GP=["35.01"])}
    want = script_points(["sPPLT", "PPLN"], arrays)
    got = [l.split("\t") for l in lines[1:]]
    assert len(got) == len(want)
    for k, (g, (folder, argv)) in enumerate(zip(got, want)):
        assert int(g[0]) == k
        assert g[1] == folder
        assert g[2].split()[1:] == argv
        assert g[3] == f"--seed={100 + k}"


def test_spec_errors_exit_1(tmp_path):
    spec = tmp_path / "bad.txt"
    spec.write_text(SCRIPT_SPEC.replace("MD = 0.5 0.6 0.7 0.8\n", ""))
    r = run([exe("cuopo_sweep"), str(spec), "--dry-run"], tmp_path)
    assert r.returncode == 1 and "Error:" in r.stderr and "MD" in r.stderr
    spec.write_text(SCRIPT_SPEC.replace("sPPLT.GP = 7.97", "sPPLT.GP = 7.97 8.0"))
    r = run([exe("cuopo_sweep"), str(spec), "--dry-run"], tmp_path)
    assert r.returncode == 1 and "same length" in r.stderr


def read_c(stem):
    return np.loadtxt(stem + "_r.dat") + 1j * np.loadtxt(stem + "_i.dat")


@pytest.mark.gpu
@pytest.mark.parametrize("eqs", ["2", "3"])
def test_batched_sweep_equals_single_runs(gpu, tmp_path, eqs):
    N, NZ, NRT = 1024, 4, 70
    spec = tmp_path / "spec.txt"
    spec.write_text(f"""CRYSTAL = PPLN
REG = cw
EQS = {eqs}
LP = 532
TCR = 0
GP = 6.97
NN = 3 5
RE = 98
D = 0.05 0.1
GD = 0
TD = 0
UPM = 0
MD = 0.25
FM = 4
LCR = 5
""")
    sweep_dir = tmp_path / "sweep"
    r = run([exe("cuopo_sweep"), str(spec), f"--out={sweep_dir}", f"--N={N}", f"--nz={NZ}", f"--nrt={NRT}",
             "--seed=40", "--noise=1e-6", "--max-batch=3"], tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    dry = run([exe("cuopo_sweep"), str(spec), "--dry-run", "--seed=40"], tmp_path).stdout.strip().splitlines()[1:]
    # points differing only in D share a folder: the last one's files stay (cuOPO.sh:162-164)
    last = {}
    for line in dry:
        k, folder, args, seed = line.split("\t")
        last[folder] = (args.split()[1:], seed)
    assert len(last) == 2 and len(dry) == 4
    single = "cuOPO_cw_ppln_3eq" if eqs == "3" else "cuOPO_cw_ppln_2eq"
    fields = ("signal", "pump", "idler") if eqs == "3" else ("signal", "pump")
    for folder, (argv, seed) in last.items():
        one = tmp_path / ("one_" + folder)
        one.mkdir()
        r1 = run([exe(single)] + argv + [f"--N={N}", f"--nz={NZ}", f"--nrt={NRT}", seed, "--noise=1e-6"], one)
        assert r1.returncode == 0, r1.stderr[-2000:]
        d = sweep_dir / folder
        assert (d / (folder + ".txt")).exists()
        log = (d / (folder + ".txt")).read_text()
        assert "Power threshold" in log and seed.split("=")[1] in log
        for f in fields:
            for kind in ("input", "output"):
                if kind == "input" and f == "pump":
                    continue
                a, b = read_c(str(d / f"{f}_{kind}")), read_c(str(one / f"{f}_{kind}"))
                assert a.size == b.size
                assert rel_l2(a, b) < 1e-13, (folder, f, kind)
        for name in ("Tp.dat", "freq.dat", "T.dat"):
            assert np.array_equal(np.loadtxt(d / name), np.loadtxt(one / name))


@pytest.mark.gpu
@pytest.mark.parametrize("reg,eqs", [("cw", "3"), ("cw", "2"), ("ns", "2"), ("ns", "3")])
def test_batched_sweep_matches_oracle(gpu, tmp_path, reg, eqs):
    """Every point's output files of one batched cuopo_sweep run against the
    FP64 oracle run on the inputs the driver wrote (src/cuOPO.sh:148-164,
    src/headers/files.cuh:21-75): the saved signal, pump (and idler) slots,
    64 round trips for CW and all NRT round trips (SIZEL = N x NRT values)
    for the nanosecond regime with the EOM and GDD compensation on, and the
    T, Tp and freq grids."""
    from oracle import oracle as O
    N, NZ = 1024, 4
    ns = reg == "ns"
    NRT = 8 if ns else 70
    spec = tmp_path / "spec.txt"
    spec.write_text(f"""CRYSTAL = PPLN
REG = {reg}
EQS = {eqs}
LP = 532
TCR = 0
GP = 6.97
NN = 3 5
RE = 97
D = 0.05
GD = {100 if ns else 0}
TD = 0
UPM = {1 if ns else 0}
MD = 0.25
FM = 4
LCR = 5
""")
    out = tmp_path / "sweep"
    r = run([exe("cuopo_sweep"), str(spec), f"--out={out}", f"--N={N}", f"--nz={NZ}", f"--nrt={NRT}", "--seed=70",
             "--noise=1e-6"], tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    dry = run([exe("cuopo_sweep"), str(spec), "--dry-run", "--seed=70"], tmp_path).stdout.strip().splitlines()[1:]
    assert len(dry) == 2
    three = eqs == "3"
    for line in dry:
        k, folder, args, seed = line.split("\t")
        v = [float(x) for x in args.split()[1:]]
        cfg = O.Config(three_eqs=three, regime=O.NS if ns else O.CW, crystal=O.PPLN, N=N, NZ=NZ, NRT=NRT,
                       lp_nm=v[0], temp_arg=v[1], Lambda=v[2], Nth=v[3], R_pct=v[4], delta=v[5], gdd_pct=v[6],
                       tod_pct=v[7], upm=int(v[8]), md=v[9], fm=v[10], lcr_mm=v[11])
        p = O.setup(cfg)
        T, w = O.grids(p)
        Ap_in = O.pump_input(p)
        d = out / folder
        s0 = int(seed.split("=")[1])
        As = read_c(str(d / "signal_input"))
        assert rel_l2(As, O.noise(s0, 0, 1e-6, N)) < 1e-15
        Ai = read_c(str(d / "idler_input")) if three else None
        assert three or not (d / "idler_input_r.dat").exists()
        sv, _ = O.run(p, w, T, Ap_in, Ap_in[:N].copy(), As, Ai, 0, NRT, save=True)
        slots = NRT if ns else 64
        stems = (("signal_output", "s"), ("pump_output", "p")) + ((("idler_output", "i"),) if three else ())
        for stem, key in stems:
            got = read_c(str(d / stem))
            assert got.size == slots * N == (p.SIZEL if ns else 64 * N)
            assert rel_l2(got, sv[key].reshape(-1)) < 1e-9, (folder, stem)
        Tpo, Fpo = O.grids_long(p)
        assert np.allclose(np.loadtxt(d / "Tp.dat"), Tpo, rtol=1e-15)
        assert np.allclose(np.loadtxt(d / "freq.dat"), Fpo, rtol=1e-15)
        assert np.allclose(np.loadtxt(d / "T.dat"), T, rtol=1e-15)


def test_ns_two_equations_is_accepted(tmp_path):
    """The nanosecond regime with two equations is a build of the reference's
    switches (src/cuOPO.sh:34-45), so the batched driver takes it too."""
    spec = tmp_path / "spec.txt"
    spec.write_text(SCRIPT_SPEC.replace("REG = cw", "REG = ns"))
    r = run([exe("cuopo_sweep"), str(spec), "--dry-run", "--seed=1"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "2 equations, ns" in r.stdout.splitlines()[0]
