"""The drop-in cuOPO executables: argument contract, parameter dump, error
convention (CPU) and the .dat outputs against the oracle (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT
from tests.helpers import rel_l2, to_args

CLI = os.path.join(ROOT, "cuopo_amd", "cli")


def exe(name):
    path = os.path.join(CLI, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (make -C cuopo_amd/cli)")
    return path


def run(args, cwd):
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=600)


def test_usage_on_missing_arguments(tmp_path):
    r = run([exe("cuOPO_cw_ppln_2eq"), "532", "0", "6.97"], tmp_path)
    assert r.returncode == 1
    assert "usage:" in r.stderr


def test_bad_flags_are_usage_errors(tmp_path):
    for bad in ("--format=csv", "--converge-now", "--gdd=3"):
        r = run([exe("cuOPO_cw_ppln_2eq"), "532", "0", "6.97", "4", "98", "0", "0", "0", "0", "0", "0", "5", bad],
                tmp_path)
        assert r.returncode == 1 and "usage:" in r.stderr, bad
    assert "--converge[=tol]" in r.stderr and "--gdd-fs2=" in r.stderr


def test_bad_output_format_is_a_usage_error(tmp_path):
    r = run([exe("cuOPO_cw_ppln_2eq"), "532", "0", "6.97", "4", "98", "0", "0", "0", "0", "0", "0", "5",
             "--format=csv"], tmp_path)
    assert r.returncode == 1
    assert "usage:" in r.stderr and "--format=dat|bin|both" in r.stderr


def test_parameter_dump_and_error_convention(tmp_path):
    """The parameter dump carries the setup scalars; a rejected grid exits 1
    with the reference's 'Error: file:line, code:, reason:' line."""
    r = run([exe("cuOPO_cw_spplt_3eq"), "532", "0", "7.97", "3", "98", "0.1", "0", "0", "0", "0", "0", "20",
             "--N=1000", "--ls=1.0", "--pm=3wave", "--no-inputs"], tmp_path)
    assert r.returncode == 1
    assert "Three equations" in r.stdout and "Regime: continuous wave" in r.stdout
    assert "Error: " in r.stderr and "code: -1" in r.stderr
    cfg = O.Config(three_eqs=True, regime=O.CW, crystal=O.SPPLT, N=1000, NZ=150, NRT=10000, lp_nm=532,
                   Lambda=7.97, Nth=3, R_pct=98, delta=0.1, lcr_mm=20, ls_um=1.0, pm_mode=O.PM_THREE_WAVE)
    p = O.setup(cfg)
    line = [l for l in r.stdout.splitlines() if l.startswith("Power threshold")][0]
    assert float(line.split("=")[1].split()[0]) == pytest.approx(p.Pth, rel=1e-5)
    line = [l for l in r.stdout.splitlines() if l.startswith("Temp ")][0]
    assert float(line.split("=")[1].split()[0]) == pytest.approx(p.Temp, rel=1e-5)


def read_c(stem):
    return np.loadtxt(stem + "_r.dat") + 1j * np.loadtxt(stem + "_i.dat")


# every drop-in build, one per switch combination: (three equations, regime,
# crystal); the sPPLT builds use the sweep script's grating period with a
# 1000 nm signal and the three-wave phase-matching scan
CLI_VARIANTS = {
    "cuOPO_cw_ppln_2eq": (False, O.CW, O.PPLN),
    "cuOPO_cw_spplt_2eq": (False, O.CW, O.SPPLT),
    "cuOPO_cw_ppln_3eq": (True, O.CW, O.PPLN),
    "cuOPO_cw_spplt_3eq": (True, O.CW, O.SPPLT),
    "cuOPO_ns_ppln_2eq": (False, O.NS, O.PPLN),
    "cuOPO_ns_spplt_2eq": (False, O.NS, O.SPPLT),
    "cuOPO_ns_ppln_3eq": (True, O.NS, O.PPLN),
    "cuOPO_ns_spplt_3eq": (True, O.NS, O.SPPLT),
}


@pytest.mark.gpu
@pytest.mark.parametrize("variant", sorted(CLI_VARIANTS))
def test_cli_outputs_match_oracle(gpu, tmp_path, variant):
    three, regime, crystal = CLI_VARIANTS[variant]
    ns = regime == O.NS
    spplt = crystal == O.SPPLT
    N, NZ, NRT = 1024, 4, (8 if ns else 70)
    lam = "7.97" if spplt else "6.97"
    argv = ["532", "0", lam, "4", "98", "0.05", "100" if ns else "0", "0", "1" if ns else "0", "0.25", "4", "5"]
    extra = ["--ls=1.0", "--pm=3wave"] if spplt else []
    r = run([exe(variant)] + argv + [f"--N={N}", f"--nz={NZ}", f"--nrt={NRT}", "--seed=3", "--noise=1e-6"] + extra,
            tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    cfg = O.Config(three_eqs=three, regime=regime, crystal=crystal, N=N, NZ=NZ, NRT=NRT, lp_nm=532,
                   Lambda=float(lam), Nth=4, R_pct=98, delta=0.05, gdd_pct=100 if ns else 0, upm=1 if ns else 0,
                   md=0.25, fm=4, lcr_mm=5)
    if spplt:
        cfg.ls_um, cfg.pm_mode = 1.0, O.PM_THREE_WAVE
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap_in = O.pump_input(p)
    As = read_c(str(tmp_path / "signal_input"))
    assert rel_l2(As, O.noise(3, 0, 1e-6, N)) < 1e-15
    assert rel_l2(read_c(str(tmp_path / "pump_input")), Ap_in) < 1e-15
    if three:
        Ai = read_c(str(tmp_path / "idler_input"))
        assert rel_l2(Ai, O.noise(3, 1, 1e-6, N)) < 1e-15
    else:
        Ai = None
        assert not (tmp_path / "idler_input_r.dat").exists()
    Ap = Ap_in[:N].copy()
    sv, _ = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, NRT, save=True)
    slots = NRT if ns else 64
    stems = (("signal_output", "s"), ("pump_output", "p")) + ((("idler_output", "i"),) if three else ())
    for stem, key in stems:
        out = read_c(str(tmp_path / stem))
        assert out.size == slots * N
        assert rel_l2(out, sv[key].reshape(-1)) < 1e-9, stem
    if not three:
        assert not (tmp_path / "idler_output_r.dat").exists()
    Tp = np.loadtxt(tmp_path / "Tp.dat")
    Tpo, Fpo = O.grids_long(p)
    assert Tp.size == p.SIZEL and np.allclose(Tp, Tpo, rtol=1e-15)
    assert np.allclose(np.loadtxt(tmp_path / "T.dat"), T, rtol=1e-15)
    assert np.allclose(np.loadtxt(tmp_path / "freq.dat"), Fpo, rtol=1e-15)


@pytest.mark.gpu
def test_cli_fp32c_against_fp64_oracle(gpu, tmp_path):
    """--fp32c runs the FP32-complex variant behind the same arguments and
    files: the saved fields stay close to the FP64 oracle, and they differ
    from it by more than FP64 rounding (so the flag really switched the
    arithmetic)."""
    N, NZ, NRT = 1024, 4, 70
    argv = ["532", "0", "6.97", "4", "98", "0.05", "0", "0", "0", "0.25", "4", "5"]
    r = run([exe("cuOPO_cw_ppln_3eq")] + argv + [f"--N={N}", f"--nz={NZ}", f"--nrt={NRT}", "--seed=3",
                                                  "--noise=1e-2", "--fp32c"], tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    cfg = O.Config(three_eqs=True, regime=O.CW, crystal=O.PPLN, N=N, NZ=NZ, NRT=NRT, lp_nm=532, Lambda=6.97,
                   Nth=4, R_pct=98, delta=0.05, md=0.25, fm=4, lcr_mm=5)
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap_in = O.pump_input(p)
    As = read_c(str(tmp_path / "signal_input"))
    Ai = read_c(str(tmp_path / "idler_input"))
    sv, _ = O.run(p, w, T, Ap_in, Ap_in[:N].copy(), As, Ai, 0, NRT, save=True)
    errs = {}
    for stem, key in (("signal_output", "s"), ("pump_output", "p"), ("idler_output", "i")):
        out = read_c(str(tmp_path / stem))
        assert out.size == 64 * N
        errs[key] = rel_l2(out, sv[key].reshape(-1))
    print("FP32C CLI rel L2 over the 64 saved round trips:", errs)  # measured: s, i 1.9e-5; p 2.3e-7
    assert max(errs.values()) < 1e-3
    assert max(errs.values()) > 1e-9


@pytest.mark.gpu
def test_out_directory_receives_every_file(gpu, tmp_path):
    """--out=<dir> puts every output in that folder and nothing in the
    working directory (the reference writes to the working directory)."""
    sub = tmp_path / "run1"
    sub.mkdir()
    argv = ["532", "0", "6.97", "4", "98", "0.05", "0", "0", "0", "0.25", "4", "5",
            "--N=256", "--nz=2", "--nrt=70", "--seed=3", "--noise=1e-6", f"--out={sub}"]
    r = run([exe("cuOPO_cw_ppln_3eq")] + argv, tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    assert sorted(f.name for f in tmp_path.iterdir()) == ["run1"]
    names = sorted(os.listdir(sub))
    assert len(names) == 15 and all(n.endswith(".dat") for n in names)
    assert {"T.dat", "Tp.dat", "freq.dat", "signal_output_r.dat", "idler_input_i.dat"} <= set(names)


@pytest.mark.gpu
def test_binary_side_cars_equal_the_text_files(gpu, tmp_path):
    """--format=both writes X.bin (raw float64) beside every X.dat with the
    same values; --format=bin writes only the .bin files, identical bytes for
    the same seeded run."""
    argv = ["532", "0", "6.97", "4", "98", "0.05", "0", "0", "0", "0.25", "4", "5",
            "--N=256", "--nz=2", "--nrt=70", "--seed=3", "--noise=1e-6"]
    both, only = tmp_path / "both", tmp_path / "bin"
    both.mkdir(); only.mkdir()
    r = run([exe("cuOPO_cw_ppln_3eq")] + argv + ["--format=both"], both)
    assert r.returncode == 0, r.stderr[-2000:]
    dats = sorted(f for f in os.listdir(both) if f.endswith(".dat"))
    assert len(dats) == 15  # three inputs, three outputs (re/im each), Tp, freq, T
    for f in dats:
        txt = np.loadtxt(both / f)
        binv = np.fromfile(both / (f[:-4] + ".bin"), dtype="<f8")
        assert binv.shape == txt.shape and np.array_equal(binv, txt), f
    r = run([exe("cuOPO_cw_ppln_3eq")] + argv + ["--format=bin"], only)
    assert r.returncode == 0, r.stderr[-2000:]
    assert not [f for f in os.listdir(only) if f.endswith(".dat")]
    for f in dats:
        b = f[:-4] + ".bin"
        assert (only / b).read_bytes() == (both / b).read_bytes(), b


@pytest.mark.gpu
def test_absolute_gdd_equals_its_percentage(gpu, tmp_path):
    """--gdd-fs2 sets GDD = -eps beta2_s Lcr through the percentage that gives
    it: the outputs equal a run given that percentage as argv7."""
    base = ["532", "0", "6.97", "4", "98", "0.05"]
    tail = ["0", "0", "0.25", "4", "5", "--N=256", "--nz=2", "--nrt=70", "--seed=3", "--noise=1e-6",
            "--format=bin", "--no-inputs"]
    cfg = O.Config(three_eqs=True, regime=O.CW, crystal=O.PPLN, N=256, NZ=2, NRT=70, lp_nm=532, Lambda=6.97,
                   Nth=4, R_pct=98, delta=0.05, lcr_mm=5)
    p = O.setup(cfg)
    gdd_fs2 = 300.0
    pct = -100.0 * gdd_fs2 * 1e-6 / (p.b2s * p.Lcr)
    a, b = tmp_path / "fs2", tmp_path / "pct"
    a.mkdir(); b.mkdir()
    r = run([exe("cuOPO_cw_ppln_3eq")] + base + ["0"] + tail + [f"--gdd-fs2={gdd_fs2}"], a)
    assert r.returncode == 0, r.stderr[-2000:]
    r = run([exe("cuOPO_cw_ppln_3eq")] + base + [f"{pct:.17g}"] + tail, b)
    assert r.returncode == 0, r.stderr[-2000:]
    for f in ("signal_output_r.bin", "signal_output_i.bin", "idler_output_r.bin"):
        x, y = np.fromfile(a / f, "<f8"), np.fromfile(b / f, "<f8")
        assert np.array_equal(x, y), f
    # and the GDD really acts: without it the signal differs
    c = tmp_path / "none"
    c.mkdir()
    r = run([exe("cuOPO_cw_ppln_3eq")] + base + ["0"] + tail, c)
    assert r.returncode == 0
    assert not np.array_equal(np.fromfile(c / "signal_output_r.bin", "<f8"),
                              np.fromfile(a / "signal_output_r.bin", "<f8"))


@pytest.mark.gpu
def test_converge_stops_then_saves_the_tail(gpu, tmp_path):
    """--converge=1e-3 (CW): round trips in chunks of 50 until the signal
    power changes by < 1e-3 over 50 of them, then the 64 round trips the
    reference saves.  The same schedule through the library (the same untuned
    kernels, deterministic) must stop at the same round trip and give
    identical saved slots; the library itself is held to the oracle by the
    parity tests (this multimode regime amplifies any rounding difference, so
    the oracle's own stopping round trip is not a stable reference here)."""
    N, NZ, NRT, tol = 256, 4, 2000, 1e-3
    argv = ["532", "0", "6.97", "2", "90", "0", "0", "0", "0", "0", "0", "5", f"--N={N}", f"--nz={NZ}",
            f"--nrt={NRT}", "--seed=3", "--noise=1e-3", f"--converge={tol}", "--format=bin"]
    r = run([exe("cuOPO_cw_ppln_3eq")] + argv, tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if "teady after" in l]
    assert len(line) == 1 and line[0].startswith("Steady after "), line
    K = int(line[0].split()[2])
    cfg = O.Config(three_eqs=True, regime=O.CW, crystal=O.PPLN, N=N, NZ=NZ, NRT=NRT, lp_nm=532, Lambda=6.97,
                   Nth=2, R_pct=90, delta=0.0, lcr_mm=5)
    p = O.setup(cfg)
    T, w = O.grids(p)
    # the executable's own inputs (its noise generator and the oracle's agree to ~1e-16, and this
    # regime amplifies that)
    rb = lambda stem: (np.fromfile(tmp_path / f"{stem}_r.bin", "<f8")
                       + 1j * np.fromfile(tmp_path / f"{stem}_i.bin", "<f8"))
    Ap_in, As, Ai = rb("pump_input"), rb("signal_input"), rb("idler_input")
    P = gpu
    s = P.Solver(P.params_from(P.derive(to_args(cfg)), 0))
    s.upload(w, T, Ap_in, Ap_in[:N].copy(), As, Ai)
    s.enable_save(True)
    nn, prev = 0, -1.0
    while nn < NRT - 64:
        s.run(nn, 50)
        P_end = s.power_history(50)[-1]
        nn += 50
        if prev > 0 and abs(P_end - prev) <= tol * P_end:
            break
        prev = P_end
    assert 100 <= K < NRT - 64 and nn == K
    s.run(NRT - 64, 64)
    sv = s.download_saved()
    s.close()
    for stem, ref in (("signal_output", sv["s"]), ("idler_output", sv["i"]), ("pump_output", sv["p"])):
        out = np.fromfile(tmp_path / f"{stem}_r.bin", "<f8") + 1j * np.fromfile(tmp_path / f"{stem}_i.bin", "<f8")
        assert np.array_equal(out, np.asarray(ref).reshape(-1)), stem
