"""Spectrum post-processing of saved round trips (CPU): a tone written in the
run's file layout comes back at its frequency on freq.dat's axis, the
spectrum obeys Parseval, and a single slot uses its own axis."""
import numpy as np
import pytest

from cuopo_amd import spectrum as SP


def write_run(folder, A, Tp, Fp, N):
    np.savetxt(folder / "signal_output_r.dat", A.real, fmt="%.20g")
    np.savetxt(folder / "signal_output_i.dat", A.imag, fmt="%.20g")
    np.savetxt(folder / "Tp.dat", Tp, fmt="%.20g")
    np.savetxt(folder / "freq.dat", Fp, fmt="%.20g")
    np.savetxt(folder / "T.dat", np.linspace(-0.5, 0.5, N), fmt="%.20g")


@pytest.mark.parametrize("nslots", [4, 64])
def test_tone_peaks_at_its_frequency(tmp_path, nslots):
    N = 256
    SIZEL = N * nslots
    t_rt = 3.3                                   # ps
    T_width = t_rt * nslots                      # src/cuOPO.cu:161-165
    Tp = np.linspace(-0.5 * T_width, 0.5 * T_width, SIZEL)
    dFp = 1 / T_width
    Fp = np.linspace(-0.5 * SIZEL * dFp, 0.5 * SIZEL * dFp, SIZEL)
    k0 = 37
    f0 = k0 / T_width                            # an exact bin of the record
    dt = T_width / SIZEL
    A = 0.7 * np.exp(2j * np.pi * f0 * dt * np.arange(SIZEL))
    write_run(tmp_path, A, Tp, Fp, N)
    out = SP.main([str(tmp_path)])
    S = np.loadtxt(out)
    assert S.size == SIZEL
    k = int(np.argmax(S))
    assert k == SIZEL // 2 + k0
    # freq.dat's axis carries the N/(N-1) stretch of linspace: one bin of tolerance
    assert abs(Fp[k] - f0) <= dFp
    assert S.sum() == pytest.approx(np.mean(np.abs(A) ** 2), rel=1e-12)   # Parseval
    # one slot: N samples, its own axis with spacing nslots/T_width = 1/t_rt
    out1 = SP.main([str(tmp_path), "--slot", "1"])
    S1 = np.loadtxt(out1)
    axis = SP.slot_axis(Fp, N, nslots)
    assert S1.size == N
    assert abs(axis[int(np.argmax(S1))] - f0) <= nslots * dFp
    assert S1.sum() == pytest.approx(0.49, rel=1e-12)


def test_mismatched_files_are_rejected(tmp_path):
    N = 64
    A = np.ones(4 * N, complex)
    write_run(tmp_path, A, np.arange(4 * N), np.arange(3 * N), N)
    with pytest.raises(SystemExit):
        SP.main([str(tmp_path)])
    write_run(tmp_path, A, np.arange(4 * N), np.arange(4 * N), N)
    with pytest.raises(SystemExit):
        SP.main([str(tmp_path), "--slot", "4"])


def test_binary_side_cars_are_read_when_the_text_files_are_absent(tmp_path):
    N, nslots = 64, 4
    SIZEL = N * nslots
    A = np.exp(2j * np.pi * 5 * np.arange(SIZEL) / SIZEL)
    Fp = np.linspace(-0.5, 0.5, SIZEL)
    for name, v in (("signal_output_r", A.real), ("signal_output_i", A.imag), ("freq", Fp),
                    ("T", np.linspace(-0.5, 0.5, N))):
        v.astype("<f8").tofile(tmp_path / f"{name}.bin")
    out = SP.main([str(tmp_path)])
    S = np.loadtxt(out)
    assert int(np.argmax(S)) == SIZEL // 2 + 5
    assert S.sum() == pytest.approx(1.0, rel=1e-12)
