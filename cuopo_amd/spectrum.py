"""Power spectra of a run's saved round trips: the downstream step cuOPO's
users do with external scripts, from the files a cuOPO / cuopo_sweep run
writes (src/cuOPO.cu:620-635: the saved fields `<field>_output_{r,i}.dat` of
SIZEL samples, the time axis `Tp.dat` and the frequency axis `freq.dat`,
Fp = linspace(-SIZEL dFp/2, SIZEL dFp/2, SIZEL), dFp = 1/T_width,
src/cuOPO.cu:174-180).

    python -m cuopo_amd.spectrum <run folder> [--field signal] [--slot K]

writes `<field>_spectrum.dat` (one value per line, aligned with freq.dat, or
with the slot's own N-point axis for --slot) and prints the peak frequency and
the mean power.  Host-side post-processing (numpy), not part of the solver.
"""
from __future__ import annotations

import argparse
import os

import numpy as np


def read_dat(path):
    """A .dat file of the run, or its binary side-car (the same name with
    .bin: raw little-endian float64, written with --format=bin|both) when the
    .dat file is absent."""
    if not os.path.exists(path) and path.endswith(".dat") and os.path.exists(path[:-4] + ".bin"):
        return np.fromfile(path[:-4] + ".bin", dtype="<f8")
    return np.loadtxt(path, dtype=np.float64)


def read_field(folder, field):
    re = read_dat(os.path.join(folder, f"{field}_output_r.dat"))
    im = read_dat(os.path.join(folder, f"{field}_output_i.dat"))
    return re + 1j * im


def power_spectrum(A):
    """|DFT(A)|^2 / n^2 in centred order (the reference's forward sign,
    sum_j A_j e^{-2 pi i jk/n}); sums to the mean of |A|^2 (Parseval)."""
    n = A.size
    X = np.fft.fftshift(np.fft.fft(A)) / n
    return np.abs(X) ** 2


def slot_axis(freq, n_slot, nslots):
    """Centred frequency axis of one round-trip slot: the record's axis has
    spacing 1/T_width, a slot's spacing is nslots times wider."""
    dF = (freq[-1] - freq[0]) / (freq.size - 1) * nslots
    return (np.arange(n_slot) - n_slot // 2) * dF


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("folder")
    ap.add_argument("--field", default="signal", choices=["signal", "idler", "pump"])
    ap.add_argument("--slot", type=int, default=None, help="spectrum of one saved round trip only")
    ap.add_argument("--N", type=int, default=None, help="samples per round trip (default: from T.dat)")
    a = ap.parse_args(argv)
    A = read_field(a.folder, a.field)
    freq = read_dat(os.path.join(a.folder, "freq.dat"))
    if freq.size != A.size:
        raise SystemExit(f"freq.dat has {freq.size} values, the saved field {A.size}")
    if a.slot is not None:
        n = a.N or read_dat(os.path.join(a.folder, "T.dat")).size
        nslots = A.size // n
        if not 0 <= a.slot < nslots:
            raise SystemExit(f"slot {a.slot} outside 0..{nslots - 1}")
        A = A[a.slot * n:(a.slot + 1) * n]
        axis = slot_axis(freq, n, nslots)
        out = os.path.join(a.folder, f"{a.field}_spectrum_slot{a.slot}.dat")
    else:
        axis = freq
        out = os.path.join(a.folder, f"{a.field}_spectrum.dat")
    S = power_spectrum(A)
    with open(out, "w") as f:
        f.write("".join(f"{v:.20g}\n" for v in S))
    k = int(np.argmax(S))
    print(f"{out}: {S.size} bins; peak at {axis[k]:.9g} THz; mean |A|^2 = {S.sum():.9g}")
    return out


if __name__ == "__main__":
    main()
