"""Absorption-spectrum post-processing of a propagation's dipole output.

A Python 3 restatement of the reference's Plots.py (:17-25): the time series
mu_real(t), mu_imag(t) from H2O.csv are Fourier transformed with the zero
frequency shifted to the centre, the frequency axis is converted to eV
(k / (N dt) * 2 pi * 27.211 for k = -N/2 .. N/2 - 1), and the positive-energy
part (index N/2 + 2 onwards, as the reference plots it) is returned with each
spectrum's real part normalized by its maximum modulus.  The reference's
`range(-N/2, N/2)` only runs under Python 2; integer division is used here.
Plotting is left to the caller.

    python -m python_code_to_check_c_amd.spectrum [H2O.csv] [out.csv]
"""
from __future__ import annotations

import sys

import numpy as np

HARTREE_EV = 27.211  # the constant Plots.py uses


def spectrum(time, mu_real, mu_imag):
    """Return (energy_eV, FT of Im mu, FT of Re mu) over the positive energies:
    np.real of the shifted FFT divided by its max modulus on that range."""
    time = np.asarray(time, dtype=float)
    yr = np.asarray(mu_real, dtype=float)
    yi = np.asarray(mu_imag, dtype=float)
    n = len(time)
    if n < 4:
        raise ValueError("need at least 4 samples")
    dt = time[1] - time[0]
    Yr = np.fft.fftshift(np.fft.fft(yr))
    Yi = np.fft.fftshift(np.fft.fft(yi))
    k = np.arange(-(n // 2), n - n // 2)
    energy = k / (n * dt) * (2 * np.pi) * HARTREE_EV
    lo = n // 2 + 2
    fi = np.real(Yi[lo:]) / np.max(np.abs(Yi[lo:]))
    fr = np.real(Yr[lo:]) / np.max(np.abs(Yr[lo:]))
    return energy[lo:], fi, fr


def main(argv):
    import pandas as pd
    src = argv[1] if len(argv) > 1 else "H2O.csv"
    dst = argv[2] if len(argv) > 2 else "spectrum.csv"
    data = pd.read_csv(src)
    e, fi, fr = spectrum(data["time"], data["mu_real"], data["mu_imag"])
    pd.DataFrame({"energy_eV": e, "FT_mu_imag": fi, "FT_mu_real": fr}).to_csv(dst, index=False)
    print(f"wrote {dst}: {len(e)} points up to {e[-1]:.1f} eV")


if __name__ == "__main__":
    main(sys.argv)
