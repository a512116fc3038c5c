"""Spectrum post-processing (port of the reference's Plots.py :17-25)."""
import numpy as np

from python_code_to_check_c_amd.spectrum import HARTREE_EV, spectrum


def test_single_line_lands_at_its_energy():
    # a damped oscillation at omega = 0.5 Eh: the FT of Im mu peaks at 0.5 * 27.211 eV
    dt, n, omega = 0.01, 4096, 0.5
    t = dt * np.arange(1, n + 1)
    yi = np.sin(omega * t) * np.exp(-0.01 * t)
    yr = 1.0 + 0.1 * np.cos(omega * t) * np.exp(-0.01 * t)
    e, fi, fr = spectrum(t, yr, yi)
    de = 2 * np.pi * HARTREE_EV / (n * dt)
    assert abs(e[np.argmax(np.abs(fi))] - omega * HARTREE_EV) <= de
    assert abs(e[np.argmax(np.abs(fr))] - omega * HARTREE_EV) <= de
    assert np.max(np.abs(fi)) <= 1.0 + 1e-15


def test_axis_and_normalization_follow_the_reference_formula():
    rng = np.random.default_rng(0)
    n, dt = 64, 0.02
    t = dt * np.arange(n)
    yr, yi = rng.standard_normal(n), rng.standard_normal(n)
    e, fi, fr = spectrum(t, yr, yi)
    # the reference: xaxis = k / (N dt) * 2 pi * 27.211 for k in range(-N/2, N/2), plotted from N/2 + 2
    k = np.arange(-n // 2, n // 2)
    assert np.allclose(e, (k / (n * dt) * 2 * np.pi * 27.211)[n // 2 + 2:], rtol=0, atol=1e-12)
    Yi = np.fft.fftshift(np.fft.fft(yi))[n // 2 + 2:]
    assert np.allclose(fi, Yi.real / np.abs(Yi).max(), rtol=0, atol=1e-14)
    # a constant-free check independent of numpy's FFT: the DFT sum at k = 5
    k5 = 5
    direct = np.sum(yr * np.exp(-2j * np.pi * k5 * np.arange(n) / n))
    Yr = np.fft.fftshift(np.fft.fft(yr))
    assert abs(Yr[n // 2 + k5] - direct) <= 1e-12
