"""Pins the Turing-pattern oracle (oracle/turing.py): the implicit-diffusion symbol against
the stencil's closed-form eigenvalues, one step on small single Fourier modes against the
linearised 2x2 map, the real-part quirk (odd modes vanish in one step), the peak count
against scipy's argrelmax, and the threshold polynomial at a hand-computed point."""
import numpy as np
import pytest

from oracle import turing as tp


def test_inverse_symbol_is_the_stencil_eigenvalue():
    for n, d in ((12, 0.0002), (16, 0.01), (9, 0.003)):
        a = tp.inverse_symbol(n, d, 1.0, 1 / 40)
        b = tp.inverse_symbol_closed_form(n, d, 1.0, 1 / 40)
        assert np.allclose(a, b, rtol=1e-12)
        assert a[0, 0, 0] == pytest.approx(1.0)


@pytest.mark.parametrize("mode", [(1, 0, 0), (2, 3, 1), (0, 0, 5)])
def test_small_even_mode_follows_the_linear_map(mode):
    # around (0, 0): [u, v] -> diag(ku(k), kv(k)) [[1 + 0.6 dt, -dt], [1.5 dt, 1 - 2 dt]] [u, v]
    n, dt, dx, du, dv = 16, 1.0, 1 / 40, 0.0002, 0.01
    x = np.arange(n)
    ph = 2 * np.pi * (mode[0] * x[:, None, None] + mode[1] * x[None, :, None] + mode[2] * x[None, None, :]) / n
    eps = 1e-5
    u0, v0 = eps * np.cos(ph), 0.5 * eps * np.cos(ph)
    ku, kv = tp.inverse_symbol(n, du, dt, dx), tp.inverse_symbol(n, dv, dt, dx)
    u1, v1 = tp.step(u0, v0, ku, kv, dt)
    lu, lv = ku[mode], kv[mode]
    J = np.array([[1 + 0.6 * dt, -dt], [1.5 * dt, 1 - 2 * dt]])
    a = np.array([lu, lv]) * (J @ np.array([eps, 0.5 * eps]))
    assert np.allclose(u1, a[0] * np.cos(ph), atol=1e-13)
    assert np.allclose(v1, a[1] * np.cos(ph), atol=1e-13)


def test_real_part_of_the_spectrum_removes_odd_modes():
    n = 16
    x = np.arange(n)
    ph = 2 * np.pi * (2 * x[:, None, None] + x[None, None, :]) / n
    ku = tp.inverse_symbol(n, 0.0002, 1.0, 1 / 40)
    kv = tp.inverse_symbol(n, 0.01, 1.0, 1 / 40)
    u1, v1 = tp.step(1e-5 * np.sin(ph), 0 * ph, ku, kv, 1.0)
    assert np.abs(u1).max() < 1e-18 and np.abs(v1).max() < 1e-18
    # in general a step keeps only the even part: the result is symmetric under x -> -x
    rng = np.random.default_rng(0)
    u, v = rng.uniform(0, 1, (n, n, n)), rng.uniform(0, 1, (n, n, n))
    u1, v1 = tp.step(u, v, ku, kv, 1.0)
    flip = lambda a: np.roll(a[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))
    assert np.allclose(u1, flip(u1), atol=1e-12) and np.allclose(v1, flip(v1), atol=1e-12)


def test_peak_count_matches_scipy_argrelmax():
    from scipy.signal import argrelmax
    rng = np.random.default_rng(3)
    u = rng.uniform(0, 1, (6, 7, 40))
    u[u < 0.4] = 0
    want = sum(argrelmax(row, order=5)[0].shape[0] for sl in u for row in sl)
    assert tp.count_peaks_last_axis(u, 5) == want
    assert tp.count_peaks_last_axis(np.zeros((2, 2, 20))) == 0


def test_threshold_polynomial_at_the_default_bv_tv():
    x = 0.138
    want = -7.67681281 * x ** 3 + 1.65199492 * x ** 2 - 0.45314158 * x + 0.60424417
    assert tp.threshold(x) == pytest.approx(want, rel=1e-14)
    # by hand: x^2 = 0.019044, x^3 = 0.0026281 -> 0.60424 - 0.06253 + 0.03146 - 0.02018
    assert tp.threshold(x) == pytest.approx(0.55300, abs=1e-4)
