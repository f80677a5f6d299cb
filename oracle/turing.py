"""TEST INFRASTRUCTURE ONLY -- numpy restatement of pyMonteOpt's Turing-pattern bone model.

pymopt/modeling_gpu/_classes.py:25-176 (TuringPattern) grows a trabecular-bone label volume
from a reaction-diffusion system. The diffusion is implicit, solved in Fourier space, and
the reaction is explicit:

    u <- Re F^-1[ ku * Re F[f(u, v)] ],   v <- Re F^-1[ kv * Re F[g(u, v)] ]     (:112-115)

Here f(u, v) = u + dt (0.6 u - v - u^3) and g(u, v) = v + dt (1.5 u - 2 v) (:51-52), and
ku = 1 / Re F[K_u], with K_u the 7-point stencil of I - dt du Laplacian on a periodic grid
of spacing dx (:83-103).

The real part taken of the forward transform (:113-114) is kept as the reference has it.
It makes each step act on the even part of f, i.e. (f(x) + f(-x)) / 2 on the periodic grid.

After the steps come:
* the threshold poly1d(coef)(bv_tv), below or at which u is set to 0 (:128-138);
* the peak count along the last axis, scipy argrelmax of order 5 (:140-160);
* binarisation, u > 0 -> 1 (:162-165).

This is the checker of libpmo's pmo_turing_* entry points.
"""
from __future__ import annotations

import numpy as np

COEF = np.array([-7.67681281, 1.65199492, -0.45314158, 0.60424417])   # _classes.py:45


def stencil(n, d, dt, dx, dtype=np.float32):
    """The reference's kernel array (_classes.py:89-99)."""
    ke = dtype(-dt * d / dx ** 2)
    k = np.zeros((n, n, n), dtype)
    k[0, 0, 0] = dtype(1 + 6 * dt * d / dx ** 2)
    for idx in ((1, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        k[idx] = ke
    return k


def inverse_symbol(n, d, dt, dx):
    """1 / Re F[stencil], as the reference computes it (_classes.py:101-104), float64."""
    return 1.0 / np.fft.fftn(stencil(n, d, dt, dx, np.float64)).real


def inverse_symbol_closed_form(n, d, dt, dx):
    """The same from the stencil's eigenvalues, 1 + 2c (3 - cos a - cos b - cos g), with
    c = dt d / dx^2 and a = 2 pi i / n, etc."""
    c = dt * d / dx ** 2
    t = np.cos(2 * np.pi * np.arange(n) / n)
    lam = 1 + 2 * c * (3 - t[:, None, None] - t[None, :, None] - t[None, None, :])
    return 1.0 / lam


def reaction(u, v, dt, a=(0.6, -1.0, -1.0), b=(1.5, -2.0)):
    """f and g of _classes.py:51-52, written as u + dt (a1 u + a2 v + a3 u^3) and
    v + dt (b1 u + b2 v)."""
    f = u + dt * (a[0] * u + a[1] * v + a[2] * u ** 3)
    g = v + dt * (b[0] * u + b[1] * v)
    return f, g


def step(u, v, ku, kv, dt, a=(0.6, -1.0, -1.0), b=(1.5, -2.0)):
    f, g = reaction(u, v, dt, a, b)
    un = np.real(np.fft.ifftn(ku * np.real(np.fft.fftn(f))))
    vn = np.real(np.fft.ifftn(kv * np.real(np.fft.fftn(g))))
    return un, vn


def initial_fields(n, seed):
    """_get_inital_vector (:117-126): 1 - rand, from np.random seeded with seed and 2*seed."""
    if seed:
        np.random.seed(seed=int(seed))
    ui = 1 - np.random.rand(n, n, n).astype(np.float32)
    if seed:
        np.random.seed(seed=int(seed * 2))
    vi = 1 - np.random.rand(n, n, n).astype(np.float32)
    return ui, vi


def threshold(bv_tv, coef=COEF):
    return float(np.poly1d(coef)(bv_tv))


def count_peaks_last_axis(u, order=5):
    """Strict local maxima along the last axis against `order` neighbours each side, out
    of range neighbours clipped to the edge element (scipy.signal.argrelmax, mode 'clip')."""
    n = u.shape[-1]
    idx = np.arange(n)
    peak = np.ones(u.shape, bool)
    for s in range(1, order + 1):
        for sh in (s, -s):
            j = np.clip(idx + sh, 0, n - 1)
            peak &= u > u[..., j]
    return int(peak.sum())


def grow(n, steps, du, dv, dx, dt, seed, bv_tv, a=(0.6, -1.0, -1.0), b=(1.5, -2.0), dtype=np.float64):
    """The whole model: the field after `steps`, its threshold, BV/TV real, the peak count
    (on the thresholded field) and the binarised labels."""
    u, v = (x.astype(dtype) for x in initial_fields(n, seed))
    ku, kv = inverse_symbol(n, du, dt, dx), inverse_symbol(n, dv, dt, dx)
    for _ in range(steps):
        u, v = step(u, v, ku, kv, dt, a, b)
    thr = threshold(bv_tv)
    t = np.where(u <= thr, 0.0, u)
    return dict(u=u, threshold=thr, bv_tv_real=1 - float((u <= thr).mean()), peaks=count_peaks_last_axis(t),
                labels=(t > 0).astype(np.int8))
