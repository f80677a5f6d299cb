"""Exit-state detectors on the host, for the optional per-packet dump of get_result().

Same definitions as pymopt/metrics/_scoring.py:16-51; the walk accumulates the same
quantities on the device (Rd_r, Rd_a, line_z), these serve code that post-processes dumps.
"""
from __future__ import annotations

import numpy as np

__all__ = ["radialDistance", "angularyResolved", "lineDistance"]


def _binned(x, w, edges):
    """Sum of w over bins (edges[i], edges[i+1]]."""
    idx = np.searchsorted(edges, x, side="left") - 1
    ok = (idx >= 0) & (idx < edges.size - 1) & (x > edges[0])
    return np.bincount(idx[ok], weights=np.asarray(w, np.float64)[ok], minlength=edges.size - 1)


def radialDistance(p, w, nn, dr, nPh):
    """Rd(r) in annuli (i dr, (i+1) dr] per unit area and packet (_scoring.py:16-25)."""
    alpha = np.array([i * dr for i in range(nn + 1)])
    da = np.array([2 * np.pi * (i + 0.5) * dr ** 2 for i in range(nn)])
    r = np.sqrt(p[0] ** 2 + p[1] ** 2)
    return alpha[:-1], _binned(r, w, alpha) / (da * nPh)


def angularyResolved(v, w, nn, nPh):
    """Rd(alpha) per steradian and packet (_scoring.py:27-38)."""
    da = np.pi / (2 * nn)
    alpha = np.array([(i + 0.5) * da for i in range(nn + 1)])
    alpha2 = np.array([i * da for i in range(nn + 1)])
    do = 4 * np.pi * np.sin(alpha) * np.sin(da / 2)
    at = np.arccos(np.sign(v[2]) * v[2])
    return alpha[:-1], _binned(at, w, alpha2) / (do[:-1] * nPh)


def lineDistance(p, w, nn, dr, nPh, y_range=5):
    """Side-face line profile over z for |y| < y_range (_scoring.py:40-51)."""
    alpha = np.array([i * dr for i in range(nn + 1)])
    da = np.ones(nn) * dr * y_range * 2
    sel = np.abs(p[1]) < y_range
    return alpha[:-1], _binned(p[2][sel], np.asarray(w)[sel], alpha) / (da * nPh)
