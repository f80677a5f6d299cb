"""Host restatements of the uniform-cube radius (pmo_skip_radius) for the tests: the
separable form the device computes, and a brute-force search that defines it."""
from __future__ import annotations

import numpy as np


def boundary(labels):
    """True on the outer face, where one of the 26 neighbours holds another label, and on
    label 255 (the device's marker of a mixed neighbourhood)."""
    lab = np.asarray(labels)
    b = np.ones(lab.shape, bool)
    c = lab[1:-1, 1:-1, 1:-1]
    inner = np.zeros(c.shape, bool)
    nx, ny, nz = lab.shape
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                inner |= lab[1 + dx:nx - 1 + dx, 1 + dy:ny - 1 + dy, 1 + dz:nz - 1 + dz] != c
    b[1:-1, 1:-1, 1:-1] = inner
    return b | (lab == 255)


def skip_radius(labels, cap):
    """min(cap, Chebyshev distance to the boundary set), by three windowed min-max passes."""
    g = np.where(boundary(labels), 0, cap + 1).astype(np.int32)
    for axis in (2, 1, 0):
        n = g.shape[axis]
        out = g.copy()
        for d in range(1, cap + 1):
            if d >= n:
                break
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[axis], hi[axis] = slice(0, n - d), slice(d, n)
            # out[i] = min(out[i], max(d, g[i - d])) and min(out[i], max(d, g[i + d]))
            out[tuple(hi)] = np.minimum(out[tuple(hi)], np.maximum(d, g[tuple(lo)]))
            out[tuple(lo)] = np.minimum(out[tuple(lo)], np.maximum(d, g[tuple(hi)]))
        g = out
    return np.minimum(g, cap).astype(np.uint8)


def skip_radius_brute(labels, cap):
    """The largest k <= cap whose cube of half-width k around the voxel lies inside the
    volume and holds one label (0 on the outer face and on label 255)."""
    lab = np.asarray(labels)
    out = np.zeros(lab.shape, np.uint8)
    nx, ny, nz = lab.shape
    for x in range(nx):
        for y in range(ny):
            for z in range(nz):
                k = 0
                while k < cap and lab[x, y, z] != 255:
                    j = k + 1
                    if x - j < 0 or y - j < 0 or z - j < 0 or x + j >= nx or y + j >= ny or z + j >= nz:
                        break
                    if np.any(lab[x - j:x + j + 1, y - j:y + j + 1, z - j:z + j + 1] != lab[x, y, z]):
                        break
                    k = j
                out[x, y, z] = k
    return out


def blob_volume(shape, n_labels, seed, smooth=2):
    """Labels in blobs: a smoothed random field cut into n_labels bands."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape)
    for axis in range(3):
        for _ in range(smooth):
            f = (np.roll(f, 1, axis) + f + np.roll(f, -1, axis)) / 3
    q = np.quantile(f, np.linspace(0, 1, n_labels + 1)[1:-1])
    return np.searchsorted(q, f).astype(np.uint8)
