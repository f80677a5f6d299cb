"""Geometry store: label volumes, optics tables and launch sources.

A volume is a C-order uint8 array [nx][ny][nz] (z fastest) whose 1-voxel outer shell holds
the end label (= number of tissue labels). The optics table is float32 [n_labels+1][4] of
(mu_a /mm, mu_s /mm, g, n); the last row holds the ambient index. This is the layout the
HIP walk reads: labels as bytes through L2, the table copied to LDS.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _abi


@dataclass
class Source:
    """Launch description (reference BaseVoxelMonteCarlo._generate_initial_coodinate,
    pymopt/voxel/_classes.py:107-213)."""
    voxel: tuple            # (ix, iy, iz) launch voxel, the beam centre
    w0: float               # launch weight 1 - Rsp
    kind: int = _abi.PMO_BEAM_PENCIL
    radius: float = 0.0     # flat-disc radius / Gaussian waist (mm)
    dir: tuple = (0.0, 0.0, 1.0)
    # per-packet launch state (add, p, v, w) of the reference kernels: add, p, v (3, N) and
    # w (N,), host arrays or device tensors; `voxel` then only centres the LDS tally window
    states: tuple = None

    def as_dict(self):
        d = dict(kind=self.kind, voxel=tuple(int(x) for x in self.voxel), w0=float(self.w0),
                 radius=float(self.radius), dir=tuple(float(x) for x in self.dir))
        if self.states is not None:
            d["states"] = tuple(x.cpu().numpy() if hasattr(x, "cpu") else x for x in self.states)
        return d


def center_add(length: int) -> int:
    """Voxel index of the beam centre along an axis of `length` padded voxels
    (pymopt/voxel/_classes.py:124-128)."""
    return int((length - 1) / 2)


def specular(n_ambient: float, n_tissue: float) -> float:
    """Rsp at normal incidence (pymopt/voxel/_classes.py:163-175)."""
    if n_ambient == n_tissue:
        return 0.0
    return ((n_ambient - n_tissue) / (n_ambient + n_tissue)) ** 2


def fresnel_oblique(ai: float, n1: float, n2: float):
    """Refraction angle and reflectance of the oblique launch
    (pymopt/voxel/_classes.py:139-147, :178-181)."""
    at = float(np.arcsin(np.sin(ai) * n1 / n2))
    r = ((np.sin(ai - at) / np.sin(ai + at)) ** 2 + (np.tan(ai - at) / np.tan(ai + at)) ** 2) / 2
    return at, float(r)


def plate_volume(thickness, voxel_space, x_size, y_size):
    """Layered plate label volume, pymopt/voxel/_classes.py:372-415 (PlateModel).

    Layer i fills round(th_i / l) z-slices from z = 1; the outer shell is the end label
    len(thickness). The reference leaves uninitialised slices when the per-layer rounding
    and the total rounding disagree; here that is an error instead.
    """
    l = float(voxel_space)
    nx = int(np.int32(round(x_size / l)))
    ny = int(np.int32(round(y_size / l)))
    nz = int(np.int32(round(float(np.array(thickness).sum()) / l)))
    layers = [int(round(t / l)) for t in thickness]
    if sum(layers) != nz:
        raise ValueError(
            f"layer voxel counts {layers} (sum {sum(layers)}) differ from round(total/l) = {nz}; "
            "choose thicknesses that are multiples of voxel_space")
    if nx < 1 or ny < 1 or nz < 1:
        raise ValueError("plate must be at least one voxel in each direction")
    end = len(thickness)
    vol = np.empty((nx + 2, ny + 2, nz + 2), dtype=np.uint8)
    z = 1
    for i, k in enumerate(layers):
        vol[:, :, z:z + k] = i
        z += k
    shell(vol, end)
    return vol


def shell(vol: np.ndarray, end: int) -> np.ndarray:
    vol[0, :, :] = end
    vol[-1, :, :] = end
    vol[:, 0, :] = end
    vol[:, -1, :] = end
    vol[:, :, 0] = end
    vol[:, :, -1] = end
    return vol


def optics_table(ma, ms, g, n, n_ambient) -> np.ndarray:
    """float32 [n_labels+1][4] table (mu_a, mu_s, g, n); row n_labels holds n_ambient."""
    ma = np.asarray(ma, np.float64).reshape(-1)
    ms = np.asarray(ms, np.float64).reshape(-1)
    g = np.asarray(g, np.float64).reshape(-1)
    n = np.asarray(n, np.float64).reshape(-1)
    L = ma.size
    if not (ms.size == L and g.size == L and n.size == L):
        raise ValueError("ma, ms, g and n need one entry per layer/label")
    if np.any(ma < 0) or np.any(ms < 0):
        raise ValueError("mu_a and mu_s must be >= 0")
    if np.any(ma + ms <= 0):
        raise ValueError("every label needs mu_a + mu_s > 0")
    if np.any(np.abs(g) >= 1):
        raise ValueError("|g| must be < 1")
    if np.any(n <= 0) or n_ambient <= 0:
        raise ValueError("refractive indices must be > 0")
    t = np.zeros((L + 1, 4), np.float32)
    t[:L, 0] = ma
    t[:L, 1] = ms
    t[:L, 2] = g
    t[:L, 3] = n
    t[L, 3] = n_ambient
    return t


def validate_volume(vol: np.ndarray, n_labels: int, shell: bool = True):
    """Checks the contract of the walk: uint8, C order, labels <= n_labels, end-label shell
    (shell=False: the type, shape and n_labels checks only)."""
    if vol.dtype != np.uint8 or vol.ndim != 3:
        raise ValueError("label volume must be a 3-D uint8 array")
    if min(vol.shape) < 3:
        raise ValueError("label volume needs at least 3 voxels per axis (1-voxel end shell)")
    if not (1 <= n_labels <= 255):
        raise ValueError("n_labels must be in [1, 255]")
    if not shell:
        return
    if int(vol.max()) > n_labels:
        raise ValueError(f"label {int(vol.max())} exceeds the end label {n_labels}")
    end = n_labels
    faces = [vol[0], vol[-1], vol[:, 0], vol[:, -1], vol[:, :, 0], vol[:, :, -1]]
    if any(np.any(f != end) for f in faces):
        raise ValueError("the outer 1-voxel shell of the volume must hold the end label")


# ---------------------------------------------------------------- synthetic phantoms

def gaussian_random_field(shape, pitch, corr_len, seed):
    """Zero-mean unit-variance 3-D Gaussian random field: white noise filtered by a
    Gaussian kernel exp(-r^2 / (2 corr_len^2)) in Fourier space."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    k = [np.fft.fftfreq(s, d=pitch) * 2 * np.pi for s in shape]
    kx, ky, kz = np.meshgrid(*k, indexing="ij", sparse=True)
    filt = np.exp(-0.5 * (kx ** 2 + ky ** 2 + kz ** 2) * corr_len ** 2)
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    f -= f.mean()
    f /= f.std()
    return f


def bone_phantom(seed=0, n=256, pitch=0.02, corr_len=0.15, porosity=0.7, cortical=0.5, soft=1.0):
    """Synthetic cancellous-bone phantom on an n^3 grid (labels 0 bone, 1 marrow,
    2 cortical, 3 soft tissue, 4 end).

    Inside the n^3 volume: a soft-tissue top layer of `soft` mm, then a cortical shell of
    `cortical` mm around (sides, bottom, top) a trabecular core whose voxels are marrow
    where a GRF drawn on the core is below its `porosity` quantile and bone elsewhere. The
    outer 1-voxel shell of the n^3 grid is the end label, so the returned array is n^3.
    """
    inner = n - 2
    ns = int(round(soft / pitch))
    nc = int(round(cortical / pitch))
    if inner - 2 * nc < 1 or inner - ns - 2 * nc < 1:
        raise ValueError("phantom too small for its cortical shell and soft-tissue layer")
    vol = np.empty((n, n, n), np.uint8)
    interior = vol[1:-1, 1:-1, 1:-1]
    interior[:] = 2                                            # cortical everywhere first
    interior[:, :, :ns] = 3                                    # soft-tissue top layer
    core_shape = (inner - 2 * nc, inner - 2 * nc, inner - ns - 2 * nc)
    f = gaussian_random_field(core_shape, pitch, corr_len, seed)
    thr = np.quantile(f, porosity)
    # trabecular core: marrow (1) where the field is below its porosity quantile, bone (0) above
    interior[nc:inner - nc, nc:inner - nc, ns + nc:inner - nc] = np.where(f < thr, 1, 0).astype(np.uint8)
    shell(vol, 4)
    return vol


def bone_phantom_wide(seed=0, n=256, tiles=5, pitch=0.02, corr_len=0.15, porosity=0.7, cortical=0.5, soft=1.0):
    """The bone phantom of `bone_phantom(seed, n, ...)` widened laterally: its trabecular
    core, a GRF drawn on a periodic grid, is tiled `tiles` x `tiles` times across x and y
    (seamlessly), with the same soft-tissue top layer, cortical shell and depth. For odd
    `tiles` the centre voxel falls on the same core cell as in the n^3 phantom. Returns
    labels of shape (tiles * core + 2 * cortical + 2, same, n)."""
    if tiles < 1 or tiles % 2 == 0:
        raise ValueError("tiles must be odd and >= 1")
    inner = n - 2
    ns = int(round(soft / pitch))
    nc = int(round(cortical / pitch))
    if inner - 2 * nc < 1 or inner - ns - 2 * nc < 1:
        raise ValueError("phantom too small for its cortical shell and soft-tissue layer")
    core_shape = (inner - 2 * nc, inner - 2 * nc, inner - ns - 2 * nc)
    f = gaussian_random_field(core_shape, pitch, corr_len, seed)
    core = np.where(f < np.quantile(f, porosity), 1, 0).astype(np.uint8)
    lat = tiles * core_shape[0] + 2 * nc
    vol = np.empty((lat + 2, lat + 2, n), np.uint8)
    interior = vol[1:-1, 1:-1, 1:-1]
    interior[:] = 2
    interior[:, :, :ns] = 3
    interior[nc:lat - nc, nc:lat - nc, ns + nc:inner - nc] = np.tile(core, (tiles, tiles, 1))
    shell(vol, 4)
    return vol


def bone_optics():
    """Label table of the bone phantom: bone, marrow, cortical, soft tissue; ambient 1.0."""
    return optics_table(ma=[0.02, 0.1, 0.01, 0.03], ms=[25.0, 10.0, 30.0, 12.0],
                        g=[0.92, 0.9, 0.92, 0.9], n=[1.55, 1.37, 1.55, 1.40], n_ambient=1.0)
