"""Named workloads: the slab, skin and bone geometries the walk is measured and checked on.

Each builder returns a `Workload` (label volume, pitch, optics table, launch source and
the exit-tally layout) so the same case can be given to the HIP walk and to a checker.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .geometry import (Source, bone_optics, bone_phantom, bone_phantom_wide, center_add, optics_table,
                       plate_volume, specular)


@dataclass
class Workload:
    name: str
    labels: np.ndarray
    pitch: float
    optics: np.ndarray
    source: Source
    tallies: dict = field(default_factory=dict)   # TallySpec keyword arguments

    @property
    def n_labels(self):
        return self.optics.shape[0] - 1

    def optics_columns(self):
        o = self.optics.astype(np.float64)
        L = self.n_labels
        return o[:L, 0], o[:L, 1], o[:L, 2], o[:, 3]


def plate(name, thickness, ma, ms, g, n, n_air, voxel_space, x_size, y_size, beam="pencil", radius=0.0,
          tallies=None) -> Workload:
    vol = plate_volume(thickness, voxel_space, x_size, y_size)
    opt = optics_table(ma, ms, g, n, n_air)
    w0 = float(np.float32(1.0) - np.float32(specular(n_air, n[0])))
    kind = {"pencil": _abi.PMO_BEAM_PENCIL, "flat": _abi.PMO_BEAM_FLAT, "gauss": _abi.PMO_BEAM_GAUSS}[beam]
    src = Source(voxel=(center_add(vol.shape[0]), center_add(vol.shape[1]), 1), w0=w0, kind=kind,
                 radius=float(radius))
    return Workload(name, vol, float(voxel_space), opt, src, dict(tallies or {}))


def slab_plate() -> Workload:
    """Homogeneous slab: 101 x 101 x 50 voxels at 0.1 mm (5 mm thick), mu_a 0.1, mu_s 10,
    g 0.9, n 1.37 in air; pencil beam; Rd(r) in 50 bins of 0.1 mm."""
    return plate("slab 101x101x50 @0.1mm", [5.0], [0.1], [10.0], [0.9], [1.37], 1.0, 0.1, 10.1, 10.1,
                 tallies=dict(nr=50, dr=0.1))


def skin_plate(beam="pencil") -> Workload:
    """3-layer skin: 201 x 201 x 140 voxels at 0.05 mm; epidermis 0.1 mm (0.5, 30, 0.8, 1.44),
    dermis 2 mm (0.05, 20, 0.9, 1.40), subcutis 4.9 mm (0.02, 12, 0.9, 1.44), air above;
    pencil beam or a 1 mm-radius uniform flat beam; Rd(r), Tt(r) in 40 bins of 0.1 mm."""
    return plate(f"3-layer skin 201x201x140 @0.05mm {beam}", [0.1, 2.0, 4.9], [0.5, 0.05, 0.02],
                 [30.0, 20.0, 12.0], [0.8, 0.9, 0.9], [1.44, 1.40, 1.44], 1.0, 0.05, 10.05, 10.05,
                 beam=beam, radius=1.0 if beam == "flat" else 0.0, tallies=dict(nr=40, dr=0.1))


def bone_workload(seed=0, n=256) -> Workload:
    """Cancellous-bone phantom on n^3 voxels at 0.02 mm (GRF corr. length 0.15 mm, porosity
    0.7, 0.5 mm cortical shell, 1 mm soft-tissue top); pencil beam at the top-face centre;
    ten 0.25 mm-wide annular top-face detectors at 0.25..2.5 mm source-detector separation
    (the phantom is 5.12 mm wide, so separations are limited to 2.5 mm)."""
    vol = bone_phantom(seed=seed, n=n)
    opt = bone_optics()
    top = int(vol[center_add(n), center_add(n), 1])
    w0 = float(np.float32(1.0) - np.float32(specular(1.0, float(opt[top, 3]))))
    src = Source(voxel=(center_add(n), center_add(n), 1), w0=w0)
    rings = tuple((0.25 * i - 0.125, 0.25 * i + 0.125) for i in range(1, 11))
    return Workload(f"bone phantom {n}^3 @0.02mm", vol, 0.02, opt, src,
                    dict(nr=50, dr=0.05, rings=rings))


def bone_wide_workload(seed=0, n=256, tiles=5) -> Workload:
    """The cancellous-bone phantom of bone_workload with annular detectors at 1-10 mm
    source-detector separation: the n^3 phantom's trabecular core (0.02 mm voxels, GRF
    correlation length 0.15 mm, porosity 0.7) tiled 5 x 5 across the top face (21.4 mm
    wide for n = 256), under the same 1 mm soft-tissue layer, in the same 0.5 mm cortical
    shell; pencil beam at the top-face centre; ten 0.5 mm-wide top-face rings centred at
    1, 2, ..., 10 mm; Rd(r) in 110 bins of 0.1 mm."""
    vol = bone_phantom_wide(seed=seed, n=n, tiles=tiles)
    opt = bone_optics()
    c = center_add(vol.shape[0])
    top = int(vol[c, c, 1])
    w0 = float(np.float32(1.0) - np.float32(specular(1.0, float(opt[top, 3]))))
    src = Source(voxel=(c, center_add(vol.shape[1]), 1), w0=w0)
    rings = tuple((i - 0.25, i + 0.25) for i in range(1, 11))
    return Workload(f"bone phantom {vol.shape[0]}x{vol.shape[1]}x{vol.shape[2]} @0.02mm (core of {n}^3 tiled "
                    f"{tiles}x{tiles})", vol, 0.02, opt, src, dict(nr=110, dr=0.1, rings=rings))


def grf_grid(seed=0, n=64, pitch=0.1, corr_len=0.3) -> Workload:
    """Heterogeneous 4-label grid: n^3 voxels at `pitch` whose interior labels are the
    quartiles of a Gaussian random field (correlation length `corr_len` mm); pencil beam
    at the top-face centre; Rd(r) in 40 bins of 0.1 mm."""
    from .geometry import gaussian_random_field, shell
    f = gaussian_random_field((n - 2, n - 2, n - 2), pitch, corr_len, seed)
    q = np.quantile(f, [0.25, 0.5, 0.75])
    vol = np.empty((n, n, n), np.uint8)
    vol[1:-1, 1:-1, 1:-1] = np.digitize(f, q).astype(np.uint8)
    shell(vol, 4)
    opt = optics_table(ma=[0.02, 0.05, 0.1, 0.2], ms=[10.0, 15.0, 20.0, 25.0], g=[0.9, 0.9, 0.85, 0.8],
                       n=[1.37, 1.40, 1.44, 1.40], n_ambient=1.0)
    top = int(vol[center_add(n), center_add(n), 1])
    w0 = float(np.float32(1.0) - np.float32(specular(1.0, float(opt[top, 3]))))
    src = Source(voxel=(center_add(n), center_add(n), 1), w0=w0)
    return Workload(f"GRF 4-label grid {n}^3 @{pitch}mm", vol, pitch, opt, src, dict(nr=40, dr=0.1))
