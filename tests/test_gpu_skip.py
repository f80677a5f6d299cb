"""Face skipping in the walk: the device uniform-cube radius equals its host restatement bit
for bit, and a walk that crosses same-label cubes in one step makes the same draws and
events as the voxel-by-voxel walk, so its tallies differ by rounding only."""
import ctypes as C

import numpy as np
import pytest

from _parity import gpu_run, ledger_error, oracle_run, zscore
from _skip import blob_volume, skip_radius

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def device_radius(torch, labels, cap):
    from pymonteopt_amd import _abi
    lab = torch.from_numpy(np.ascontiguousarray(labels).reshape(-1)).cuda()
    out = torch.empty_like(lab)
    scratch = torch.empty_like(lab)
    L = _abi.lib()
    _abi.check(L.pmo_skip_radius(C.c_void_p(lab.data_ptr()), *labels.shape, cap, C.c_void_p(out.data_ptr()),
                                 C.c_void_p(scratch.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(labels.shape)


@pytest.mark.parametrize("shape,n_labels,cap,seed", [
    ((40, 37, 33), 3, 15, 0), ((40, 37, 33), 3, 3, 1), ((5, 64, 9), 2, 127, 2), ((3, 3, 3), 1, 1, 3)])
def test_device_radius_equals_host(torch_cuda, shape, n_labels, cap, seed):
    lab = blob_volume(shape, n_labels, seed)
    print(f"skip radius {shape} labels {n_labels} cap {cap}", flush=True)
    assert np.array_equal(device_radius(torch_cuda, lab, cap), skip_radius(lab, cap))


def test_device_radius_with_label_255_and_long_axes(torch_cuda):
    lab = blob_volume((7, 9, 300), 2, 4)
    lab[:, 5:, 100:180] = 255
    for cap in (15, 127):
        assert np.array_equal(device_radius(torch_cuda, lab, cap), skip_radius(lab, cap)), cap
    lab = blob_volume((300, 6, 5), 3, 6)
    assert np.array_equal(device_radius(torch_cuda, lab, 40), skip_radius(lab, 40))


def test_device_radius_on_model_volumes(torch_cuda):
    from pymonteopt_amd import presets
    for wl in (presets.skin_plate(), presets.bone_workload(seed=1, n=128)):
        print(f"skip radius {wl.name}", flush=True)
        assert np.array_equal(device_radius(torch_cuda, wl.labels, 15), skip_radius(wl.labels, 15))


def test_radius_argument_checks(torch_cuda):
    lab = blob_volume((8, 8, 8), 2, 0)
    for cap in (0, 128):
        with pytest.raises(ValueError):
            device_radius(torch_cuda, lab, cap)


@pytest.fixture(scope="module")
def slab_pair(torch_cuda):
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import Walker
    wl = presets.slab_plate()
    return wl, Walker(wl.labels, wl.pitch, wl.optics, skip=True), Walker(wl.labels, wl.pitch, wl.optics, skip=False)


def test_skipping_changes_rounding_only(slab_pair):
    """Same seeds with and without skipping: the events agree (a rare packet may part at a
    rounding tie), and each seed's Rd, Tt and A differ far less than the seed-to-seed
    spread; the FP64 voxel-by-voxel walk is no closer to the voxel-by-voxel FP32 walk than
    to the skipping one."""
    wl, ws, wv = slab_pair
    N = 100000
    S, V, D = [], [], []
    for s in range(6):
        S.append(gpu_run(ws, wl, N, s))
        V.append(gpu_run(wv, wl, N, s))
        D.append(gpu_run(wv, wl, N, s, fp64=True))
    for a, b in zip(S, V):
        assert ledger_error(a["scalars"]) < 1e-6
        for k in ("packets", "escaped", "killed"):
            assert a["events"][k] == b["events"][k], k
        for k in ("interactions", "crossings", "fresnel"):
            print(k, a["events"][k], b["events"][k], flush=True)
            assert abs(a["events"][k] - b["events"][k]) <= 1e-3 * b["events"][k], (k, a["events"][k], b["events"][k])
    for k in ("Rd", "Tt", "A"):
        x = np.array([r["scalars"][k] / N for r in S])
        y = np.array([r["scalars"][k] / N for r in V])
        z = np.array([r["scalars"][k] / N for r in D])
        spread = np.std(y, ddof=1)
        print(k, "skip-voxel", np.max(np.abs(x - y)) / spread, "skip-f64", np.max(np.abs(x - z)) / spread,
              "voxel-f64", np.max(np.abs(y - z)) / spread, flush=True)
        assert np.max(np.abs(x - y)) < 0.5 * spread, (k, x - y, spread)
        assert np.max(np.abs(x - z)) < 0.5 * spread, (k, x - z, spread)
    ga = np.sum([r["absorbed"] for r in S], axis=0)
    va = np.sum([r["absorbed"] for r in V], axis=0)
    print("absorbed L1", np.abs(ga - va).sum() / va.sum(), flush=True)
    assert np.abs(ga - va).sum() / va.sum() < 1e-2


def test_voxel_stepping_matches_oracle(slab_pair, oracle):
    """The voxel-by-voxel walk (skip off) against the oracle on the slab, as the default
    (skipping) walk is in the other parity tests."""
    wl, _, wv = slab_pair
    N = 100000
    G = [gpu_run(wv, wl, N, s) for s in range(10)]
    O = [oracle_run(oracle, wl, N, s) for s in range(10)]
    for k in ("Rd", "Tt", "A"):
        assert zscore([g["scalars"][k] for g in G], [o["scalars"][k] for o in O]) < 3, k


def test_default_skips_where_it_pays(torch_cuda):
    """Walker(skip=None) skips on the plates and the bone phantom, not on the smallest volume
    (12^3) nor on the fragmented 64^3 four-label grid (mean radius 0.24)."""
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import Walker
    cases = ((presets.slab_plate(), True), (presets.bone_workload(seed=0, n=128), True),
             (presets.grf_grid(seed=0), False),
             (presets.plate("slab 12^3", [5], [0.01], [10], [0.9], [1.5], 1.0, 0.5, 5, 5), False))
    for wl, want in cases:
        wk = Walker(wl.labels, wl.pitch, wl.optics)
        assert wk.skip is want, wl.name
        assert (wk._skip_radius is not None) is want, wl.name
        assert Walker(wl.labels, wl.pitch, wl.optics, skip=True).skip


def test_skipping_on_the_bone_phantom(torch_cuda):
    """On the 128^3 bone phantom (bone/marrow index steps at every trabecula) the skipping
    and the voxel-stepping walks, same seeds: the same packet counts, Rd, Tt and A within a
    fraction of the seed-to-seed spread, and the energy ledger."""
    from pymonteopt_amd import presets
    from pymonteopt_amd.runtime import Walker
    wl = presets.bone_workload(seed=2, n=128)
    ws = Walker(wl.labels, wl.pitch, wl.optics, skip=True)
    wv = Walker(wl.labels, wl.pitch, wl.optics, skip=False)
    N = 100000
    S = [gpu_run(ws, wl, N, s) for s in range(6)]
    V = [gpu_run(wv, wl, N, s) for s in range(6)]
    for a, b in zip(S, V):
        assert ledger_error(a["scalars"]) < 1e-6
        assert a["events"]["packets"] == b["events"]["packets"] == N
        for k in ("interactions", "crossings", "fresnel"):
            assert abs(a["events"][k] - b["events"][k]) <= 2e-3 * b["events"][k], (k, a["events"][k], b["events"][k])
    for k in ("Rd", "Tt", "A"):
        x = np.array([r["scalars"][k] / N for r in S])
        y = np.array([r["scalars"][k] / N for r in V])
        spread = np.std(y, ddof=1)
        print(k, "skip-voxel", np.max(np.abs(x - y)) / spread, flush=True)
        assert np.max(np.abs(x - y)) < 0.5 * spread, (k, x - y, spread)


def test_packed_label_radius(torch_cuda):
    """pmo_pack_label_radius writes labels | radius << 8 for every voxel, on a volume larger
    than one pass of its grid (2e7 voxels), and the Walker's packed volume is that of its
    labels and radius; pmo_walk reads it (skip alone is refused on the host)."""
    from pymonteopt_amd import _abi, presets
    from pymonteopt_amd.runtime import Walker
    rng = np.random.default_rng(4)
    n = 20_000_003
    lab = torch_cuda.from_numpy(rng.integers(0, 256, n, dtype=np.uint8)).cuda()
    rad = torch_cuda.from_numpy(rng.integers(0, 16, n, dtype=np.uint8)).cuda()
    out = torch_cuda.full((n,), -1, dtype=torch_cuda.int16, device="cuda")
    print("pmo_pack_label_radius", n, flush=True)
    _abi.check(_abi.lib().pmo_pack_label_radius(C.c_void_p(lab.data_ptr()), C.c_void_p(rad.data_ptr()),
                                                C.c_void_p(out.data_ptr()), n,
                                                C.c_void_p(torch_cuda.cuda.current_stream().cuda_stream)))
    torch_cuda.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, lab.cpu().numpy().astype(np.uint16) | (rad.cpu().numpy().astype(np.uint16) << 8))
    wl = presets.bone_workload(seed=2, n=128)
    wk = Walker(wl.labels, wl.pitch, wl.optics, skip=True)
    lr = wk.label_radius().cpu().numpy().view(np.uint16)
    assert np.array_equal(lr, wk.labels.cpu().numpy().astype(np.uint16)
                          | (wk.skip_radius().cpu().numpy().astype(np.uint16) << 8))
    vol = wk.volume_struct(skip=True)
    assert bool(vol.label_radius) and bool(vol.skip)
