"""Host logic of the batched sweep: schedule construction and instance descriptors."""
import ctypes as C

import numpy as np

from pymonteopt_amd import _abi
from pymonteopt_amd.batched import Plate, slab_sweep


def schedule(budgets, batch):
    L = _abi.lib()
    b = (C.c_uint64 * len(budgets))(*budgets)
    n = L.pmo_plate_schedule(b, len(budgets), batch, None, 0)
    s = (C.c_uint32 * (4 * n))()
    assert L.pmo_plate_schedule(b, len(budgets), batch, s, n) == n
    return np.array(list(s), np.uint32).reshape(-1, 4)


def test_schedule_covers_every_packet_once():
    budgets = [10, 1000, 1, 513, 0, 256]
    s = schedule(budgets, 256)
    for i, N in enumerate(budgets):
        rows = s[s[:, 0] == i]
        ids = np.concatenate([np.arange(f, f + c) for _, f, c, _ in rows]) if len(rows) else np.array([])
        assert np.array_equal(np.sort(ids), np.arange(N))
        assert np.all(rows[:, 2] <= 256)
    # round-robin: the first entries visit every non-empty instance once
    assert s[:5, 0].tolist() == [0, 1, 2, 3, 5]


def test_schedule_errors():
    L = _abi.lib()
    b = (C.c_uint64 * 2)(5, 1 << 33)
    assert L.pmo_plate_schedule(b, 2, 64, None, 0) == _abi.PMO_ERR_UNSUPPORTED
    assert L.pmo_plate_schedule(b, 0, 64, None, 0) == _abi.PMO_ERR_INVALID
    assert L.pmo_plate_schedule(b, 1, 0, None, 0) == _abi.PMO_ERR_INVALID


def test_plate_column_matches_volume():
    p = Plate(thickness=[0.1, 0.2], ma=[0.1, 0.2], ms=[10, 20], g=[0.9, 0.8], n=[1.4, 1.3], voxel_space=0.05,
              x_size=1.0, y_size=1.0)
    v = p.volume()
    assert v.shape == p.shape()
    col = p.z_labels()
    assert np.array_equal(v[10, 10, :], col)
    assert np.all(v[0] == 2) and np.all(v[:, -1] == 2)


def test_slab_sweep_distribution():
    ps = slab_sweep(512, seed=0)
    th = np.array([p.thickness[0] for p in ps])
    N = np.array([p.n_packets for p in ps])
    g = np.array([p.g[0] for p in ps])
    msp = np.array([p.ms[0] * (1 - p.g[0]) for p in ps])
    assert th.min() >= 0.5 - 0.025 and th.max() <= 10.0 + 0.025
    assert N.min() >= 1e5 - 1 and N.max() <= 1e6
    assert np.all((g >= 0.7) & (g <= 0.95)) and np.all((msp >= 0.5 - 1e-9) & (msp <= 3 + 1e-9))
    assert 1.5e8 < N.sum() < 2.6e8
    assert np.allclose(np.round(th / 0.05), th / 0.05)


def test_instance_struct_layout():
    assert C.sizeof(_abi.PmoPlateInstance) == 96
    assert _abi.PmoPlateInstance.n_packets.offset == 72
    assert _abi.PmoPlateInstance.layers.offset == 80


def test_layer_table_matches_z_labels():
    from pymonteopt_amd.batched import Plate
    p = Plate([0.1, 0.4, 0.25], [0.5, 0.05, 0.1], [30.0, 20.0, 10.0], [0.8, 0.9, 0.9], [1.44, 1.40, 1.37],
              voxel_space=0.05, x_size=1.05, y_size=1.05)
    col, tab = p.z_labels(), p.layer_table()
    l = p.voxel_space
    assert tab.shape == (3, 4)
    assert tab[0, 0] == 0.0 and abs(tab[-1, 1] - l * (col.size - 2)) < 1e-6
    for j in range(tab.shape[0]):
        z0, z1, lab = int(round(tab[j, 0] / l)) + 1, int(round(tab[j, 1] / l)) + 1, int(tab[j, 2])
        assert np.all(col[z0:z1] == lab)
        if j + 1 < tab.shape[0]:
            assert tab[j, 1] == tab[j + 1, 0] and int(tab[j + 1, 2]) != lab
