"""PlateExModel geometry, the first_layer_clear launch weight and the bookkeeping of
packets reflected at entry (initial_refrect_by_angle), on the host."""
import numpy as np
import pytest

from pymonteopt_amd.voxel import PlateExModel, VoxelPlateExModel


def test_plate_ex_geometry():
    m = PlateExModel()
    m.set_params(x_size=2.0, y_size=1.5, voxel_space=0.1, thickness=[0.3, 1.0])
    m.build()
    vm = m.voxel_model
    # intralipid int(2/0.1) x int(1.5/0.1) x int(1/0.1) = 20 x 15 x 10; the glass is
    # int(0.3/0.1) + 1 = 3 voxels around (0.3/0.1 = 2.9999999999999996 truncates to 2)
    assert vm.shape == (26, 21, 16)
    assert np.all(vm[0] == 2) and np.all(vm[:, -1] == 2) and np.all(vm[:, :, 0] == 2)
    assert np.all(vm[3:23, 3:18, 3:13] == 1)
    core = np.zeros_like(vm, bool)
    core[3:23, 3:18, 3:13] = True
    shell = np.zeros_like(vm, bool)
    shell[1:-1, 1:-1, 1:-1] = True
    assert np.all(vm[shell & ~core] == 0)
    assert m.get_second_layer_addz() == 3


def test_first_layer_clear_launch():
    v = VoxelPlateExModel(nPh=10, device="cpu")
    v.build(x_size=2.0, y_size=2.0, voxel_space=0.1, thickness=[0.3, 1.0], n=[1.5, 1.4], n_air=1.0)
    s = v.source()
    assert s.voxel == (12, 12, 3) and v.model.voxel_model[s.voxel] == 1 and v.model.voxel_model[12, 12, 2] == 0
    # air|glass R1 = (0.5/2.5)^2 = 0.04, glass|intralipid r2 = (0.1/2.9)^2 = 0.00118906;
    # Rsp = R1 + r2 (1 - R1)^2 / (1 - R1 r2) = 0.0410959
    assert s.w0 == pytest.approx(1 - 0.0410959, abs=2e-7)


def test_entry_reflections_join_the_results():
    v = VoxelPlateExModel(nPh=10, beam_angle=0.5, initial_refrect_by_angle=True, first_layer_clear=False,
                          na=9, exit_state=True, device="cpu")
    raw = {"scalars": {"Rd": 1.0, "face_z-": 0.5, "w0": 7.0}, "events": {"packets": 7, "escaped": 6},
           "rd_a": np.zeros(9), "exit_state": np.ones((7, 7), np.float32)}
    v._add_entry_reflections(raw, 3)
    assert raw["scalars"] == {"Rd": 4.0, "face_z-": 3.5, "w0": 10.0}
    assert raw["events"] == {"packets": 10, "escaped": 9}
    assert raw["rd_a"][int(np.ceil(0.5 / (np.pi / 18))) - 1] == 3.0 and raw["rd_a"].sum() == 3.0
    es = raw["exit_state"]
    assert es.shape == (7, 10) and np.allclose(es[:, 0], [0, 0, 0, np.sin(0.5), 0, -np.cos(0.5), 1])
