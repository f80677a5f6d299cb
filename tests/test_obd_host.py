"""Host side of the OBD receiver: glass indices and marginal ray, the per-step geometry
table the kernel reads, the parameter contract and the device-call validation (no GPU)."""
import ctypes as ct
import math

import numpy as np
import pytest

from oracle import obd
from pymonteopt_amd import _abi
from pymonteopt_amd.optics import OBD, Grass, SideOBD


def test_grass_matches_the_oracle():
    g = Grass()
    for wl in (500, 700, 850, 900, 1300):
        assert g.get_refractive_index(wl, "N-BK7") == pytest.approx(obd.n_bk7(wl), rel=1e-14)
        assert g.get_refractive_index(wl, "N-SF11") == pytest.approx(obd.n_sf11(wl), rel=1e-14)
        assert g.get_inital_angle(10, 51.68, wl) == pytest.approx(obd.initial_angle(10, 51.68, wl), rel=1e-12)
        assert g.get_real_bfl(10, 51.68, 10.0, wl) == pytest.approx(obd.real_bfl(10, 51.68, 10.0, wl), rel=1e-12)
    assert g.get_refractive_index(850, "fused silica") == 1.517
    with pytest.raises(ValueError):
        g.get_refractive_index(200, "N-BK7")


def test_params_contract_and_indices():
    o = OBD.__new__(OBD)
    OBD.__init__(o, device="cpu")
    assert o.params["n_1"] == pytest.approx(obd.n_bk7(850)) and o.params["n_2"] == pytest.approx(obd.n_sf11(850))
    o.set_params(wavelength=800, n_1=1.6)          # indices follow the substrates, as in the reference
    assert o.params["n_1"] == pytest.approx(obd.n_bk7(800))
    with pytest.raises(KeyError):
        o.set_params(focal_length=3)
    o.set_monte_data(nPh=123)
    assert o.nPh == 123
    with pytest.raises(KeyError):
        o.set_monte_data(photons=1)
    # a model's whole result dict: the exit states and nPh are taken, the tallies ignored
    o.set_monte_data({"p": 1, "v": 2, "w": 3, "nPh": 7, "Rd": 0.4, "Rd_r": [0.1]})
    assert (o.data["p"], o.data["v"], o.data["w"], o.nPh) == (1, 2, 3, 7)


def test_step_table_matches_the_oracle_geometry():
    o = OBD(device="cpu")
    o.set_params(start=-2.0, end=3.0, split=0.5, distance_2slits=30, pd_poit_correction=0.7)
    Z, t = o.step_table()
    assert t.shape == (10, _abi.PMO_OBD_STEP_FIELDS) and t.dtype == np.float32
    keys = ["z_lens1", "c_lens1", "z_slit1", "z_lfp", "z_slit2", "z_lens2", "c_lens2", "z_pd"]
    for i, z in enumerate(Z):
        G = obd.geometry(o.params, z)
        assert np.allclose(t[i, :8], [G[k] for k in keys], rtol=1e-6)
        assert t[i, 8] == 0.0


def test_side_obd_shift_follows_the_marginal_ray():
    s = SideOBD(device="cpu")
    Z, t = s.step_table()
    P = s.params
    th0 = obd.initial_angle(P["slit_D"] / 2, P["r_1"], P["wavelength"])
    dz = obd.real_bfl(P["slit_D"] / 2, P["r_1"], P["ct_1"], P["wavelength"]) - P["bfl_1"]
    assert np.allclose(t[:, 8], -(Z + dz) * math.tan(th0), rtol=1e-6, atol=1e-6)
    assert Z.size == 260


def test_obd_trace_argument_errors():
    L = _abi.lib()
    c = OBD(device="cpu")._config(0, 0, False)
    buf = (ct.c_float * 16)()
    out = (ct.c_double * 4)()
    assert L.pmo_obd_trace(None, None, None, None, 0, 0, buf, 1, out, None) == _abi.PMO_ERR_INVALID
    assert L.pmo_obd_trace(ct.byref(c), buf, buf, buf, 5, 4, buf, 1, out, None) == _abi.PMO_ERR_INVALID
    assert L.pmo_obd_trace(ct.byref(c), buf, buf, buf, 1, 1, buf, 0, out, None) == _abi.PMO_ERR_INVALID
    assert L.pmo_obd_trace(ct.byref(c), buf, buf, buf, 1, 1, buf, 2000, out, None) == _abi.PMO_ERR_INVALID
    bad = OBD(device="cpu")._config(0, 0, False)
    bad.outer_r1 = 60.0
    assert L.pmo_obd_trace(ct.byref(bad), buf, buf, buf, 1, 1, buf, 1, out, None) == _abi.PMO_ERR_INVALID
    assert b"aperture" in L.pmo_last_error()
    # nothing to trace: returns before touching the device
    assert L.pmo_obd_trace(ct.byref(c), buf, buf, buf, 0, 0, buf, 1, out, None) == _abi.PMO_OK

