"""Whole-walk known answers from the reference's own notebooks and the literature they
cite, run through the oracle (10 seed batches, the notebooks' protocol)."""
import json
import math
import os

import numpy as np

from pymonteopt_amd.geometry import optics_table, plate_volume, specular

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def batches(oracle, params, n, seeds=range(10), voxel_space=None, precision="f64"):
    p = dict(params)
    l = voxel_space if voxel_space is not None else p["voxel_space"]
    vol = plate_volume(p["thickness"], l, p["x_size"], p["y_size"])
    opt = optics_table(p["ma"], p["ms"], p["g"], p["n"], p["n_air"])
    L = opt.shape[0] - 1
    src = dict(kind=0, voxel=((vol.shape[0] - 1) // 2, (vol.shape[1] - 1) // 2, 1),
               w0=float(np.float32(1) - np.float32(specular(p["n_air"], p["n"][0]))))
    out = {"Rd": [], "Tt": []}
    for s in seeds:
        r = oracle.walk(vol, l, opt[:L, 0], opt[:L, 1], opt[:L, 2], opt[:, 3], src, n, s, absorbed=False,
                        precision=precision)
        out["Rd"].append(r["scalars"]["Rd"] / n)
        out["Tt"].append(r["scalars"]["Tt"] / n)
    return {k: (float(np.mean(v)), float(np.std(v, ddof=1) / math.sqrt(len(v)))) for k, v in out.items()}


def z(a, b):
    return abs(a[0] - b[0]) / math.sqrt(a[1] ** 2 + b[1] ** 2)


def test_thin_slab_van_de_hulst(oracle):
    c = KA["thin_slab"]
    r = batches(oracle, c["params"], 20000)
    for k in ("Rd", "Tt"):
        assert z(r[k], (c["literature"][k], 0.0)) < 4
        assert z(r[k], tuple(c["reference_cpu"][k])) < 3
        assert z(r[k], tuple(c["reference_gpu"][k])) < 3


def test_semi_infinite_mismatched(oracle):
    c = KA["semi_infinite"]
    r = batches(oracle, c["params"], 20000)
    assert z(r["Rd"], tuple(c["reference_cpu"]["Rd"])) < 3
    assert z(r["Rd"], tuple(c["reference_gpu"]["Rd"])) < 3
    assert r["Tt"][0] == 0.0


def test_three_layer(oracle):
    c = KA["three_layer"]
    nb = 20000
    r = batches(oracle, c["params"], nb)
    for k in ("Rd", "Tt"):
        # the reference ran one 1e6-packet batch: its SE is ours scaled to 1e6 packets
        se_ref = r[k][1] * math.sqrt(10 * nb / c["n_packets"])
        assert z(r[k], (c["reference_gpu"][k], se_ref)) < 3


def test_voxel_size_invariance(oracle):
    c = KA["voxel_size_invariance"]
    nb = 5000
    for pitch, key in ((5.0, "5mm"), (0.05, "0.005mm")):
        r = batches(oracle, c["params"], nb, voxel_space=pitch)
        ref = c["reference_gpu"][key]
        for k in ("Rd", "Tt"):
            se_ref = r[k][1] * math.sqrt(10 * nb / c["n_packets"])
            assert z(r[k], (ref[k], se_ref)) < 3, (pitch, k, r[k], ref[k])
