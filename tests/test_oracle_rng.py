"""Pins the oracle's Philox4x32-10 and uniform conversion (the draws every walk uses)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _vectors():
    return json.load(open(os.path.join(HERE, "golden", "philox_vectors.json")))


def test_philox_random123_known_answers(oracle):
    for v in _vectors()["random123_kat"]:
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"]


def test_philox_matches_rocrand(oracle):
    vec = _vectors()["vectors"]
    assert len(vec) >= 40
    for v in vec:
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"], v


def test_u01_range_and_exactness(oracle):
    assert oracle.u01(0) == 2.0 ** -24
    assert oracle.u01(0xFFFFFFFF) == 1.0
    assert oracle.u01(0xFF) == 2.0 ** -24          # the low 8 bits are dropped
    x = 0x12345678
    assert oracle.u01(x) == ((x >> 8) + 1) * 2.0 ** -24
    assert np.float32(oracle.u01(x)) == oracle.u01(x)   # exactly representable in float


def test_philox_uniformity(oracle):
    # 4096 blocks of the walk stream: uniforms pass a coarse chi-square test
    u = []
    for i in range(4096):
        u.extend(oracle.u01(x) for x in oracle.philox([i, 0, 7, 0], [3, 0]))
    h, _ = np.histogram(u, bins=32, range=(0, 1))
    e = len(u) / 32
    chi2 = ((h - e) ** 2 / e).sum()
    assert chi2 < 70          # 31 dof, p ~ 1e-4
