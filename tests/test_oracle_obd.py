"""Pins the OBD-receiver oracle (oracle/obd.py) against what optics fixes independently:
published glass indices, Snell's law, the walk's pinned Fresnel, paraxial focusing of both
lenses, the reference's own closed-form marginal ray (Grass.get_real_bfl /
get_inital_angle, pymopt/optics/_classes.py:229-258) traced back through lens 1, hand-computed
LD-mount cut points, and its Philox against the C oracle's."""
import math

import numpy as np
import pytest

from oracle import obd

# Schott catalogue Sellmeier coefficients (B1, B2, B3), (C1, C2, C3) [um^2]
SELLMEIER = {"N-BK7": ((1.03961212, 0.231792344, 1.01046945), (0.00600069867, 0.0200179144, 103.560653)),
             "N-SF11": ((1.73759695, 0.313747346, 1.89878101), (0.013188707, 0.0623068142, 155.23629))}


def sellmeier(wl, glass):
    B, Cc = SELLMEIER[glass]
    l2 = (wl * 1e-3) ** 2
    return math.sqrt(1 + sum(b * l2 / (l2 - c) for b, c in zip(B, Cc)))


@pytest.mark.parametrize("wl", [450, 500, 600, 700, 800, 850, 900, 1000, 1300, 1550, 2000])
def test_glass_fits_follow_sellmeier(wl):
    assert obd.n_bk7(wl) == pytest.approx(sellmeier(wl, "N-BK7"), abs=3e-4)
    assert obd.n_sf11(wl) == pytest.approx(sellmeier(wl, "N-SF11"), abs=3e-4)
    with pytest.raises(ValueError):
        obd.n_bk7(3000)


def test_numpy_philox_matches_the_c_oracle(oracle):
    rng = np.random.default_rng(4)
    ctr = rng.integers(0, 2 ** 32, size=(16, 4), dtype=np.uint64)
    key = rng.integers(0, 2 ** 32, size=(16, 2), dtype=np.uint64)
    for c, k in zip(ctr, key):
        got = [int(x[0]) for x in obd.philox(*[np.array([x]) for x in c], int(k[0]), int(k[1]))]
        assert got == list(oracle.philox([int(x) for x in c], [int(x) for x in k]))


def test_fresnel_matches_the_reference_form_and_the_walk_oracle(oracle):
    for ni, nt in ((1.0, 1.5098), (1.7619, 1.0), (1.0, 1.7619), (1.5098, 1.0)):
        for ai in (0.05, 0.3, 0.6, 0.9):
            st = ni / nt * math.sin(ai)
            if st >= 1:
                continue
            at = math.asin(st)
            ref = 0.5 * ((math.sin(ai - at) / math.sin(ai + at)) ** 2 + (math.tan(ai - at) / math.tan(ai + at)) ** 2)
            got = obd.fresnel(math.cos(ai), math.cos(at), ni, nt)
            assert got == pytest.approx(ref, rel=1e-10)
            r_walk, _ = oracle.fresnel(ni, nt, math.cos(ai))
            assert got == pytest.approx(r_walk, rel=1e-9)


def test_refraction_obeys_snell():
    rng = np.random.default_rng(1)
    for ni, nt in ((1.0, 1.5098), (1.7619, 1.0)):
        nn = obd._norm(rng.normal(size=(3, 200)) * np.array([[0.2], [0.2], [1.0]]) + np.array([[0], [0], [3.0]]))
        v = obd._norm(nn + 0.3 * rng.normal(size=(3, 200)))
        vt, R, tir = obd.refract(v, nn, ni, nt)
        vt = obd._norm(vt)
        ok = ~tir
        sin_i = np.linalg.norm(np.cross(v.T, nn.T), axis=1)
        sin_t = np.linalg.norm(np.cross(vt.T, nn.T), axis=1)
        assert np.allclose(ni * sin_i[ok], nt * sin_t[ok], rtol=1e-9, atol=1e-12)
        # v, nn and the refracted ray are coplanar; the ray goes on across the surface
        assert np.allclose(np.einsum("ij,ij->j", np.cross(v.T, nn.T).T, vt)[ok], 0, atol=1e-12)
        assert np.all(np.einsum("ij,ij->j", vt, nn)[ok] > 0)
        if ni > nt:
            assert tir.any() and np.all(sin_i[tir] * ni / nt >= 1 - 1e-12)


def _params(**kw):
    P = dict(obd.DEFAULT_PARAMS)
    P.update(kw)
    return P


def _through_lens1(P, Z, p, v):
    G = obd.geometry(P, Z)
    nz = np.array([0.0, 0.0, -1.0])[:, None] * np.ones(p.shape[1])[None, :]
    p = obd.plane_hit(p, v, G["z_lens1"])
    v, _, _ = obd.refract(v, nz, 1.0, P["n_1"])
    v = obd._norm(v)
    p, ok = obd.sphere_hit(p, v, G["c_lens1"], P["r_1"], outward=True)
    v, _, tir = obd.refract(v, obd.sphere_normal(p, P["r_1"], outward=True), P["n_1"], 1.0)
    assert ok.all() and not tir.any()
    return p, obd._norm(v)


def test_lens1_collimates_a_source_at_its_front_focus():
    # design values: n = 1.517 gives f = r/(n-1) = 99.96 and bfl = f - ct/n = 93.37 ~ bfl_1
    P = _params(n_1=1.517)
    th = np.array([0.005, 0.01, 0.02, -0.015])
    v = np.array([np.sin(th), np.zeros_like(th), -np.cos(th)])
    p = np.zeros((3, th.size))
    _, vo = _through_lens1(P, 0.0, p, v)
    out_angle = np.arctan2(vo[0], -vo[2])
    assert np.all(np.abs(out_angle) < 0.02 * np.abs(th))


def test_lens2_focuses_collimated_light_on_the_photodiode_plane():
    # design values: n = 1.785 gives f = 49.99 and bfl = f - ct/n = 43.27 ~ bfl_2
    P = _params(n_2=1.785)
    G = obd.geometry(P, 0.0)
    h = np.array([0.5, 1.0, 2.0, -1.5])
    p = np.array([h, np.zeros_like(h), np.full_like(h, G["z_slit2"])])
    v = np.array([np.zeros_like(h), np.zeros_like(h), -np.ones_like(h)])
    p, ok = obd.sphere_hit(p, v, G["c_lens2"], P["r_2"], outward=False)
    v, _, _ = obd.refract(v, obd.sphere_normal(p, P["r_2"], outward=False), 1.0, P["n_2"])
    v = obd._norm(v)
    p = obd.plane_hit(p, v, G["z_lens2"])
    nz = np.array([0.0, 0.0, -1.0])[:, None] * np.ones(h.size)[None, :]
    v, _, _ = obd.refract(v, nz, P["n_2"], 1.0)
    p = obd.plane_hit(p, obd._norm(v), G["z_pd"])
    assert ok.all()
    assert np.all(np.abs(p[0]) < 0.02 * np.abs(h))


def test_marginal_ray_of_get_real_bfl_traces_back_to_the_slit_radius():
    # Grass.get_real_bfl gives where a ray parallel to the axis at the slit radius crosses
    # the axis after lens 1 (reversed), and get_inital_angle its angle there; a ray from
    # that axis point at that angle must leave lens 1 parallel to the axis at that radius
    wl, h = 850, 10.0
    P = obd.set_indices(_params(wavelength=wl))
    z_rel = obd.real_bfl(h, P["r_1"], P["ct_1"], wl, "N-BK7")
    a4 = obd.initial_angle(h, P["r_1"], wl, "N-BK7")
    Z = P["bfl_1"] - z_rel                  # puts the plano face z_rel below the source
    p = np.zeros((3, 1))
    v = np.array([[math.sin(a4)], [0.0], [-math.cos(a4)]])
    po, vo = _through_lens1(P, Z, p, v)
    assert abs(vo[0, 0]) < 1e-9 and vo[2, 0] == pytest.approx(-1.0, abs=1e-12)
    assert po[0, 0] == pytest.approx(h, abs=1e-9)


def test_ld_mount_cut_at_hand_computed_points():
    # angle 68/180 rad: a1 = -0.39659, b1 = 6.45522, a2 = 5.23064; keep where
    # y > a1 x + b1, or y < a2 x - 6, or y < -6
    # at x = 1 the upper edge is y = 6.0586 and the lower one y = -0.7694
    pts = np.array([[0.0, 0.0, 2.0, -2.0, 0.0, 1.0, 1.0, 1.0, 1.0],
                    [0.0, 7.0, 0.0, -3.0, -6.5, 6.0, 6.1, -0.7, -0.8]])
    want = np.array([False, True, True, False, True, False, True, False, True])
    assert np.array_equal(obd.lfp_keep(pts), want)


def test_stochastic_cut_averages_to_the_transmission():
    rng = np.random.default_rng(7)
    n = 40000
    P = obd.set_indices(_params())
    # rays from a 1 mm spot toward the slit annulus of lens 1 at Z = 0
    r0 = rng.uniform(0, 0.5, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    p = np.array([r0 * np.cos(ph), r0 * np.sin(ph), np.zeros(n)])
    ang = rng.uniform(0.09, 0.13, n)
    ph2 = rng.uniform(0, 2 * np.pi, n)
    v = np.array([np.sin(ang) * np.cos(ph2), np.sin(ang) * np.sin(ph2), -np.cos(ang)])
    w = np.ones(n)
    idx = np.arange(n)
    P["ld_fix_part"] = False
    s = obd.trace_step(P, 0.0, p, v, w, u=obd.step_uniforms(0, idx, 5, 0), mode="stochastic")
    e = obd.trace_step(P, 0.0, p, v, w, mode="expect")
    hit = e > 0
    assert hit.sum() > 1000
    assert np.all(s[~hit] == 0)
    # each detected ray passes with probability e (its product of 1 - R)
    mean, var = e.sum(), (e * (1 - e)).sum()
    assert abs(s.sum() - mean) < 4 * math.sqrt(var)
    assert np.all((e[hit] > 0.75) & (e[hit] < 1))
