"""TEST INFRASTRUCTURE ONLY -- numpy restatement of pyMonteOpt's OBD receiver.

The OBD (pymopt/optics/_classes.py:381-567) traces the packets a walk sends back out of
the top face (exit states p, v, w) through a receiver at focus offset Z. The receiver is:
* a plano-convex lens with its plano side toward the sample;
* two annular slits with the LD mount between them;
* a second plano-convex lens with its convex side toward the sample;
* a photodiode.

The receiver is scanned over Z, giving I(Z) = the weight reaching the photodiode. This is
the checker of libpmo's pmo_obd_trace. Only tests/ and bench.py's baseline leg use it.

The restatement follows the reference's operations in order, in float64:
* plane hits, `_classes.py:64-67,133-136`; sphere hits, `:138-150`;
* lens-aperture and slit cuts, `:69-77,127-131`;
* surface normals, `:194-210`; vector refraction, `:152-188`;
* the LD-mount cut, `:336-373`; the photodiode, `:316-334`;
* the set-up of `opticalUnit` and `opticalAnalysisMeth`, `:454-525`, and of SideOBD, `:527-567`.

Two choices differ from the reference, and both are stated where they apply:
* The reference draws its Fresnel decisions from np.random. Here, and in the kernel, the
  uniform of surface s at focus step i for packet j is Philox4x32-10 with counter
  (i, 3, j_lo, j_hi) and key (seed, stream), output word s. `mode="expect"` instead
  multiplies the weight by the transmission 1 - R, which is the mean of the stochastic
  cut.
* At exactly normal incidence the reference's sin/tan Fresnel is 0/0 = NaN and rejects
  the ray. Here the normal-incidence limit is used. The case is a single direction,
  measure zero in the reference's own float64 arithmetic.
"""
from __future__ import annotations

import math

import numpy as np

TAG_OBD = 3

# the reference's OBD.params defaults (pymopt/optics/_classes.py:384-397)
DEFAULT_PARAMS = {
    "start": -2.04, "end": 62.96, "split": 0.25, "wavelength": 850,
    "outerD_1": 50, "efl_1": 100, "bfl_1": 93.41, "ct_1": 10, "et_1": 3.553, "r_1": 51.68, "n_1": 1.517,
    "substrate_1": "N-BK7",
    "outerD_2": 50, "efl_2": 50, "bfl_2": 43.28, "ct_2": 12, "et_2": 3.01, "r_2": 39.24, "n_2": 1.758,
    "substrate_2": "N-SF11",
    "slit_outerD": 50, "slit_D": 20, "slit_width": 2, "slit_thickness": 5,
    "d_pd": 3, "ld_fix_part": True, "distance_2slits": 37, "pd_poit_correction": 0,
    "inversion": False, "side": False,
}


# ------------------------------------------------------------------ glass (_classes.py:223-293)

def n_bk7(wl):
    x = wl * 1e-3
    if 312.6 <= wl <= 486.1:
        return -2.4880 * x ** 3 + 3.5910 * x ** 2 - 1.8102 * x + 1.8395
    if 486.1 < wl <= 852.1:
        return -0.1605 * x ** 3 + 0.3866 * x ** 2 - 0.3307 * x + 1.6102
    if 852.1 < wl <= 2325:
        return -0.0029 * x ** 3 + 0.0125 * x ** 2 - 0.0304 * x + 1.5284
    raise ValueError("wavelength outside the N-BK7 fit")


def n_sf11(wl):
    x = wl * 1e-3
    if 404.7 <= wl <= 589.3:
        return -4.4506 * x ** 3 + 7.8207 * x ** 2 - 4.7512 * x + 2.7792
    if 589.3 < wl <= 852.1:
        return -0.4659 * x ** 3 + 1.1864 * x ** 2 - 1.0622 * x + 2.0937
    if 852.1 < wl <= 1060:
        return 0.0427 * x ** 2 - 0.1192 * x + 1.8324
    if 1060 < wl <= 2325.4:
        return -0.0046 * x ** 3 + 0.0257 * x ** 2 - 0.0648 * x + 1.7993
    raise ValueError("wavelength outside the N-SF11 fit")


def refractive_index(wl, glass):
    return {"N-BK7": n_bk7, "N-SF11": n_sf11}.get(glass, lambda _: 1.517)(wl)


def initial_angle(slit_radius, r, wl, glass="N-BK7"):
    n1 = refractive_index(wl, glass)
    a1 = math.asin(slit_radius / r)
    a2 = math.asin(math.sin(a1) / n1)
    return math.asin(n1 * math.sin(a1 - a2))


def real_bfl(slit_radius, r, ct, wl, glass="N-BK7"):
    n1 = refractive_index(wl, glass)
    a1 = math.asin(slit_radius / r)
    a2 = math.asin(math.sin(a1) / n1)
    a3 = a1 - a2
    a4 = math.asin(n1 * math.sin(a3))
    l = math.sqrt(r * r - slit_radius * slit_radius) - r + ct
    dx = l * math.tan(a3)
    return (slit_radius - dx) / math.tan(a4)


def set_indices(params):
    """OBD._set_refrective_index: n_1, n_2 follow the substrates at the wavelength."""
    params["n_1"] = refractive_index(params["wavelength"], params["substrate_1"])
    params["n_2"] = refractive_index(params["wavelength"], params["substrate_2"])
    return params


# ------------------------------------------------------------------ Philox draws

_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on uint32 arrays (Salmon et al., SC'11), vectorised."""
    c = [np.asarray(x, np.uint64) & _MASK for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0) & _MASK, np.uint64(k1) & _MASK
    for _ in range(10):
        p0 = np.uint64(_M0) * c[0]
        p1 = np.uint64(_M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & _MASK, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & _MASK]
        k0 = (k0 + np.uint64(_W0)) & _MASK
        k1 = (k1 + np.uint64(_W1)) & _MASK
    return [x.astype(np.uint32) for x in c]


def u01(x):
    return ((np.asarray(x, np.uint64) >> np.uint64(8)).astype(np.float64) + 1.0) * 2.0 ** -24


def step_uniforms(step, pid, seed, stream):
    """The four lens-surface uniforms of focus step `step` for packets `pid`."""
    pid = np.asarray(pid, np.uint64)
    n = pid.size
    out = philox(np.full(n, step, np.uint64), np.full(n, TAG_OBD, np.uint64), pid & _MASK, pid >> np.uint64(32),
                 seed, stream)
    return np.stack([u01(x) for x in out])


# ------------------------------------------------------------------ optics

def fresnel(cos_i, cos_t, ni, nt):
    """Unpolarised reflectance, 0.5 [sin^2(ai-at)/sin^2(ai+at) + tan^2(ai-at)/tan^2(ai+at)]
    (_classes.py:212-215) written with cosines; finite at normal incidence."""
    rs = (ni * cos_i - nt * cos_t) / (ni * cos_i + nt * cos_t)
    rp = (ni * cos_t - nt * cos_i) / (ni * cos_t + nt * cos_i)
    return 0.5 * (rs * rs + rp * rp)


def _norm(v):
    return v / np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)


def refract(v, nn, ni, nt):
    """(_classes.py:152-188) v' = a v + (cos_t - a cos_i) nn with a = ni/nt, cos_i = nn.v and
    cos_t = sign(cos_i) sqrt(1 - a^2 (1 - cos_i^2)). Returns v', R and the total-internal-
    reflection mask (ai >= asin(nt/ni))."""
    a = ni / nt
    ci = nn[0] * v[0] + nn[1] * v[1] + nn[2] * v[2]
    s2 = a * a * (1 - ci * ci)
    tir = s2 >= 1.0 if ni > nt else np.zeros(ci.shape, bool)
    ct = np.sign(ci) * np.sqrt(np.clip(1 - s2, 0.0, None))
    R = np.where(tir, 1.0, fresnel(np.abs(ci), np.abs(ct), ni, nt))
    vn = a * v + (ct - a * ci) * nn
    return vn, R, tir


def plane_hit(p, v, z):
    # rays parallel to the plane give inf/nan here; the callers drop them by their masks
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (z - p[2]) / v[2]
        return p + a * v


def sphere_hit(p, v, center, r, outward):
    """Lens1 (outward) leaves through the far side of the sphere, lens2 enters the near side."""
    pz = p[2] - center
    A = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
    B = 2 * (p[0] * v[0] + p[1] * v[1] + pz * v[2])
    Cc = p[0] ** 2 + p[1] ** 2 + pz ** 2 - r * r
    disc = B * B - 4 * A * Cc
    ok = disc >= 0
    sq = np.sqrt(np.where(ok, disc, 0.0))
    t = (-B + sq) / (2 * A) if outward else (-B - sq) / (2 * A)
    return p + v * t, ok


def sphere_normal(p, r, outward):
    d = np.sqrt(r * r - (p[0] ** 2 + p[1] ** 2))
    dx, dy = -p[0] / d, -p[1] / d
    nn = np.array([-dx, -dy, -np.ones_like(dx)]) if outward else np.array([dx, dy, -np.ones_like(dx)])
    return _norm(nn)


def geometry(params, Z):
    """Planes and sphere centres of opticalUnit at focus offset Z (_classes.py:454-460)."""
    P = params
    zl1 = -P["bfl_1"] + Z
    zl2 = zl1 - P["ct_1"] - P["slit_thickness"] * 2 - P["ct_2"] - P["distance_2slits"]
    zs1 = zl1 - P["ct_1"]
    zs2 = zl2 + P["ct_2"] + P["slit_thickness"]
    zpd = zl2 - P["bfl_2"] - P["pd_poit_correction"]
    return dict(z_lens1=zl1, c_lens1=zl1 - (P["ct_1"] - P["r_1"]), z_slit1=zs1, z_lfp=zs1 - P["slit_thickness"] - 10,
                z_slit2=zs2, z_lens2=zl2, c_lens2=zl2 + (P["ct_2"] - P["r_2"]), z_pd=zpd)


LFP_ANGLE = 68 / 180      # the reference's LdFixingPart.angle, in radians as written (_classes.py:338)


def lfp_keep(p):
    ang, wd = LFP_ANGLE, 6.0
    a1 = 1 / math.tan(ang) - 1 / (math.sin(ang) * math.cos(ang))
    b1 = wd / math.cos(ang)
    a2 = (math.cos(ang) + 1) / math.sin(ang)
    return (p[1] > a1 * p[0] + b1) | (p[1] < a2 * p[0] - wd) | (p[1] < -wd)


def trace_step(params, Z, p, v, w, u=None, mode="stochastic"):
    """Weight of each packet that reaches the photodiode at focus offset Z (0 elsewhere).
    p, v: [3][n] at the sample surface (p_z = 0), w: [n]; u: [4][n] surface uniforms."""
    P = params
    G = geometry(P, Z)
    n1, n2 = P["n_1"], P["n_2"]
    r_out1, r_out2 = P["outerD_1"] / 2, P["outerD_2"] / 2
    d_out = P["slit_D"] / 2 + P["slit_width"] / 2
    d_in = P["slit_D"] / 2 - P["slit_width"] / 2
    alive = np.ones(w.shape, bool)
    wt = w.astype(np.float64).copy()
    p = p.astype(np.float64)
    v = _norm(v.astype(np.float64))
    nz = np.array([0.0, 0.0, -1.0])[:, None] * np.ones_like(w)[None, :]

    def cut(R, s, tir):
        nonlocal wt, alive
        alive &= ~tir
        if mode == "expect":
            wt = wt * (1 - R)
        else:
            alive &= u[s] > R

    def radial(q):
        return np.sqrt(q[0] ** 2 + q[1] ** 2)

    # lens 1: plano face (air -> glass), convex face (glass -> air)
    p = plane_hit(p, v, G["z_lens1"])
    alive &= radial(p) <= r_out1
    v, R, tir = refract(v, nz, 1.0, n1)
    cut(R, 0, tir)
    v = _norm(v)
    p, ok = sphere_hit(p, v, G["c_lens1"], P["r_1"], outward=True)
    alive &= ok & (radial(p) <= r_out1)
    with np.errstate(invalid="ignore"):
        nn = sphere_normal(p, P["r_1"], outward=True)
    v, R, tir = refract(v, nn, n1, 1.0)
    cut(R, 1, tir)
    v = _norm(v)
    # slit 1, LD mount, slit 2: front and back face of each
    for z0, th, keep in ((G["z_slit1"], P["slit_thickness"], "slit"),
                         (G["z_lfp"], 5.0, "lfp"),
                         (G["z_slit2"], P["slit_thickness"], "slit")):
        if keep == "lfp" and not P["ld_fix_part"]:
            continue
        for z in (z0, z0 - th):
            p = plane_hit(p, v, z)
            if keep == "slit":
                rr = radial(p)
                alive &= (rr < d_out) & (rr > d_in)
            else:
                alive &= lfp_keep(p)
    # lens 2: convex face (air -> glass), plano face (glass -> air)
    p, ok = sphere_hit(p, v, G["c_lens2"], P["r_2"], outward=False)
    alive &= ok & (radial(p) <= r_out2)
    with np.errstate(invalid="ignore"):
        nn = sphere_normal(p, P["r_2"], outward=False)
    v, R, tir = refract(v, nn, 1.0, n2)
    cut(R, 2, tir)
    v = _norm(v)
    p = plane_hit(p, v, G["z_lens2"])
    alive &= radial(p) <= r_out2
    v, R, tir = refract(v, nz, n2, 1.0)
    cut(R, 3, tir)
    v = _norm(v)
    # photodiode
    p = plane_hit(p, v, G["z_pd"])
    alive &= radial(p) < P["d_pd"] / 2
    alive &= np.isfinite(p).all(axis=0)
    return np.where(alive, wt, 0.0)


def select(params, p, v, w):
    """opticalAnalysisMeth's preparation (_classes.py:498-520): optional side view, the
    packets leaving toward the receiver, p_z = 0, optional inversion. Returns the kept
    packets and their indices in the exit arrays (the Philox packet ids)."""
    P = params
    p = np.array(p, np.float64)
    v = np.array(v, np.float64)
    if P.get("side"):
        v = v[[2, 1, 0]]
        p = p[[2, 1, 0]]
        p[0] -= P["side"] / 2
        p[2] += P["side"] / 2
    idx = np.where(v[2] > 0)[0] if P.get("inversion") else np.where(v[2] < 0)[0]
    p = p[:, idx]
    p[2] = 0
    v = v[:, idx]
    if P.get("inversion"):
        p[1] *= -1
        v[2] *= -1
        v[1] *= -1
    return p, v, np.asarray(w, np.float64)[idx], idx


def steps(params):
    return np.arange(params["start"], params["end"], params["split"])


def obd_curve(params, p, v, w, seed=0, stream=0, mode="stochastic", side_obd=False):
    """I(Z) over the focus steps: OBD (side_obd False) or SideOBD (side_obd True, whose x
    shifts by -(Z + dz) tan(theta0) per step, _classes.py:545-565)."""
    P = dict(params)
    q, d, wt, idx = select(P, p, v, w) if not side_obd else _select_side(P, p, v, w)
    Zs = steps(P)
    out = np.empty(Zs.size)
    shift = None
    if side_obd:
        th0 = initial_angle(P["slit_D"] / 2, P["r_1"], P["wavelength"], "N-BK7")
        dz = real_bfl(P["slit_D"] / 2, P["r_1"], P["ct_1"], P["wavelength"], "N-BK7") - P["bfl_1"]
        shift = (th0, dz)
    x0 = q[0].copy()
    for i, Z in enumerate(Zs):
        if shift is not None:
            q[0] = x0 - (Z + shift[1]) * math.tan(shift[0])
        u = step_uniforms(i, idx, seed, stream) if mode == "stochastic" else None
        out[i] = trace_step(P, Z, q, d, wt, u=u, mode=mode).sum()
    return Zs, out


def _select_side(P, p, v, w):
    p = np.array(p, np.float64)
    v = np.array(v, np.float64)
    idx = np.where(v[2] < 0)[0]
    p = p[:, idx]
    p[2] = 0
    return p, v[:, idx], np.asarray(w, np.float64)[idx], idx
