"""Pins of the oracle's sky for scenes with a directional light
(shade_atmosphere, commons.glsl:156-168; atmosphere.glsl:40-205).

- an independent float64 implementation of the single-scattering model
  written here with numpy from its definition (densities exp(-h/8 km),
  exp(-h/1.2 km) and a 30 km ozone tent at 25 km; an 8-sample midpoint
  optical depth to the top of a 100 km shell; 64 view samples at
  (i/64)^(1 + 8 clamp(1 - h/100 km)) of the length; Rayleigh and Schlick
  phase functions; exposure 20) agrees with the float32 oracle;
- the phase functions integrate to 1 over the sphere;
- the sky is invariant under a rotation of view and light about the
  vertical (the planet's centre lies straight below the origin);
- zenith sky under a high sun is blue (B > G > R); looking toward a sun at
  the horizon it is red (R > G > B);
- scenes without a directional light see the constant sky colour.
A swapped coefficient, a planet centre on another axis, a dropped phase
factor or a wrong height scale each fail one of these."""
import math

import numpy as np
import pytest

import lumen_amd as la
import oracle

R, H, HR, HM = 6371000.0, 100000.0, 8000.0, 1200.0
CR = np.array([5.802, 13.558, 33.100]) * 1e-6
CM = np.array([3.996, 3.996, 3.996]) * 1e-6
CO = np.array([0.650, 1.881, 0.085]) * 1e-6


def _sphere(o, d, c, r):
    oc = o - c
    a, b, cc = d @ d, 2.0 * (oc @ d), oc @ oc - r * r
    disc = b * b - 4.0 * a * cc
    if disc < 0:
        return -1.0, -1.0
    s = math.sqrt(disc)
    return (-b - s) / (2 * a), (-b + s) / (2 * a)


def _height(p):
    return np.linalg.norm(p - np.array([0.0, -R, 0.0])) - R


def _density(h):
    return np.array([math.exp(-max(0.0, h / HR)), math.exp(-max(0.0, h / HM)), max(0.0, 1.0 - abs(h - 25000.0) / 15000.0)])


def _absorb(od):
    return np.exp(-(od[0] * CR + od[1] * CM * 1.1 + od[2] * CO))


def _optical_depth(p, d):
    length = _sphere(p, d, np.array([0.0, -R, 0.0]), R + H)[1]
    step = length / 8
    return sum(_density(_height(p + d * (i + 0.5) * step)) * step for i in range(8))


def sky64(o, d, length, ldir, Lc):
    o, d, ldir, Lc = (np.asarray(x, np.float64) for x in (o, d, ldir, Lc))
    expo = 1.0 + min(max(1.0 - _height(o) / H, 0.0), 1.0) * 8.0
    near, far = _sphere(o, d, np.array([0.0, -R, 0.0]), R + H)
    length = min(length, far)
    if near > 0:
        o = o + d * near
        length -= near
    c = d @ ldir
    pr = 3 * (1 + c * c) / (16 * math.pi)
    k = 1.55 * 0.85 - 0.55 * 0.85 ** 3
    pm = (1 - k * k) / (4 * math.pi * (1 - k * c) ** 2)
    od = np.zeros(3)
    ray = np.zeros(3)
    mie = np.zeros(3)
    prev = 0.0
    for i in range(64):
        t = (i / 64) ** expo * length
        step = t - prev
        p = o + d * t
        dn = _density(_height(p))
        od = od + dn * step
        tt = _absorb(od) * _absorb(_optical_depth(p, ldir))
        ray += tt * pr * dn[0] * step
        mie += tt * pm * dn[1] * step
        prev = t
    return (ray * CR + mie * CM) * Lc * 20.0


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


CASES = [  # origin, view direction, length, light direction
    ([0, 0, 0], [0, 1, 0], 1e4, [0, 0.7071, 0.7071]),
    ([0, 0, 0], [0, 1, 0], 1e9, [0, 0.7071, 0.7071]),
    ([1.5, 2.0, -3.0], [0.3, 0.4, 0.866], 1e4, [-0.2, 0.9, 0.1]),
    ([0, 0, 0], [0, 0.05, 0.9987], 1e9, [0, 0.05, 0.9987]),
    ([0, 50, 0], [0.6, 0.2, -0.77], 1e9, [0.1, 0.3, 0.95]),
    ([0, 2000, 0], [0, 0.01, 1], 1e9, [1, 0.2, 0]),
]


@pytest.mark.parametrize("o,d,length,ldir", CASES)
def test_oracle_sky_matches_a_float64_implementation(o, d, length, ldir):
    d, ldir = _unit(d), _unit(ldir)
    Lc = np.array([5.0, 4.0, 3.0])
    got = oracle.scattering(o, d, length, ldir, Lc).astype(np.float64)
    want = sky64(o, d, length, ldir, Lc)
    assert np.all(want > 0)
    assert np.allclose(got, want, rtol=3e-4, atol=0), (got, want)


def test_phase_functions_are_normalised():
    from scipy.integrate import quad
    for which in ("rayleigh", "mie"):
        total, _ = quad(lambda c: 2 * math.pi * oracle.phase(which, c), -1, 1, points=[0.99, 0.999], limit=200,
                        epsabs=1e-6, epsrel=1e-5)
        assert total == pytest.approx(1.0, rel=2e-4), which
    # the Schlick lobe peaks forward for g > 0
    assert oracle.phase("mie", 1.0) > 50 * oracle.phase("mie", -1.0)


def test_sky_is_invariant_about_the_vertical():
    o, d, ldir = np.array([0.0, 3.0, 0.0]), _unit([0.5, 0.3, 0.2]), _unit([0.1, 0.6, -0.4])
    base = oracle.scattering(o, d, 1e9, ldir, [1, 1, 1])
    for phi in (0.7, 2.0, 4.5):
        c, s = math.cos(phi), math.sin(phi)
        rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        got = oracle.scattering(o, rot @ d, 1e9, rot @ ldir, [1, 1, 1])
        assert np.allclose(got, base, rtol=2e-4), (phi, got, base)
    # tilting the pair away from the vertical changes it
    c, s = math.cos(0.5), math.sin(0.5)
    tilt = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    assert not np.allclose(oracle.scattering(o, tilt @ d, 1e9, tilt @ ldir, [1, 1, 1]), base, rtol=1e-2)


def test_blue_zenith_and_red_sunset():
    z = oracle.scattering([0, 0, 0], [0, 1, 0], 1e9, _unit([0, 1, 1]), [1, 1, 1])
    assert z[2] > z[1] > z[0]
    s = oracle.scattering([0, 0, 0], _unit([0, 0.02, 1]), 1e9, _unit([0, 0.02, 1]), [1, 1, 1])
    assert s[0] > s[1] > s[2]
    # linear in the light's colour
    a = oracle.scattering([0, 0, 0], [0, 1, 0], 1e9, _unit([0, 1, 1]), [2, 0.5, 4])
    assert np.allclose(a, z * np.array([2, 0.5, 4]), rtol=1e-6)


def test_scene_without_directional_light_sees_the_sky_colour():
    sc = la.Scene.procedural("cornell", seed=1)
    acc = oracle.Accel(sc.desc)
    got = acc.atmosphere([0.25, 0.5, 0.75], [0, 0, 0], [0, 1, 0], 1e4)
    assert np.array_equal(got, np.array([0.25, 0.5, 0.75], np.float32))
