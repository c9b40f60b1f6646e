"""Pins the oracle BDPT integrator (bdpt.rgen / bdpt_commons.glsl restated)
with closed forms and with agreement against the independently pinned Path
oracle.  CPU only.

BDPT estimates each path length with every (s, t) strategy and combines them
with calc_mis_weight; a wrong pdf, a dropped strategy, a mis-projected splat
or MIS weights that do not sum to one change the expectation, which the
closed-form scenes below detect.
"""
import numpy as np
import pytest

import oracle
from scenes import DescScene, box_faces, camera, quad


def _emissive_box(rho, Le):
    s = DescScene()
    for face in box_faces((-1, -1, -1), (1, 1, 1)):
        s.mesh(face, s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3, emissive_factor=(Le,) * 3))
    return s.build()


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("power", [False, True])
def test_emissive_furnace_closed_form(D, power):
    """Camera inside a closed, emitting Lambert box (albedo rho, radiance Le).
    Paths of up to D - 1 bounces (s + t - 2 <= max_depth - 1) give
    L = Le (1 - rho^D) / (1 - rho).  Every strategy contributes here: s = 0
    emission, s = 1 NEE, s >= 2 connections and t = 1 splats through the
    raster projection, so the balance (reference) and power weights must both
    partition unity."""
    rho, Le = 0.5, 1.0
    acc = oracle.Accel(_emissive_box(rho, Le))
    cam = camera([0.1, -0.2, 0.3], [0.3, 0.2, -1.0], 60)
    img, _, _, _ = acc.render_bdpt(cam, 24, 24, D, frames=128, power=power)
    want = Le * (1 - rho ** D) / (1 - rho)
    got = img[..., :3].mean()
    assert got == pytest.approx(want, rel=0.01)
    if D == 1:  # only the s = 0, t = 2 strategy: exact
        assert np.all(img[..., :3] == np.float32(Le))


def test_spot_lit_lambert_plane_closed_form():
    """Spot light of intensity I at height h over a Lambert plane: the point
    below it has radiance rho/pi * I / h^2 (NEE with s = 1 plus the t = 1
    light-tracing splats, MIS-weighted against each other)."""
    rho, I, h = 0.7, 20.0, 3.0
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3)
    s.mesh(quad([-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]), m)
    s.analytic.append(dict(pos=(0, h, 0), to=(0, 0, 0), L=(I, I, I), flags=1 | 16 | 32))
    acc = oracle.Accel(s.build())
    cam = camera([0.0, 10.0, 0.0], [0.0, -1.0, 0.0], 1.0)
    img, _, _, _ = acc.render_bdpt(cam, 9, 9, 2, frames=16)
    want = rho / np.float32(3.14159265359) * I / h ** 2
    assert img[4, 4, 0] == pytest.approx(want, rel=0.01)


def _ceiling_rect(xa, xb, za, zb, y=1.0):
    return quad([xa, y, za], [xb, y, za], [xb, y, zb], [xa, y, zb])  # normal -y


def sealed_cornell():
    """Cornell-like box open at +z whose quad light (albedo 0) fills a hole of
    the ceiling, so its unlit back face is never seen: the reference Path
    (two-sided NEE, stale-payload MIS hits on a reflective light) and BDPT
    (one-sided emission sampling) then estimate the same image."""
    s = DescScene()
    white = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.73,) * 3)
    red = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.65, 0.05, 0.05))
    green = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.12, 0.45, 0.15))
    light = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.0,) * 3, emissive_factor=(17.0, 12.0, 4.0))
    f = box_faces((-1, -1, -1), (1, 1, 1))
    for i, mat in ((0, white), (2, white), (4, red), (5, green)):
        s.mesh(f[i], mat)
    for r in [(-1, -0.25, -1, 1), (0.25, 1, -1, 1), (-0.25, 0.25, -1, -0.25), (-0.25, 0.25, 0.25, 1)]:
        s.mesh(_ceiling_rect(*r), white)
    s.mesh(box_faces((-0.6, -1, -0.6), (-0.1, 0.2, -0.1)), white)
    s.mesh(_ceiling_rect(-0.25, 0.25, -0.25, 0.25), light)
    return s.build()


def test_agrees_with_path_and_power_heuristic():
    """BDPT (balance), BDPT (power) and Path converge to the same radiance
    away from the directly seen light (BDPT samples pixel centres, Path
    jitters, so the light's edge pixels alias differently)."""
    acc = oracle.Accel(sealed_cornell())
    cam = camera([0, 0, 3.9], [0, 0, -1], 45)
    W = H = 32
    D = 5
    b, _, _, _ = acc.render_bdpt(cam, W, H, D, frames=128)
    bp, _, _, _ = acc.render_bdpt(cam, W, H, D, frames=128, power=True)
    p, _, _, _ = acc.render_path(cam, W, H, D, frames=512)
    rows = slice(0, 24)  # the light is seen in rows 25-26
    mb, mbp, mp = (x[rows, :, :3].mean() for x in (b, bp, p))
    assert mb == pytest.approx(mp, rel=0.03)
    assert mbp == pytest.approx(mb, rel=0.02)
    blk = lambda x: x[rows, :, :3].mean(-1).reshape(3, 8, 4, 8).mean((1, 3))
    assert np.all(np.abs(blk(b) - blk(p)) <= 0.15 * blk(p) + 0.005)


def test_running_mean_and_splat_buffer():
    """Frames accumulate as mix(old, new, 1/(n+1)) (bdpt.rgen:82-90); the
    splat buffer carries between calls and is cleared by the pixel that
    reads it (bdpt.rgen:79-81), so one call of 3 frames equals 3 calls."""
    acc = oracle.Accel(sealed_cornell())
    cam = camera([0, 0, 3.9], [0, 0, -1], 45)
    a, _, _, _ = acc.render_bdpt(cam, 12, 12, 4, frames=3, seed=7)
    b = None
    sp = np.zeros((12, 12, 3), np.float32)
    for f in range(3):
        b, _, _, _ = acc.render_bdpt(cam, 12, 12, 4, frame_start=f, frames=1, seed=7, accum=b, splat=sp)
    assert np.array_equal(a, b)
    assert not np.any(sp)
