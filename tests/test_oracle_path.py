"""Pins the oracle Path integrator (path.rgen / pt_commons.glsl restated)
with scenes whose radiance has a closed form.  CPU only."""
import numpy as np
import pytest

import oracle
from scenes import DescScene, box_faces, camera, quad


def test_emissive_furnace_closed_form():
    """Camera inside a closed box whose six faces are diffuse (albedo rho)
    and emit Le.  The Path estimator adds Le at depth 0 and, by NEE + MIS, one
    more scattering order per vertex up to max_depth - 1 vertices, so
    L = Le (1 - rho^D) / (1 - rho)."""
    rho, Le, D = 0.5, 1.0, 6
    s = DescScene()
    for face in box_faces((-1, -1, -1), (1, 1, 1)):
        m = s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3, emissive_factor=(Le,) * 3)
        s.mesh(face, m)
    d = s.build()
    acc = oracle.Accel(d)
    cam = camera([0.1, -0.2, 0.3], [0.3, 0.2, -1.0], 60)
    img, _, _, _ = acc.render_path(cam, 24, 24, D, frames=64)
    want = Le * (1 - rho ** D) / (1 - rho)
    got = img[..., :3].mean()
    assert got == pytest.approx(want, rel=0.01)


def test_spot_lit_lambert_plane_closed_form():
    """A spot light of intensity I straight above a Lambert plane (albedo rho)
    at height h: the radiance of the point below it is rho/pi * I / h^2
    (one NEE sample, delta light, MIS weight 1; depth 2 stops after it)."""
    rho, I, h = 0.7, 20.0, 3.0
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3)
    s.mesh(quad([-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]), m)
    s.analytic.append(dict(pos=(0, h, 0), to=(0, 0, 0), L=(I, I, I), flags=1 | 16 | 32))
    d = s.build()
    acc = oracle.Accel(d)
    cam = camera([0.0, 10.0, 0.0], [0.0, -1.0, 0.0], 1.0)  # narrow view of the point below the light
    img, _, _, _ = acc.render_path(cam, 9, 9, 2, frames=4)
    want = rho / np.float32(3.14159265359) * I / h ** 2
    assert img[4, 4, :3] == pytest.approx([want] * 3, rel=2e-3)


def test_nan_guard_and_running_mean():
    """Frames accumulate as the running mean mix(old, new, 1/(n+1)) (path.rgen:106-112)."""
    s = DescScene()
    for face in box_faces((-1, -1, -1), (1, 1, 1)):
        s.mesh(face, s.material(bsdf_type=1, bsdf_props=9, albedo=(0.4,) * 3, emissive_factor=(1, 1, 1)))
    d = s.build()
    acc = oracle.Accel(d)
    cam = camera([0, 0, 0], [0, 0, -1], 50)
    a, _, _, _ = acc.render_path(cam, 8, 8, 4, frames=3)
    b = None
    for f in range(3):
        b, _, _, _ = acc.render_path(cam, 8, 8, 4, frame_start=f, frames=1, accum=b)
    assert np.array_equal(a, b)


def test_nee_pick_reference_bias_and_corrected_weight():
    """A point light and a two-triangle area light over a Lambert plane, Path
    at depth 2 (direct light only).  Rendered alone, each light is estimated
    without bias by the reference's weight (one emitter).  Together the
    reference weight is light_triangle_count = 3 for every pick (path.rgen:76-78),
    while the pick probabilities are 1/2 (point) and 1/4 (each area triangle):
    the point light comes out 3/2 too bright and the area light 3/4 too dark.
    The corrected weight (nee_pick=1) gives the sum of the two."""
    rho = 0.6

    def scene(point, area):
        s = DescScene()
        m = s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3)
        s.mesh(quad([-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50]), m)
        if area:
            lm = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.0,) * 3, emissive_factor=(8.0, 8.0, 8.0))
            s.mesh(quad([-1.5, 2.0, -0.5], [-0.5, 2.0, -0.5], [-0.5, 2.0, 0.5], [-1.5, 2.0, 0.5]), lm)
        if point:
            s.analytic.append(dict(pos=(1.0, 2.5, 0.0), to=(0, 0, 0), L=(12.0,) * 3, flags=4 | 16 | 32))
        return oracle.Accel(s.build())

    cam = camera([0.0, 10.0, 0.0], [0.0, -1.0, 0.0], 1.0)
    render = lambda acc, pick: acc.render_path(cam, 5, 5, 2, frames=2048, nee_pick=pick)[0][2, 2, 0]
    p_only = render(scene(True, False), 0)
    a_only = render(scene(False, True), 0)
    both = scene(True, True)
    assert render(both, 1) == pytest.approx(p_only + a_only, rel=0.02)
    assert render(both, 0) == pytest.approx(1.5 * p_only + 0.75 * a_only, rel=0.02)
