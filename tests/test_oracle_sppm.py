"""Pins the oracle SPPM integrator (oracle/oracle_sppm.cpp: sppm_eye.rgen,
sppm_light.rgen, gather.comp, composite.comp and the grid passes restated).
CPU only.

The reference's photon table keeps one photon per hash bucket with an
insertion count, and the gather weighs the bucket's photon by that count and
skips it when its path is too long: in expectation over which photon wins a
bucket this sums the flux of every photon of an admissible length, so on a
closed emitting Lambert box (a furnace) SPPM converges to the Path
integrator's Le (1 - rho^D) / (1 - rho) away from the walls' edges."""
import numpy as np
import pytest

import oracle
from scenes import DescScene, box_faces, camera


def _furnace(rho, Le):
    s = DescScene()
    for face in box_faces((-1, -1, -1), (1, 1, 1)):
        s.mesh(face, s.material(bsdf_type=1, bsdf_props=9, albedo=(rho,) * 3, emissive_factor=(Le,) * 3))
    return s.build()


@pytest.mark.parametrize("D", [2, 3, 5])
def test_emissive_furnace_closed_form(D):
    """Camera at the box centre looking at the middle of a wall: emission at
    depth 0, NEE at the first vertex and photons of every admissible length
    add up to Le (1 - rho^D) / (1 - rho) (the photon radius, 0.1, is small
    against the wall: the central pixels' disks stay on it).  Measured over
    seeds and depths: within 1.1%."""
    rho, Le = 0.5, 1.0
    acc = oracle.Accel(_furnace(rho, Le))
    cam = camera([0.0, 0.0, 0.0], [0.0, 0.0, -1.0], 30)
    img, _, _, _, _ = acc.render_sppm(cam, 128, 128, D, frames=24, base_radius=0.1, seed=3)
    want = Le * (1 - rho ** D) / (1 - rho)
    got = img[32:96, 32:96, :3].mean()
    assert got == pytest.approx(want, rel=0.025)


def test_grid_spans_every_pixel_point_and_radius():
    """The grid of a frame spans min(p - r) .. max(p + r) over all pixels
    with cells of the largest radius (max.comp, min.comp, calc_bounds.comp):
    in the furnace every pixel sees a wall, so the bounds are those of the
    recorded points (within the box) grown by the base radius."""
    acc = oracle.Accel(_furnace(0.5, 1.0))
    cam = camera([0.0, 0.0, 0.0], [0.0, 0.0, -1.0], 120)
    r = 0.1
    _, _, _, grid, _ = acc.render_sppm(cam, 32, 32, 3, frames=1, base_radius=r)
    mn, mx, res = grid[:3], grid[3:6], grid[6:]
    assert np.all(mn >= -1 - r - 1e-5) and np.all(mx <= 1 + r + 1e-5)
    assert mn[2] == pytest.approx(-1 - r, abs=1e-5)  # the back wall is seen
    ext = mx - mn
    base = int(ext.max() / r)
    assert list(res) == [max(int(np.float32(base) * e / ext.max()), 1) for e in ext]


def test_running_mean_and_state_between_calls():
    """Frames in one call equal frames in several calls with the state carried
    over (composite.comp running mean; per-pixel radius, N and visible
    points persist)."""
    acc = oracle.Accel(_furnace(0.6, 1.0))
    cam = camera([0.1, -0.2, 0.3], [0.3, 0.2, -1.0], 60)
    a, _, _, _, _ = acc.render_sppm(cam, 16, 16, 4, frames=3, seed=5, base_radius=0.2)
    st = oracle.SppmState(16, 16)
    b = None
    for f in range(3):
        b, _, _, _, st = acc.render_sppm(cam, 16, 16, 4, frame_start=f, frames=1, seed=5, base_radius=0.2,
                                         state=st, accum=b)
    assert np.array_equal(a, b)
