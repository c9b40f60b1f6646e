"""Pins the oracle VCM integrator (vcm_light.rgen / vcm_eye.rgen /
vcm_commons.glsl restated) with closed forms, agreement with the Path
oracle, and a brute-force check of the photon grid.  CPU only."""
import numpy as np
import pytest

import oracle
from scenes import camera
from test_oracle_bdpt import _emissive_box, sealed_cornell


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("corrected", [False, True])
@pytest.mark.parametrize("vm", [False, True])
def test_emissive_furnace_closed_form(D, corrected, vm):
    """Emitting Lambert box: L = Le (1 - rho^D) / (1 - rho).  Every VCM
    strategy contributes (emission hits, NEE, light-vertex connections,
    camera splats and, with merging, photon gathers), so their MIS weights
    must partition unity.  For D <= 2 at most one light vertex can connect
    to an eye vertex, where the reference's overwrite of vertex-connection
    results (vcm_commons.glsl:240) is exact."""
    rho, Le = 0.5, 1.0
    acc = oracle.Accel(_emissive_box(rho, Le))
    cam = camera([0.1, -0.2, 0.3], [0.3, 0.2, -1.0], 60)
    img, _, _, _ = acc.render_vcm(cam, 24, 24, D, frames=96, corrected=corrected, use_vm=vm, radius_factor=2.0)
    want = Le * (1 - rho ** D) / (1 - rho)
    assert img[..., :3].mean() == pytest.approx(want, rel=0.01)
    if D == 1:
        assert np.all(img[..., :3] == np.float32(Le))


def test_corrected_vcm_agrees_with_path():
    """Corrected VCM (connections summed, all photons kept) converges to the
    Path image on the sealed Cornell box, with and without merging (radius
    0.01 of the scene's bounding radius, shrinking per frame)."""
    acc = oracle.Accel(sealed_cornell())
    cam = camera([0, 0, 3.9], [0, 0, -1], 45)
    W = H = 32
    D = 5
    p, _, _, _ = acc.render_path(cam, W, H, D, frames=512)
    rows = slice(0, 24)  # the directly seen light aliases differently under the two jitters
    mp = p[rows, :, :3].mean()
    blk = lambda x: x[rows, :, :3].mean(-1).reshape(3, 8, 4, 8).mean((1, 3))
    for vm in (False, True):
        v, _, _, _ = acc.render_vcm(cam, W, H, D, frames=128, corrected=True, use_vm=vm, radius_factor=1.0)
        assert v[rows, :, :3].mean() == pytest.approx(mp, rel=0.03)
        assert np.all(np.abs(blk(v) - blk(p)) <= 0.15 * blk(p) + 0.006)


def test_reference_mode_loses_overwritten_connections():
    """The reference keeps only the last vertex connection / merge per eye
    vertex, so with several connectable light vertices (D = 5) its image is
    darker than the corrected estimator's on the same samples, while the two
    agree where a single candidate exists (D = 2, checked above)."""
    acc = oracle.Accel(sealed_cornell())
    cam = camera([0, 0, 3.9], [0, 0, -1], 45)
    ref, _, _, _ = acc.render_vcm(cam, 24, 24, 5, frames=32, corrected=False, use_vm=True, radius_factor=1.0)
    cor, _, _, _ = acc.render_vcm(cam, 24, 24, 5, frames=32, corrected=True, use_vm=True, radius_factor=1.0)
    assert ref[..., :3].mean() < 0.99 * cor[..., :3].mean()


def test_photon_grid_matches_brute_force():
    """The hash grid (cell = floor((p - min) / (max - min) * res), hashed by
    utils.glsl:62-65 into 10 W H buckets) returns exactly the photons a
    brute-force radius search finds."""
    acc = oracle.Accel(sealed_cornell())
    cam = camera([0, 0, 3.9], [0, 0, -1], 45)
    for factor in (1.0, 5.0):
        assert acc.vcm_grid_selfcheck(cam, 32, 32, 5, factor, nq=300) == 0


def test_radius_schedule_and_grid():
    """VCM.cpp:140-158: r_n = R f / 100 / (n + 1)^((1 - alpha) / 2) with R the
    half diagonal of the scene box; the grid resolution is the box extent
    over r along the longest axis, scaled per axis."""
    d = sealed_cornell()
    R = float(np.linalg.norm(np.array(d.scene_max[:]) - np.array(d.scene_min[:])) / 2)
    for n, alpha in ((0, 2 / 3), (9, 2 / 3), (99, 0.75)):
        r, g = oracle.vcm_radius_grid(d, 1.0, alpha, n)
        assert r == pytest.approx(R * 0.01 / (n + 1) ** (0.5 * (1 - alpha)), rel=1e-6)
        ext = np.array(d.scene_max[:]) - np.array(d.scene_min[:])
        base = int(ext.max() / r)
        assert g == tuple(max(int(e * base / ext.max()), 1) for e in ext)


def test_running_mean_and_splat_buffer():
    """Frames accumulate as mix(old, new, 1/(n+1)) (vcm_eye.rgen:88-97) and the
    splat buffer is read and cleared by the eye pass, so one call of 3 frames
    equals 3 calls of one."""
    acc = oracle.Accel(sealed_cornell())
    cam = camera([0, 0, 3.9], [0, 0, -1], 45)
    kw = dict(seed=5, use_vm=True, radius_factor=2.0)
    a, _, _, _ = acc.render_vcm(cam, 12, 12, 4, frames=3, **kw)
    b = None
    sp = np.zeros((12, 12, 3), np.float32)
    for f in range(3):
        b, _, _, _ = acc.render_vcm(cam, 12, 12, 4, frame_start=f, frames=1, accum=b, splat=sp, **kw)
    assert np.array_equal(a, b)
