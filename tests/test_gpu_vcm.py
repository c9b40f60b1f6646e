"""GPU VCM integrator vs the oracle (vcm_light.rgen / vcm_eye.rgen /
vcm_commons.glsl restated), both estimators.

Same scene, camera, seed and frame indices: the light pass of path p and the
eye pass of pixel (x, y) draw their pcg4d streams from per-frame words that
both sides derive from (seed, frame) the same way, so the GPU replays the
oracle's subpaths, connections and merges.  The reference estimator's
photon table keeps, per hash bucket, the photon with the largest (path,
vertex) reference on both sides.  What differs is rounding: the order in
which float atomics sum the splats (and, for the corrected estimator, the
order of photons within a bucket), and last-ulp differences of the device
transcendentals, which flip a rare discrete decision.  The bar is the one of
the Path and BDPT replays: per-pixel |gpu - cpu| <= 1e-3 * (1 + |cpu|) on
>= 99.5% of the pixels and a relative image RMSE <= 1%.
"""
import numpy as np
import pytest

import lumen_amd as la
import oracle
from scenes import camera
from test_gpu_bdpt import _compare, _DescScene
from test_oracle_bdpt import sealed_cornell

pytestmark = pytest.mark.gpu


def _render_both(ctx, sc, W, H, frames, depth, estimator, vm, factor, seed=0):
    it = la.VCM(sc, ctx, W, H, path_length=depth, radius_factor=factor, enable_vm=vm, estimator=estimator)
    it.seed = seed
    it.init()
    it.create_accel()
    it.render(spp=frames)
    g = it.output_tex
    acc = oracle.Accel(sc.desc)
    o, _, _, _ = acc.render_vcm(it.camera, W, H, depth, frames=frames, seed=seed, radius_factor=factor,
                                use_vm=vm, corrected=(estimator == "corrected"), sky_col=tuple(it.sky_col))
    it.destroy()
    return g, o


@pytest.mark.parametrize("estimator", ["reference", "corrected"])
@pytest.mark.parametrize("kind,W,H,frames,kw", [
    ("cornell", 40, 40, 2, {}),
    ("caustics", 48, 40, 2, {}),
    ("terrain", 48, 27, 1, dict(terrain_tris=20000, num_instances=100, num_point_lights=8)),
])
def test_vcm_replays_oracle(gpu_ctx, estimator, kind, W, H, frames, kw):
    print(f"[gpu] vcm {estimator} {kind} {W}x{H} x{frames}", flush=True)
    sc = la.Scene.procedural(kind, seed=1, **kw)
    depth = 8 if kind == "caustics" else 5
    g, o = _render_both(gpu_ctx, sc, W, H, frames, depth, estimator, True, 2.0, seed=3)
    _compare(g, o, f"vcm {estimator} {kind}")


@pytest.mark.parametrize("estimator", ["reference", "corrected"])
@pytest.mark.parametrize("vm", [False, True])
def test_vcm_sealed_cornell(gpu_ctx, estimator, vm):
    print(f"[gpu] vcm sealed cornell {estimator} vm={vm}", flush=True)
    sc = _DescScene(sealed_cornell(), camera([0, 0, 3.9], [0, 0, -1], 45), 5)
    g, o = _render_both(gpu_ctx, sc, 32, 32, 3, 5, estimator, vm, 1.0, seed=11)
    _compare(g, o, f"vcm sealed {estimator} vm={vm}")


def test_vcm_dense_merge_queries(gpu_ctx):
    """A small image with a large merge radius over the caustic: many merge
    queries hold more than 64 photons inside the radius, so the merge
    kernel's LDS queue fills and carries over (corrected estimator)."""
    print("[gpu] vcm dense merge queries", flush=True)
    sc = la.Scene.procedural("caustics", seed=7)
    g, o = _render_both(gpu_ctx, sc, 64, 64, 2, 8, "corrected", True, 4.0, seed=5)
    q, over, mx = oracle.vcm_merge_stats()
    print(f"[gpu] merge queries {q}, with > 64 photons in the radius {over}, at most {mx}", flush=True)
    assert over >= 100 and mx > 1000
    _compare(g, o, "vcm dense merge")
