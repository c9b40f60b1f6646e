"""GPU Path integrator vs the oracle (path.rgen restated on the CPU).

Same scene, camera, seed and frame indices: the per-pixel pcg4d stream is a
pure function of (x, y, frame) (path.rgen:23), so the GPU image replays the
oracle's.  Transcendentals (cos/sin/pow) differ in the last ulp between the
device library and libm, which flips a rare discrete decision; the bar is
therefore per-pixel |gpu - cpu| <= 1e-3 * (1 + |cpu|) on >= 99.5% of the
pixels and a relative image RMSE <= 1%.
"""
import numpy as np
import pytest

import lumen_amd as la
from lumen_amd import _abi
import oracle

pytestmark = pytest.mark.gpu


def _render_both(ctx, kind, W, H, frames, depth=None, **kw):
    sc = la.Scene.procedural(kind, seed=1, **kw)
    it = la.Path(sc, ctx, W, H, path_length=depth)
    it.init()
    it.create_accel()
    it.render(spp=frames)
    g = it.output_tex
    acc = oracle.Accel(sc.desc)
    icfg = sc.integrator_cfg()
    o, _, _, _ = acc.render_path(it.camera, W, H, it.path_length, sky_col=tuple(icfg.sky_col), frames=frames)
    it.destroy()
    return g, o


@pytest.mark.parametrize("kind,W,H,frames,kw", [
    ("cornell", 64, 64, 4, {}),
    ("caustics", 48, 40, 2, {}),
    ("diff_wall", 48, 48, 2, {}),
    ("terrain", 64, 36, 1, dict(terrain_tris=20000, num_instances=100, num_point_lights=8)),
])
def test_path_replays_oracle(gpu_ctx, kind, W, H, frames, kw):
    print(f"[gpu] path {kind} {W}x{H} x{frames}", flush=True)
    g, o = _render_both(gpu_ctx, kind, W, H, frames, **kw)
    assert np.all(np.isfinite(g))
    d = np.abs(g[..., :3] - o[..., :3])
    ok = np.all(d <= 1e-3 * (1.0 + np.abs(o[..., :3])), axis=-1)
    rel = np.sqrt(np.mean(d ** 2)) / max(np.sqrt(np.mean(o[..., :3] ** 2)), 1e-6)
    print(f"[gpu] {kind}: pixels matching {ok.mean():.4f}, rel rmse {rel:.2e}", flush=True)
    assert ok.mean() >= 0.995
    assert rel <= 0.01


def test_running_mean_equals_single_launch(gpu_ctx):
    """Frames rendered one launch at a time accumulate exactly like one
    multi-frame launch (path.rgen:106-112 running mean)."""
    sc = la.Scene.procedural("cornell")
    a = la.Path(sc, gpu_ctx, 32, 32)
    a.init()
    a.create_accel()
    a.render(spp=3)
    one = a.output_tex
    a.reset()
    for _ in range(3):
        a.render(spp=1)
        a.update()
    three = a.output_tex
    a.destroy()
    assert np.array_equal(one, three)


@pytest.mark.parametrize("kind,W,H,frames,kw", [
    ("cornell", 64, 64, 3, {}),
    ("terrain", 96, 54, 2, dict(terrain_tris=20000, num_instances=100, num_point_lights=8)),
    ("caustics", 48, 40, 2, {}),
])
@pytest.mark.parametrize("strategy", ["wavefront", "persistent"])
def test_execution_forms_equal_megakernel(gpu_ctx, kind, W, H, frames, kw, strategy):
    """Same per-path random streams and decisions: the execution forms differ
    only in the float summation order of the per-vertex terms (wavefront) or
    in which lane runs a path (persistent, path regeneration)."""
    print(f"[gpu] {strategy} vs megakernel {kind}", flush=True)
    sc = la.Scene.procedural(kind, seed=1, **kw)
    imgs = []
    for strat in ("megakernel", strategy):
        it = la.Path(sc, gpu_ctx, W, H, strategy=strat)
        it.init()
        it.create_accel()
        it.render(spp=frames)
        imgs.append(it.output_tex)
        it.destroy()
    a, b = imgs
    assert np.all(np.isfinite(b))
    d = np.abs(a[..., :3] - b[..., :3])
    assert np.all(d <= 1e-4 * (1.0 + np.abs(a[..., :3]))), d.max()


@pytest.mark.parametrize("kind,kw", [
    ("cornell", {}),
    ("caustics", {}),
    ("terrain", dict(terrain_tris=20000, num_instances=100, num_point_lights=8)),
])
def test_wavefront_material_sort_is_exact(gpu_ctx, kind, kw):
    """LT_WF_SORT_BSDF only moves where a path is shaded (the hit paths of a
    depth grouped by BSDF type): every path draws the same random numbers and
    writes only its own state, so the image is bit-identical."""
    print(f"[gpu] wavefront material sort {kind}", flush=True)
    sc = la.Scene.procedural(kind, seed=1, **kw)
    imgs = []
    for mode in (_abi.LT_WF_SORT_NONE, _abi.LT_WF_SORT_BSDF):
        gpu_ctx.set_tuning(wf_sort=mode)
        it = la.Path(sc, gpu_ctx, 72, 40, strategy="wavefront")
        it.init()
        it.create_accel()
        it.render(spp=2)
        imgs.append(it.output_tex)
        it.destroy()
    gpu_ctx.set_tuning(wf_sort=_abi.LT_WF_SORT_NONE)
    assert np.any(imgs[0][..., :3] > 0)
    assert np.array_equal(imgs[0], imgs[1])


@pytest.mark.parametrize("strategy", ["megakernel", "persistent", "wavefront"])
def test_corrected_nee_pick_replays_oracle(gpu_ctx, strategy):
    """LT_NEE_CORRECTED: the NEE sample weighted by its own pick probability,
    on the terrain's 8 point lights + 1 two-triangle area light."""
    print(f"[gpu] corrected nee pick {strategy}", flush=True)
    sc = la.Scene.procedural("terrain", seed=1, terrain_tris=20000, num_instances=100, num_point_lights=8)
    W, H = 64, 36
    it = la.Path(sc, gpu_ctx, W, H, strategy=strategy, nee_pick="corrected")
    it.init()
    it.create_accel()
    it.render(spp=2)
    g = it.output_tex
    it.destroy()
    o, _, _, _ = oracle.Accel(sc.desc).render_path(it.camera, W, H, it.path_length, sky_col=tuple(it.sky_col),
                                                   frames=2, nee_pick=1)
    d = np.abs(g[..., :3] - o[..., :3])
    ok = np.all(d <= 1e-3 * (1.0 + np.abs(o[..., :3])), axis=-1)
    rel = np.sqrt(np.mean(d ** 2)) / max(np.sqrt(np.mean(o[..., :3] ** 2)), 1e-6)
    print(f"[gpu] corrected nee {strategy}: pixels matching {ok.mean():.4f}, rel rmse {rel:.2e}", flush=True)
    assert ok.mean() >= 0.995 and rel <= 0.01


def test_counting_totals_are_consistent(gpu_ctx):
    """A counting frame of the megakernel and of the persistent kernel trace
    the same paths with the same traversal: equal ray, node and triangle
    counts; the occlusion queries' part of the nodes and triangles is
    positive and below the totals (the terrain has point lights)."""
    print("[gpu] counting totals", flush=True)
    sc = la.Scene.procedural("terrain", seed=1, terrain_tris=20000, num_instances=100, num_point_lights=8)
    stats = []
    for strat in ("megakernel", "persistent"):
        it = la.Path(sc, gpu_ctx, 96, 54, strategy=strat)
        it.init()
        it.create_accel()
        it.set_counting(True)
        it.render(spp=1)
        s = it.stats()
        stats.append((s.closest_rays, s.shadow_rays, s.nodes_visited, s.tris_tested, s.shadow_nodes_visited,
                      s.shadow_tris_tested))
        it.destroy()
    assert stats[0] == stats[1]
    closest, shadow, nodes, tris, snodes, stris = stats[0]
    assert closest >= 96 * 54 and shadow > 0
    assert 0 < snodes < nodes and 0 < stris < tris
