"""Frames replayed from a captured hipGraph (lt_render_set_graph) against the
plain launches of the same integrator.

Each graph replay is checked against the eager image of the same frame,
with eager work and synchronisations between the replays (readbacks, a raw
trace call, a frame-number jump that rewrites the device frame word, a camera
change that forces a new capture).  Path images are deterministic per pixel,
so they must be bit-identical; BDPT sums its light-tracing splats with float
atomics in a run-dependent order, so it is held to 1e-5 relative.
"""
import numpy as np
import pytest

import lumen_amd as la

pytestmark = pytest.mark.gpu


def _pair(ctx, sc, kind, W, H, **kw):
    its = []
    for graph in (False, True):
        it = la.create_integrator(kind, sc, ctx, W, H, **kw)
        it.seed = 5
        it.init()  # uploads the scene: the BVH is built once both are created
        its.append(it)
    its[1].create_accel()
    its[1].set_graph(True)
    return its


def _check(a, b, exact, tag):
    if exact:
        assert np.array_equal(a, b), tag
    else:
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6), tag


@pytest.mark.parametrize("kind,kw,exact", [
    ("path", dict(strategy="megakernel"), True),
    ("path", dict(strategy="wavefront"), True),
    ("path", dict(strategy="persistent"), True),
    ("bdpt", dict(path_length=5), False),
])
def test_graph_frames_equal_eager_frames(gpu_ctx, kind, kw, exact):
    print(f"[gpu] graph replay {kind} {kw}", flush=True)
    sc = la.Scene.procedural("cornell", seed=1)
    W, H = 48, 40
    eager, graph = _pair(gpu_ctx, sc, kind, W, H, **kw)
    rays = la.make_rays(np.zeros((64, 3), np.float32), np.tile([0, 0, -1], (64, 1)).astype(np.float32), 0.0, 10.0)
    for f in range(4):
        for it in (eager, graph):
            it.render(spp=1)
            it.update()
        gpu_ctx.trace_closest(rays)  # eager work between the replays
        _check(graph.output_tex, eager.output_tex, exact, f"{kind} frame {f}")
    # several frames per call
    for it in (eager, graph):
        it.render(spp=3)
        it.update()
    _check(graph.output_tex, eager.output_tex, exact, f"{kind} frames 4-6")
    # a jump of the frame number (the device word is rewritten)
    for it in (eager, graph):
        it.frame_num = 20
        it.render(spp=1)
        it.update()
    _check(graph.output_tex, eager.output_tex, exact, f"{kind} frame 20")
    # a new camera forces a new capture
    cam = sc.camera(W, H)
    cam.inv_view[12] += 0.05
    for it in (eager, graph):
        it.camera = cam
        it.render(spp=2)
        it.update()
    _check(graph.output_tex, eager.output_tex, exact, f"{kind} moved camera")
    for it in (eager, graph):
        it.destroy()


def test_vcm_refuses_graph_mode(gpu_ctx):
    sc = la.Scene.procedural("cornell", seed=1)
    it = la.VCM(sc, gpu_ctx, 16, 16)
    it.init()
    it.create_accel()
    with pytest.raises(la.LtError, match="LT_ERR_UNSUPPORTED"):
        it.set_graph(True)
    it.destroy()
