"""lt_trace_closest_noalloc (closest hits from a fixed device scratch array,
no allocation) against lt_trace_closest with device pointers: bit-exact
hits on two calls in a row, the second one smaller, and after a BVH of
another scene; refusals of bad arguments."""
import ctypes as C

import numpy as np
import pytest

import lumen_amd as la
from lumen_amd import _abi
from raygen import aimed_rays, random_rays

pytestmark = pytest.mark.gpu


def _noalloc(ctx, rays):
    n = len(rays)
    d_rays = ctx.alloc(max(1, rays.nbytes))
    d_out = ctx.alloc(max(1, n * _abi.HIT_DTYPE.itemsize))
    ctx.h2d(d_rays, rays)
    la.check(la.lib().lt_trace_closest_noalloc(ctx.handle, d_rays, n, d_out), ctx.handle, "lt_trace_closest_noalloc")
    ctx.synchronize()
    out = np.zeros(n, _abi.HIT_DTYPE)
    ctx.d2h(out, d_out)
    ctx.free(d_rays)
    ctx.free(d_out)
    return out


def _reference(ctx, rays):
    n = len(rays)
    d_rays = ctx.alloc(max(1, rays.nbytes))
    d_out = ctx.alloc(max(1, n * _abi.HIT_DTYPE.itemsize))
    ctx.h2d(d_rays, rays)
    ctx.trace_closest_device(d_rays, n, d_out)
    ctx.synchronize()
    out = np.zeros(n, _abi.HIT_DTYPE)
    ctx.d2h(out, d_out)
    ctx.free(d_rays)
    ctx.free(d_out)
    return out


def _same(a, b):
    return (np.array_equal(a["inst"], b["inst"]) and np.array_equal(a["prim"], b["prim"]) and
            np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32)) and
            np.array_equal(a["u"].view(np.uint32), b["u"].view(np.uint32)) and
            np.array_equal(a["v"].view(np.uint32), b["v"].view(np.uint32)))


@pytest.mark.parametrize("kind,kw", [
    ("terrain", dict(terrain_tris=200000, num_instances=500, num_point_lights=8)),
    ("caustics", {}),
])
def test_noalloc_matches_trace_closest(kind, kw):
    print(f"[gpu] no-alloc closest hits, {kind}", flush=True)
    ctx = la.Context(0)
    sc = la.Scene.procedural(kind, seed=2, **kw)
    ctx.upload(sc)
    st = ctx.build_accel()
    print(f"[gpu] bvh8 depth {st.max_depth}", flush=True)
    tris = sc.triangles()
    rays = np.concatenate([random_rays(tris, 150000, 3, tmin=0.001, tmax=1e4), aimed_rays(tris, 150000, 4)])
    first = _noalloc(ctx, rays)
    assert _same(first, _reference(ctx, rays))
    assert (first["inst"] != 0xFFFFFFFF).mean() > 0.2
    small = rays[:777]  # a second call, smaller: the counter starts at zero again
    assert _same(_noalloc(ctx, small), _reference(ctx, small))
    ctx.close()


def test_noalloc_refusals(gpu_ctx):
    lib = la.lib()
    h = gpu_ctx.handle
    sc = la.Scene.procedural("cornell")
    gpu_ctx.upload(sc)
    gpu_ctx.build_accel()
    assert lib.lt_trace_closest_noalloc(None, None, 0, None) == 1
    assert lib.lt_trace_closest_noalloc(h, None, 5, None) == 1
    assert lib.lt_trace_closest_noalloc(h, None, 0, None) == 0  # an empty batch is a no-op
