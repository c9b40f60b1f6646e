"""Pins the oracle's traversal: watertight closest / any hit by brute force,
the host BVH against brute force, closed-form hits and the tie rule.  CPU only."""
import numpy as np
import pytest

import lumen_amd as la
import oracle
from raygen import aimed_rays, random_rays
from scenes import DescScene, quad


def _eq(a, b):
    return np.array_equal(a["inst"], b["inst"]) and np.array_equal(a["prim"], b["prim"]) and \
        np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


@pytest.mark.parametrize("kind,kw", [("cornell", {}), ("caustics", dict(sphere_tris=4000)),
                                     ("terrain", dict(terrain_tris=8000, num_instances=50, num_point_lights=2))])
def test_host_bvh_equals_brute_force(kind, kw):
    sc = la.Scene.procedural(kind, seed=2, **kw)
    acc = oracle.Accel(sc.desc)
    t = sc.triangles()
    rays = np.concatenate([random_rays(t, 3000, 1), aimed_rays(t, 3000, 2), random_rays(t, 1000, 3, tmin=0.001)])
    assert _eq(acc.closest(rays), acc.closest(rays, brute=True))
    sh = rays.copy()
    sh["tmax"] = np.random.default_rng(4).random(len(sh)).astype(np.float32) * 2
    assert np.array_equal(acc.any(sh), acc.any(sh, brute=True))
    # any-hit <=> a closest hit exists in [tmin, tmax]
    assert np.array_equal(acc.any(sh, brute=True).astype(bool), acc.closest(sh, brute=True)["inst"] != 0xFFFFFFFF)


def test_closed_form_hit_on_axis_quad():
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(1, 1, 1))
    s.mesh(quad([-1, -1, -2], [1, -1, -2], [1, 1, -2], [-1, 1, -2]), m)
    d = s.build()
    acc = oracle.Accel(d)
    r = la.make_rays(np.array([[0.25, 0.5, 3.0]]), np.array([[0.0, 0.0, -1.0]]))
    h = acc.closest(r, brute=True)[0]
    assert h["inst"] == 0 and h["t"] == np.float32(5.0)
    # barycentrics reproduce the hit point: p = (1-u-v) v0 + u v1 + v v2
    tri = np.array(quad([-1, -1, -2], [1, -1, -2], [1, 1, -2], [-1, 1, -2]), np.float32)[h["prim"]]
    p = (1 - h["u"] - h["v"]) * tri[0] + h["u"] * tri[1] + h["v"] * tri[2]
    assert np.allclose(p, [0.25, 0.5, -2.0], atol=1e-6)
    # tmax just below the hit: miss; tmin just above: miss
    r["tmax"] = np.nextafter(np.float32(5.0), np.float32(0))
    assert acc.closest(r, brute=True)[0]["inst"] == 0xFFFFFFFF
    r["tmax"], r["tmin"] = 1e30, np.nextafter(np.float32(5.0), np.float32(10))
    assert acc.closest(r, brute=True)[0]["inst"] == 0xFFFFFFFF


def test_shared_edge_is_watertight_and_ties_go_to_smaller_index():
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(1, 1, 1))
    s.mesh(quad([-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]), m)   # mesh 0: diagonal (-1,-1)-(1,1)
    s.mesh(quad([-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]), m)   # mesh 1: the same quad again
    d = s.build()
    acc = oracle.Accel(d)
    # rays through points of the shared diagonal edge and the shared vertices
    pts = np.array([[x, x, 0.0] for x in np.linspace(-0.9, 0.9, 37)] + [[1, 1, 0], [-1, -1, 0]], np.float32)
    o = pts + np.array([0.013, -0.027, 2.0], np.float32)
    r = la.make_rays(o, pts - o)
    h = acc.closest(r, brute=True)
    assert np.all(h["inst"] == 0)  # coplanar duplicate: the smaller index wins every tie
    assert np.all(h["prim"] <= 1)
