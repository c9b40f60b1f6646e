"""Instance transforms (lt_scene_desc::world_matrices / index_counts, the
reference's TLAS instances with LumenPrimMesh::world_matrix,
Integrator.cpp:147-158, LumenScene.h:32).  CPU only: the oracle's
flattened triangle list, hits and payload against closed forms.

An instance here is a mesh that shares another mesh's index and vertex ranges
and has its own object-to-world matrix W.  ray.rchit:62-80 reports the
world-space point W (p, 1), normals through the inverse transpose of W's
linear part, and the world-space triangle area."""
import numpy as np
import pytest

import lumen_amd as la
import oracle
from scenes import DescScene, quad


def _rot(axis, deg):
    a = np.radians(deg)
    c, s = np.cos(a), np.sin(a)
    x, y, z = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def _world(R, scale, t):
    W = np.eye(4)
    W[:3, :3] = R @ np.diag(scale)
    W[:3, 3] = t
    return W


def _scene():
    """A unit quad in the z = 0 plane (normal +z) as the geometry, shown once
    with the identity and twice more as instances."""
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.5, 0.5, 0.5))
    s.mesh(quad([-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]), m)
    W1 = _world(np.eye(3), (2.0, 2.0, 2.0), (0.0, 0.0, -5.0))          # scaled, pushed back
    W2 = _world(_rot((0, 1, 0), 90.0), (1.0, 3.0, 1.0), (4.0, 0.0, 0.0))  # turned to face +x, stretched in y
    s.instance(0, m, W1)
    s.instance(0, m, W2)
    return s, s.build(), (np.eye(4), W1, W2)


def test_flattened_triangles_are_the_transformed_geometry():
    s, d, W = _scene()
    sc = la.Scene.__new__(la.Scene)
    sc.desc = d
    tris = sc.triangles()
    assert tris.shape == (6, 3, 3) and sc.num_triangles == 6
    base = tris[:2].astype(np.float64)
    for k in (1, 2):
        want = base @ W[k][:3, :3].T + W[k][:3, 3]
        assert np.allclose(tris[2 * k:2 * k + 2], want, atol=1e-6)


def test_hits_on_instances_closed_form():
    """A ray down -z hits the identity quad at t = 3 (from z = 3); moved off
    it, the scaled instance at z = -5 at t = 8; a ray down -x at y = 1.2
    misses the identity quad and hits the instance turned to face +x at
    x = 4 (stretched to |y| <= 1.5) at t = 6."""
    _, d, _ = _scene()
    acc = oracle.Accel(d)
    rays = la.make_rays(np.array([[0.1, 0.2, 3.0], [0.8, 0.0, 3.0], [10.0, 1.2, 0.2]], np.float32),
                        np.array([[0, 0, -1], [0, 0, -1], [-1, 0, 0]], np.float32), tmin=0.0, tmax=100.0)
    for brute in (True, False):
        h = acc.closest(rays, brute=brute)
        assert list(h["inst"]) == [0, 1, 2]
        assert list(h["t"]) == [np.float32(3.0), np.float32(8.0), np.float32(6.0)]


def test_payload_is_transformed_like_the_closest_hit_shader():
    """Instance 2 (W2: 90 degrees about +y, then stretched by 3 in y): the
    object normal +z becomes +x; the world point is W2 (p, 1); the area is
    the world triangle's (0.5 x 3 = 1.5 per triangle of the unit quad)."""
    _, d, W = _scene()
    acc = oracle.Accel(d)
    r = la.make_rays(np.array([[10.0, 1.2, 0.2]], np.float32), np.array([[-1.0, 0.0, 0.0]], np.float32), 0.0, 100.0)
    h = acc.closest(r)
    p = acc.payload(h[0], r[0])
    n_g, n_s, pos, area = p[0:3], p[3:6], p[6:9], p[11]
    assert np.allclose(np.abs(n_g), [1, 0, 0], atol=1e-6) and np.allclose(n_s, [1, 0, 0], atol=1e-6)
    assert np.allclose(pos, [4.0, 1.2, 0.2], atol=1e-5)
    assert area == pytest.approx(1.5, rel=1e-6)
    # the stretched instance's shading normal is the inverse transpose's image of the object normal
    # (a tilted object normal would tilt the other way under a non-uniform scale)
    A = W[2][:3, :3]
    n_obj = np.array([0.0, 0.0, 1.0])
    want = np.linalg.inv(A).T @ n_obj
    assert np.allclose(n_s, want / np.linalg.norm(want), atol=1e-6)


def test_light_on_an_instance_samples_world_points():
    """An emissive instance: sampled emission points lie on the world-space
    quad, the normal is the transformed one and the area pdf is
    1 / (world triangle area) / (light triangles) (commons.glsl:335-406);
    the translation does not enter the area (the edges are vectors)."""
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.5, 0.5, 0.5))
    e = s.material(bsdf_type=1, bsdf_props=9, albedo=(1, 1, 1), emissive_factor=(3, 3, 3))
    s.mesh(quad([-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]), m)
    W = _world(_rot((1, 0, 0), 90.0), (2.0, 1.0, 4.0), (0.0, 3.0, 0.0))  # facing -y at y = 3, 2 x 1
    s.instance(0, e, W)
    d = s.build()
    assert d.num_lights == 1 and d.lights[0].prim_mesh_idx == 1 and d.total_light_triangle_cnt == 2
    rng = np.random.default_rng(3)
    for _ in range(20):
        out = oracle.sample_light_Le(d, rng.random(6).astype(np.float32))
        pos, n, pdf_a = out[3:6], out[9:12], out[13]
        assert pos[1] == pytest.approx(3.0, abs=1e-5)
        assert -1.0 - 1e-5 <= pos[0] <= 1.0 + 1e-5 and -0.5 - 1e-5 <= pos[2] <= 0.5 + 1e-5
        assert np.allclose(n, [0, -1, 0], atol=1e-6)
        assert pdf_a == pytest.approx(1.0 / 1.0 / 2.0, rel=1e-5)  # each world triangle has area 1
