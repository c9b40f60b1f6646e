"""GPU BVH8 traversal vs the oracle: bit-exact closest-hit (t, inst, prim, u, v)
and any-hit on seeded ray sets, brute force over every triangle being the
ground truth."""
import numpy as np
import pytest

import lumen_amd as la
import oracle
from raygen import aimed_rays, random_rays

pytestmark = pytest.mark.gpu

SCENES = [
    ("cornell", {}),
    ("caustics", {}),
    ("diff_wall", {}),
    ("terrain", dict(terrain_tris=60000, num_instances=300, num_point_lights=8)),
]


def _setup(ctx, kind, kw, max_leaf=3):
    sc = la.Scene.procedural(kind, seed=3, **kw)
    ctx.upload(sc)
    st = ctx.build_accel(max_leaf)
    assert st.num_tris == sc.num_triangles
    return sc, st


def _compare_closest(g, o, what):
    hit_g, hit_o = g["inst"] != 0xFFFFFFFF, o["inst"] != 0xFFFFFFFF
    bad = np.nonzero((g["inst"] != o["inst"]) | (g["prim"] != o["prim"]) |
                     (g["t"].view(np.uint32) != o["t"].view(np.uint32)) |
                     (g["u"].view(np.uint32) != o["u"].view(np.uint32)) |
                     (g["v"].view(np.uint32) != o["v"].view(np.uint32)))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} mismatches, first {bad[:5]}: gpu {g[bad[:3]]} oracle {o[bad[:3]]}"
    assert np.array_equal(hit_g, hit_o)


@pytest.mark.parametrize("kind,kw", SCENES, ids=[s[0] for s in SCENES])
@pytest.mark.parametrize("max_leaf", [1, 3, 4])
def test_closest_bit_exact_vs_brute_force(gpu_ctx, kind, kw, max_leaf):
    print(f"[gpu] closest {kind} leaf={max_leaf}", flush=True)
    sc, st = _setup(gpu_ctx, kind, kw, max_leaf)
    tris = sc.triangles()
    acc = oracle.Accel(sc.desc)
    n = 4000 if sc.num_triangles < 30000 else 1500
    rays = np.concatenate([random_rays(tris, n, 11), aimed_rays(tris, n, 12)])
    g = gpu_ctx.trace_closest(rays)
    o = acc.closest(rays, brute=True)
    _compare_closest(g, o, f"{kind} brute")
    assert (g["inst"] != 0xFFFFFFFF).mean() > 0.2


@pytest.mark.parametrize("kind,kw", SCENES, ids=[s[0] for s in SCENES])
def test_closest_and_any_bit_exact_vs_host_bvh(gpu_ctx, kind, kw):
    print(f"[gpu] closest+any {kind}", flush=True)
    sc, st = _setup(gpu_ctx, kind, kw)
    tris = sc.triangles()
    acc = oracle.Accel(sc.desc)
    rays = np.concatenate([random_rays(tris, 50000, 21, tmin=0.001, tmax=1e4), aimed_rays(tris, 50000, 22)])
    g = gpu_ctx.trace_closest(rays)
    o = acc.closest(rays)
    _compare_closest(g, o, kind)
    # shadow rays: finite tmax, tmin = 0 (pt_commons.glsl:20-23)
    sh = rays.copy()
    rng = np.random.default_rng(5)
    sh["tmin"] = 0.0
    sh["tmax"] = rng.random(len(sh)).astype(np.float32) * 3.0
    ga = gpu_ctx.trace_any(sh)
    oa = acc.any(sh)
    assert np.array_equal(ga, oa), f"any-hit mismatches: {np.count_nonzero(ga != oa)}"
    # any-hit <=> a closest hit exists within [tmin, tmax]
    gc = gpu_ctx.trace_closest(sh)
    assert np.array_equal(ga.astype(bool), gc["inst"] != 0xFFFFFFFF)


def test_empty_and_ragged_batches(gpu_ctx):
    sc, _ = _setup(gpu_ctx, "cornell", {})
    acc = oracle.Accel(sc.desc)
    tris = sc.triangles()
    assert len(gpu_ctx.trace_closest(random_rays(tris, 0, 1))) == 0
    for n in (1, 63, 64, 65, 129, 1001):
        r = random_rays(tris, n, n)
        _compare_closest(gpu_ctx.trace_closest(r), acc.closest(r, brute=True), f"n={n}")


def test_counting_pass_matches_visits(gpu_ctx):
    sc, st = _setup(gpu_ctx, "terrain", dict(terrain_tris=60000, num_instances=300, num_point_lights=8))
    tris = sc.triangles()
    rays = random_rays(tris, 20000, 7)
    nodes, ntri = gpu_ctx.trace_count(rays)
    assert nodes >= len(rays) and ntri > 0
    nodes2, ntri2 = gpu_ctx.trace_count(rays)
    assert (nodes, ntri) == (nodes2, ntri2)


@pytest.mark.parametrize("kind,kw", SCENES, ids=[s[0] for s in SCENES])
def test_host_traversal_of_exported_bvh8(gpu_ctx, kind, kw):
    """The host restatement of the traversal over the exported BVH8 gives the
    GPU's hits bit for bit.  Its visit counts are the GPU's lower bound: the
    GPU tests a leaf's triangles once enough lanes of the wave have leaves
    queued, so its closest-hit bound shrinks later and it may visit more
    nodes and triangles (the roofline counts what the GPU visits)."""
    print(f"[gpu] exported bvh8 {kind}", flush=True)
    sc, st = _setup(gpu_ctx, kind, kw)
    nodes, tris = gpu_ctx.export_bvh(st)
    t = sc.triangles()
    rays = np.concatenate([random_rays(t, 3000, 31), aimed_rays(t, 3000, 32)])
    g = gpu_ctx.trace_closest(rays)
    h, nc, tc = oracle.bvh8_trace(nodes, tris, rays)
    _compare_closest(g, h, f"{kind} host-bvh8")
    nodes_g, tris_g = gpu_ctx.trace_count(rays)
    nodes_h, tris_h = int(nc.sum()), int(tc.sum())
    print(f"[gpu] {kind}: nodes gpu {nodes_g} host {nodes_h}, triangles gpu {tris_g} host {tris_h}", flush=True)
    assert nodes_h <= nodes_g <= 2 * nodes_h
    assert tris_h <= tris_g <= 2 * tris_h
    sh = rays.copy()
    sh["tmax"] = np.random.default_rng(3).random(len(sh)).astype(np.float32) * 2.0
    ga = gpu_ctx.trace_any(sh)
    ha, _, _ = oracle.bvh8_trace(nodes, tris, sh, any_hit=True)
    assert np.array_equal(ga, ha)


def test_instances_bit_exact_vs_brute_force(gpu_ctx):
    """A scene of instances (lt_scene_desc::world_matrices / index_counts): an
    icosphere-like blob and a quad, each shown 40 times under random
    rotations, non-uniform scales and translations.  The device builder
    flattens them through their transforms as the oracle does, so closest
    hits (t, inst, prim, u, v) and any hits are bit-identical to brute force;
    the payload normals go through the inverse transpose."""
    print("[gpu] instances closest+any", flush=True)
    from scenes import DescScene, quad
    rng = np.random.default_rng(5)
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.5, 0.5, 0.5))
    pts = rng.normal(size=(60, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    blob = np.stack([pts[i:i + 3] for i in range(0, 57, 3)]).astype(np.float32)
    s.mesh(blob, m)
    s.mesh(quad([-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]), m)
    for k in range(80):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        W = np.eye(4)
        W[:3, :3] = R @ np.diag(rng.uniform(0.3, 2.0, 3))
        W[:3, 3] = rng.uniform(-8, 8, 3)
        s.instance(k % 2, m, W)
    d = s.build()
    sc = la.Scene.__new__(la.Scene)
    sc.desc = d
    gpu_ctx.upload(sc)
    st = gpu_ctx.build_accel()
    assert st.num_tris == sc.num_triangles == 19 + 2 + 40 * 19 + 40 * 2
    tris = sc.triangles()
    acc = oracle.Accel(d)
    rays = np.concatenate([random_rays(tris, 20000, 31), aimed_rays(tris, 20000, 32)])
    g = gpu_ctx.trace_closest(rays)
    o = acc.closest(rays, brute=True)
    _compare_closest(g, o, "instances brute")
    assert (g["inst"] != 0xFFFFFFFF).mean() > 0.2 and (g["inst"] >= 2).mean() > 0.2
    sh = rays.copy()
    sh["tmax"] = rng.random(len(sh)).astype(np.float32) * 4
    assert np.array_equal(gpu_ctx.trace_any(sh), acc.any(sh, brute=True))
