"""BVH8 quality on the GPU builder: the SAH-optimal collapse (LT_COLLAPSE_SAH,
dynamic programming over the binary tree, Ylitie, Karras & Laine 2017)
against the greedy largest-area expansion (LT_COLLAPSE_GREEDY) over the same
binary LBVH, and the PLOC binary tree (LT_TREE_PLOC, the default: clustering
by merged surface area, Meister & Bittner 2018) against the LBVH under the
SAH collapse.

The SAH cost of a BVH8, sum over nodes of A(node) * 1 plus sum over leaf
slots of A(slot) * 0.5 * triangles (A = surface area), is computed here in
numpy from the exported 128-B nodes (fp16 child planes dequantised).  The
programme minimises exactly this cost over every collapse of the binary
tree, and the greedy result is one such collapse, so the SAH build's cost is
at most the greedy one's (up to the quantisation of the child boxes); PLOC's
tree is expected to cost less than the LBVH's.  All trees give the same hits
(the traversal is exact either way); the visits per ray are printed for the
record."""
import numpy as np
import pytest

import lumen_amd as la
from raygen import aimed_rays, random_rays

pytestmark = pytest.mark.gpu


def _sah_cost(nodes: np.ndarray) -> float:
    p = nodes[:, 0:12].view(np.float32).reshape(-1, 3).astype(np.float64)
    e = nodes[:, 12:15].view(np.int8).astype(np.int64)
    imask = nodes[:, 15]
    meta = nodes[:, 24:32]
    qlo = nodes[:, 32:80].view(np.float16).reshape(-1, 3, 8).astype(np.float64)
    qhi = nodes[:, 80:128].view(np.float16).reshape(-1, 3, 8).astype(np.float64)
    scale = np.ldexp(1.0, e)[:, :, None]
    lo = p[:, :, None] + qlo * scale
    hi = p[:, :, None] + qhi * scale
    ext = np.maximum(hi - lo, 0.0)
    area = ext[:, 0] * ext[:, 1] + ext[:, 1] * ext[:, 2] + ext[:, 2] * ext[:, 0]  # [N, 8]
    internal = ((imask[:, None] >> np.arange(8)) & 1).astype(bool)
    count = (meta & 7).astype(np.float64)
    used = internal | (count > 0)
    # a node's own area: the union of its children
    nlo = np.where(used[:, None, :], lo, np.inf).min(axis=2)
    nhi = np.where(used[:, None, :], hi, -np.inf).max(axis=2)
    next_ = np.maximum(nhi - nlo, 0.0)
    narea = next_[:, 0] * next_[:, 1] + next_[:, 1] * next_[:, 2] + next_[:, 2] * next_[:, 0]
    leaf_cost = np.where(~internal, area * 0.5 * count, 0.0).sum()
    return float(narea.sum() * 1.0 + leaf_cost)


@pytest.mark.parametrize("kind,kw", [
    ("caustics", {}),
    ("terrain", dict(terrain_tris=60000, num_instances=300, num_point_lights=8)),
])
@pytest.mark.parametrize("max_leaf", [1, 3])
def test_sah_collapse_is_cheaper_than_greedy(gpu_ctx, kind, kw, max_leaf):
    print(f"[gpu] bvh quality {kind} leaf={max_leaf}", flush=True)
    sc = la.Scene.procedural(kind, seed=3, **kw)
    gpu_ctx.upload(sc)
    tris = sc.triangles()
    rays = np.concatenate([random_rays(tris, 20000, 41), aimed_rays(tris, 20000, 42)])
    res = {}
    for mode, tree in (("greedy", "lbvh"), ("sah", "lbvh"), ("sah", "ploc")):
        st = gpu_ctx.build_accel(max_leaf, collapse=mode, tree=tree)
        mode = mode if tree == "lbvh" else "ploc"
        nodes, _ = gpu_ctx.export_bvh(st)
        hits = gpu_ctx.trace_closest(rays)
        nv, nt = gpu_ctx.trace_count(rays)
        res[mode] = (_sah_cost(nodes), st.num_nodes, nv / len(rays), nt / len(rays), hits)
        print(f"[gpu] {kind} leaf={max_leaf} {mode}: sah cost {res[mode][0]:.6g}, {st.num_nodes} nodes, "
              f"{res[mode][2]:.2f} nodes and {res[mode][3]:.2f} triangles per ray", flush=True)
    assert res["sah"][0] <= res["greedy"][0] * (1 + 2e-3)
    assert res["ploc"][0] <= res["sah"][0]
    g = res["greedy"][4]
    for other in ("sah", "ploc"):
        s = res[other][4]
        for f in ("inst", "prim"):
            assert np.array_equal(g[f], s[f])
        assert np.array_equal(g["t"].view(np.uint32), s["t"].view(np.uint32))
