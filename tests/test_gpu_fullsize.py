"""Parity at the workloads' real sizes (GPU against the oracle).

- Terrain, 1920x1080: 1,023,700 triangles (fBm heightfield + 2,000 boxes),
  Disney materials, 64 point lights + 1 area light, Path to depth 12 with
  Russian roulette after depth 3.  Closest and any hits on 20,000 rays
  bit-exact against the host traversal of the exported BVH8 and against the
  host BVH2; one persistent-kernel frame against the oracle on a band of 64
  rows; the megakernel and wavefront forms against the persistent one on the
  whole frame.
- Caustics, 1024x1024 (20k-triangle glass sphere, spot light): BDPT with the
  power heuristic, depth 8, one frame; VCM with 1,048,576 light subpaths,
  merge radius 1% of the bounding radius (0.5% of the diagonal), alpha 0.75,
  depth 8, both estimators, one frame, whose merge queries include ones with
  far more than 64 photons inside the radius (the merge kernel's carry path).
- Differentiable Cornell box: 512x512, back wall of 16,384 triangles with
  256x256 albedo and roughness textures, depth 6, variants 0 and 1 of the 32,
  images and texel gradients against the oracle; gradients against float64
  central differences (h = 1e-3) of the oracle's estimator on 64 texels.

Tolerances (per-pixel replay of the same pcg4d streams): device
transcendentals differ from libm in the last ulp and, rarely, flip a
discrete decision (a lobe choice, a roulette test, an edge hit), after which
that path differs entirely; float atomics reorder splat and merge sums.
Each image test states the fraction of pixels within 1e-3 (1 + |oracle|)
it requires and a bound on the relative RMSE of the whole compared region.
"""
import numpy as np
import pytest

import lumen_amd as la
import oracle
from raygen import random_rays

pytestmark = pytest.mark.gpu


def _stats(g, o):
    d = np.abs(g[..., :3].astype(np.float64) - o[..., :3])
    ok = np.all(d <= 1e-3 * (1.0 + np.abs(o[..., :3])), axis=-1)
    rel = np.sqrt(np.mean(d ** 2)) / max(np.sqrt(np.mean(o[..., :3].astype(np.float64) ** 2)), 1e-12)
    return ok.mean(), rel, int((~ok).sum())


def _check_image(g, o, tag, min_match, max_rel):
    assert np.all(np.isfinite(g)), tag
    frac, rel, bad = _stats(g, o)
    print(f"[gpu] {tag}: pixels matching {frac:.5f} ({bad} differ), rel rmse {rel:.2e}", flush=True)
    assert frac >= min_match, tag
    assert rel <= max_rel, tag


# ------------------------------------------------------------------ terrain
@pytest.fixture(scope="module")
def terrain():
    print("[gpu] terrain scene (1M triangles + 2,000 boxes)", flush=True)
    sc = la.Scene.procedural("terrain", seed=1)
    assert sc.num_triangles == 1_023_700
    print("[gpu] host BVH of the terrain", flush=True)
    return sc, oracle.Accel(sc.desc)


def _hit_mismatch(g, o):
    return np.nonzero((g["inst"] != o["inst"]) | (g["prim"] != o["prim"]) |
                      (g["t"].view(np.uint32) != o["t"].view(np.uint32)) |
                      (g["u"].view(np.uint32) != o["u"].view(np.uint32)) |
                      (g["v"].view(np.uint32) != o["v"].view(np.uint32)))[0]


def test_terrain_hits_bit_exact(gpu_ctx, terrain):
    sc, acc = terrain
    gpu_ctx.upload(sc)
    st = gpu_ctx.build_accel()
    print(f"[gpu] terrain bvh8: {st.num_nodes} nodes, depth {st.max_depth}", flush=True)
    nodes, tris = gpu_ctx.export_bvh(st)
    # camera rays through random pixels of the 1080p view, and rays between random points of the scene box
    W, H = 1920, 1080
    it = la.Path(sc, gpu_ctx, W, H)
    cam = it.camera
    rng = np.random.default_rng(17)
    px = rng.random((10000, 2)) * [W, H]
    ndc = px / [W, H] * 2 - 1
    inv_view = np.array(cam.inv_view, np.float64).reshape(4, 4).T
    inv_proj = np.array(cam.inv_projection, np.float64).reshape(4, 4).T
    tg = (inv_proj @ np.c_[ndc, np.ones((len(ndc), 2))].T).T[:, :3]
    tg /= np.linalg.norm(tg, axis=1, keepdims=True)
    d = (inv_view[:3, :3] @ tg.T).T
    cam_rays = np.zeros(len(d), la.RAY_DTYPE)
    cam_rays["org"], cam_rays["dir"], cam_rays["tmin"], cam_rays["tmax"] = inv_view[:3, 3], d, 0.001, 1e4
    rays = np.concatenate([cam_rays, random_rays(sc.triangles(), 10000, 18, tmin=0.001, tmax=1e4)])
    g = gpu_ctx.trace_closest(rays)
    assert (g["inst"] != 0xFFFFFFFF).mean() > 0.4
    h8, _, _ = oracle.bvh8_trace(nodes, tris, rays)
    bad = _hit_mismatch(g, h8)
    assert len(bad) == 0, f"{len(bad)} rays differ from the host BVH8 traversal"
    h2 = acc.closest(rays)
    bad = _hit_mismatch(g, h2)
    assert len(bad) == 0, f"{len(bad)} rays differ from the host BVH2"
    sh = rays.copy()
    sh["tmin"] = 0.0
    sh["tmax"] = rng.random(len(sh)).astype(np.float32) * 20.0
    ga = gpu_ctx.trace_any(sh)
    assert np.array_equal(ga, acc.any(sh))
    print(f"[gpu] terrain: {len(rays)} closest and any hits bit-exact", flush=True)


def test_terrain_path_frame_band(gpu_ctx, terrain):
    """One 1-spp persistent-kernel frame (frame 0, depth 12, roulette after 3)
    against the oracle on rows 508..571 (122,880 pixels)."""
    sc, acc = terrain
    W, H = 1920, 1080
    it = la.Path(sc, gpu_ctx, W, H, strategy="persistent")
    assert it.path_length == 12
    it.init()
    it.create_accel()
    it.render(spp=1)
    g = it.output_tex
    it.destroy()
    y0, y1 = 508, 572
    print("[gpu] terrain path: oracle band", flush=True)
    o, secs, _, _ = acc.render_path(it.camera, W, H, it.path_length, sky_col=tuple(it.sky_col), frames=1,
                                    window=(0, y0, W, y1))
    _check_image(g[y0:y1], o[y0:y1], "terrain path 1080p band", 0.999, 0.01)


def test_terrain_execution_forms_full_frame(gpu_ctx, terrain):
    """The megakernel and the wavefront form run the persistent kernel's
    estimator: same per-path streams and decisions, different summation
    order only, on every pixel of two full frames."""
    sc, _ = terrain
    imgs = {}
    for strat in ("persistent", "megakernel", "wavefront"):
        it = la.Path(sc, gpu_ctx, 1920, 1080, strategy=strat)
        it.init()
        it.create_accel()
        it.render(spp=2)
        imgs[strat] = it.output_tex
        it.destroy()
    a = imgs["persistent"]
    for strat in ("megakernel", "wavefront"):
        b = imgs[strat]
        assert np.all(np.isfinite(b))
        d = np.abs(a[..., :3] - b[..., :3])
        bad = ~np.all(d <= 1e-4 * (1.0 + np.abs(a[..., :3])), axis=-1)
        print(f"[gpu] terrain {strat} vs persistent: {int(bad.sum())} of {bad.size} pixels differ", flush=True)
        assert bad.mean() <= 1e-5, strat


# ----------------------------------------------------------------- caustics
@pytest.fixture(scope="module")
def caustics():
    sc = la.Scene.procedural("caustics", seed=7)
    return sc, oracle.Accel(sc.desc)


def test_caustics_bdpt_power_full_frame(gpu_ctx, caustics):
    sc, acc = caustics
    W = H = 1024
    it = la.BDPT(sc, gpu_ctx, W, H, path_length=8, mis="power")
    it.seed = 7
    it.init()
    it.create_accel()
    it.render(spp=1)
    g = it.output_tex
    it.destroy()
    print("[gpu] caustics bdpt: oracle frame", flush=True)
    o, _, _, _ = acc.render_bdpt(it.camera, W, H, 8, frames=1, seed=7, power=True)
    _check_image(g, o, "caustics bdpt power 1024^2", 0.999, 0.02)


@pytest.mark.parametrize("estimator", ["corrected", "reference"])
def test_caustics_vcm_full_frame(gpu_ctx, caustics, estimator):
    sc, acc = caustics
    W = H = 1024
    it = la.VCM(sc, gpu_ctx, W, H, path_length=8, radius_factor=1.0, alpha=0.75, enable_vm=True,
                estimator=estimator)
    it.seed = 7
    it.init()
    it.create_accel()
    it.render(spp=1)
    g = it.output_tex
    it.destroy()
    print(f"[gpu] caustics vcm {estimator}: oracle frame", flush=True)
    o, _, _, _ = acc.render_vcm(it.camera, W, H, 8, frames=1, seed=7, radius_factor=1.0, alpha=0.75, use_vm=True,
                                corrected=(estimator == "corrected"), sky_col=tuple(it.sky_col))
    if estimator == "corrected":
        q, over, mx = oracle.vcm_merge_stats()
        print(f"[gpu] merge queries {q}, with > 64 photons in the radius {over}, at most {mx}", flush=True)
        assert over > 1000 and mx > 10000  # the dense caustic: the merge kernel's queue carries over
    _check_image(g, o, f"caustics vcm {estimator} 1024^2", 0.999, 0.01)


# ------------------------------------------------------------------ adjoint
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def diff_wall():
    sc = la.Scene.procedural("diff_wall", seed=5)  # 16,384-triangle wall, 256x256 textures
    d = sc.desc
    tex = [np.ctypeslib.as_array(d.textures[t].data, shape=(d.textures[t].height, d.textures[t].width, 4)).copy()
           for t in range(d.num_textures)]
    alb, rough = tex[0][..., :3].astype(np.float32), tex[1][..., 0].astype(np.float32)
    assert alb.shape == (256, 256, 3) and rough.shape == (256, 256)
    rng = np.random.default_rng(1)
    V = 32  # the bench's variants; the test takes the first two
    albs = np.clip(alb[None] + rng.uniform(-0.2, 0.2, (V,) + alb.shape), 0.05, 0.95).astype(np.float32)
    roughs = np.clip(rough[None] + rng.uniform(-0.2, 0.2, (V,) + rough.shape), 0.1, 0.95).astype(np.float32)
    return sc, oracle.Accel(sc.desc), albs[:2].copy(), roughs[:2].copy()


def test_adjoint_c4_two_variants(gpu_ctx, diff_wall):
    from lumen_amd.adjoint import Adjoint
    sc, acc, albs, roughs = diff_wall
    W = H = 512
    V, spp, depth = 2, 2, 6
    adj = Adjoint(sc, gpu_ctx, W, H, spp=spp, variants=V, albedo_texture=0, roughness_texture=1, max_depth=depth)
    w = np.random.default_rng(3).normal(size=(V, H, W, 3)).astype(np.float32)
    ta, tr = torch.from_numpy(albs).cuda(), torch.from_numpy(roughs).cuda()
    img, cache = adj.forward(ta, tr, seed=7)
    ga, gr = adj.backward(ta, tr, torch.from_numpy(w).cuda(), seed=7, cache=cache)
    torch.cuda.synchronize()
    g_img, ga, gr = img.cpu().numpy(), ga.cpu().numpy(), gr.cpu().numpy()
    adj.close()
    print("[gpu] adjoint c4: oracle forward and backward", flush=True)
    kw = dict(albedo=albs, rough=roughs, albedo_id=0, rough_id=1, seed=7)
    o_img = acc.adjoint(adj.camera, W, H, depth, spp, **kw)
    o_ga, o_gr = acc.adjoint(adj.camera, W, H, depth, spp, d_images=w, **kw)
    for v in range(V):
        _check_image(g_img[v], o_img[v], f"adjoint c4 forward v{v}", 0.999, 0.005)
    for g, o, tag in ((ga, o_ga, "albedo"), (gr, o_gr, "roughness")):
        rel = np.linalg.norm(g - o) / np.linalg.norm(o)
        ok = np.abs(g - o) <= 1e-3 * np.abs(o).max() + 1e-2 * np.abs(o)
        print(f"[gpu] adjoint c4 {tag} gradient: rel L2 {rel:.2e}, texels matching {ok.mean():.5f}", flush=True)
        assert rel <= 1e-3 and ok.mean() >= 0.999, tag


def test_adjoint_gradients_match_float64_central_differences(gpu_ctx, diff_wall):
    """GPU texel gradients (C4 scene and textures, 96x96 image, 4 spp) against
    central differences with h = 1e-3 of the oracle's estimator in float64, on
    32 albedo and 32 roughness texels (roughness with detached sampling, the
    estimator the adjoint differentiates): within 1e-3 relative, with a floor
    of 1e-6 of the largest gradient for the float32 atomics."""
    from lumen_amd.adjoint import Adjoint
    sc, _, albs, roughs = diff_wall
    a64 = oracle.Accel64(sc.desc)
    W = H = 96
    spp, depth, h = 4, 6, 1e-3
    alb, rough = albs[:1], roughs[:1]
    rng = np.random.default_rng(8)
    w = rng.uniform(0.5, 1.5, size=(1, H, W, 3)).astype(np.float32)
    adj = Adjoint(sc, gpu_ctx, W, H, spp=spp, variants=1, albedo_texture=0, roughness_texture=1, max_depth=depth)
    ta, tr = torch.from_numpy(alb).cuda(), torch.from_numpy(rough).cuda()
    img, cache = adj.forward(ta, tr, seed=5)
    ga, gr = adj.backward(ta, tr, torch.from_numpy(w).cuda(), seed=5, cache=cache)
    torch.cuda.synchronize()
    ga, gr = ga.cpu().numpy().astype(np.float64), gr.cpu().numpy().astype(np.float64)
    cam = adj.camera
    adj.close()
    kw = dict(rough_id=1, albedo_id=0, seed=5)
    worst = {}
    for grad, name in ((ga, "albedo"), (gr, "roughness")):
        flat = np.flatnonzero(np.abs(grad) > 1e-4 * np.abs(grad).max())
        picks = rng.choice(flat, size=32, replace=False)
        errs = []
        for f in picks:
            idx = np.unravel_index(f, grad.shape)
            if name == "albedo":
                ap, am = alb.astype(np.float64), alb.astype(np.float64)
                ap[idx] += h
                am[idx] -= h
                ip = a64.adjoint(cam, W, H, depth, spp, albedo=ap, rough=rough, **kw)
                im = a64.adjoint(cam, W, H, depth, spp, albedo=am, rough=rough, **kw)
            else:
                rp, rm = rough.astype(np.float64), rough.astype(np.float64)
                rp[idx] += h
                rm[idx] -= h
                ip = a64.adjoint(cam, W, H, depth, spp, albedo=alb, rough=rough, eval_rough=rp, **kw)
                im = a64.adjoint(cam, W, H, depth, spp, albedo=alb, rough=rough, eval_rough=rm, **kw)
            fd = float(np.sum(w * (ip - im)) / (2 * h))
            err = abs(grad[idx] - fd)
            errs.append(err / abs(fd))
            assert err <= 1e-3 * abs(fd) + 1e-6 * np.abs(grad).max(), (name, idx, grad[idx], fd)
        worst[name] = max(errs)
    print(f"[gpu] adjoint vs float64 central differences: max rel error {worst}", flush=True)
