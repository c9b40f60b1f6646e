"""The reference's own scenes (tests/golden/scenes: scenes/cornell_box with its
Path, BDPT, VCM and SPPM integrator blocks and its directional-light Path
variant with the wood-textured floor and the atmosphere sky and its variant
with an unknown ("disney") BSDF type, scenes/caustics and its zoomed view,
the occluded-light VCM scenes occluded and occluded2, and the material test
scene material_test/materials with its dielectrics, conductors and principled
materials), loaded through lt_scene_load_json and rendered by the integrator
their JSON names on the GPU, against the oracle on the same scene, camera,
seed and frames.

The scenes come from the reference unchanged: an 8-shape Cornell box of 72
triangles and 1,920 quads (3,912 triangles) with a glass sphere, an
18,893-triangle caustics scene with a spot light, the material test scene of
23,112 triangles.  The bar is the one of the integrators' replays (per-pixel
|gpu - cpu| <= 1e-3 (1 + |cpu|) on >= 99.5 % of the pixels, relative RMSE
<= 1 %).  The VCM scenes also render with the corrected estimator at that
bar.  The zoomed caustic (depth 13, the glass sphere filling the view) is
the exception: its eye and light paths refract through the sphere so many
times that last-bit arithmetic flips discrete choices; the oracle built with
fused multiply-adds differs from itself there on 0.6 % of the pixels with a
relative RMSE of 2 % (tests/test_oracle_sensitivity.py), so the device is
held to >= 99 % of the pixels and a relative RMSE <= 2.5 % on that scene."""
import gzip
from pathlib import Path

import numpy as np
import pytest

import lumen_amd as la
import oracle

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "scenes"


def _compare_frac(g, o, tag, frac=0.995, max_rel=0.01):
    d = np.abs(g[..., :3] - o[..., :3])
    ok = np.all(d <= 1e-3 * (1.0 + np.abs(o[..., :3])), axis=-1)
    rel = np.sqrt(np.mean(d ** 2)) / max(np.sqrt(np.mean(o[..., :3] ** 2)), 1e-6)
    print(f"[gpu] {tag}: pixels matching {ok.mean():.4f} ({int((~ok).sum())} differ), rel rmse {rel:.2e}",
          flush=True)
    assert np.all(np.isfinite(g))
    assert ok.mean() >= frac and rel <= max_rel


def _load(tmp_path, json_name, obj_name):
    (tmp_path / obj_name).write_bytes(gzip.decompress((FIX / (obj_name + ".gz")).read_bytes()))
    (tmp_path / json_name).write_bytes((FIX / json_name).read_bytes())
    (tmp_path / "wood1.jpg").write_bytes((FIX / "wood1.jpg").read_bytes())
    return la.Scene.load_json(str(tmp_path / json_name))


@pytest.mark.parametrize("json_name,obj_name,frames", [
    ("cornell_box_path.json", "cornell_box.obj", 3),
    ("cornell_box_dir.json", "cornell_box.obj", 3),
    ("cornell_box_bdpt.json", "cornell_box.obj", 2),
    ("cornell_box_vcm.json", "cornell_box.obj", 2),
    ("cornell_box_sppm.json", "cornell_box.obj", 3),
    ("caustics.json", "caustics.obj", 2),
    ("cornell_box_disney.json", "cornell_box.obj", 3),
    ("caustics_zoomed.json", "caustics.obj", 2),
    ("occluded.json", "occluded.obj", 2),
    ("occluded2.json", "occluded2.obj", 2),
    ("materials.json", "materials.obj", 3),
])
def test_reference_scene_renders_like_the_oracle(gpu_ctx, tmp_path, json_name, obj_name, frames):
    sc = _load(tmp_path, json_name, obj_name)
    cfg = sc.integrator_cfg()
    name = {0: "path", 1: "bdpt", 2: "vcm", 3: "sppm"}[cfg.kind]
    W, H = 64, 48
    print(f"[gpu] reference scene {json_name}: {name}, {sc.num_triangles} triangles, {W}x{H} x{frames}",
          flush=True)
    it = la.create_integrator(name, sc, gpu_ctx, W, H)
    it.seed = 11
    it.init()
    it.create_accel()
    for _ in range(frames):
        it.render()
        it.update()
    g = it.output_tex
    it.destroy()
    acc = oracle.Accel(sc.desc)
    depth = it.path_length
    if name == "path":
        o, _, _, _ = acc.render_path(it.camera, W, H, depth, sky_col=tuple(it.sky_col), frames=frames, seed=11)
    elif name == "bdpt":
        o, _, _, _ = acc.render_bdpt(it.camera, W, H, depth, frames=frames, seed=11)
    elif name == "vcm":
        o, _, _, _ = acc.render_vcm(it.camera, W, H, depth, frames=frames, seed=11, radius_factor=it.radius_factor,
                                    alpha=it.alpha, use_vm=it.enable_vm, corrected=False,
                                    sky_col=tuple(it.sky_col))
    else:
        o, _, _, _, _ = acc.render_sppm(it.camera, W, H, depth, frames=frames, seed=11, base_radius=it.base_radius,
                                        sky_col=tuple(it.sky_col))
    assert np.abs(o[..., :3]).max() > 0
    bar = dict(frac=0.99, max_rel=0.025) if json_name == "caustics_zoomed.json" else {}
    _compare_frac(g, o, f"reference {json_name}", **bar)
    if name == "vcm":
        it = la.create_integrator(name, sc, gpu_ctx, W, H, estimator="corrected")
        it.seed = 11
        it.init()
        it.create_accel()
        for _ in range(frames):
            it.render()
            it.update()
        g = it.output_tex
        it.destroy()
        o, _, _, _ = acc.render_vcm(it.camera, W, H, depth, frames=frames, seed=11, radius_factor=it.radius_factor,
                                    alpha=it.alpha, use_vm=it.enable_vm, corrected=True, sky_col=tuple(it.sky_col))
        _compare_frac(g, o, f"reference {json_name}, corrected estimator", **bar)
