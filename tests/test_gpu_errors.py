"""Error paths of the C ABI on the device, forced from the host: each refusal
returns its status with a message (lt_last_error), and the context goes on
working afterwards (a valid scene, BVH and frame after the refusals, the
frame against the oracle).  Nothing here makes a kernel fail: every refusal
happens in the host checks before a launch."""
import ctypes as C

import numpy as np
import pytest

import lumen_amd as la
import oracle
from lumen_amd import LtError, _abi, lib

pytestmark = pytest.mark.gpu


def test_refusals_leave_the_context_usable(gpu_ctx):
    print("[gpu] error paths", flush=True)
    sc = la.Scene.procedural("cornell", seed=1)
    it = la.Path(sc, gpu_ctx, 24, 16)
    it.init()  # uploads the scene, which drops any BVH
    with pytest.raises(LtError, match="LT_ERR_NOT_READY"):
        it.render()
    assert b"accel" in lib().lt_last_error(gpu_ctx._h)
    # build options out of range
    for field, value in (("max_leaf_tris", 9), ("collapse", 9), ("binary_tree", 9)):
        o = _abi.lt_build_opts()
        setattr(o, field, value)
        s = _abi.lt_build_stats()
        assert lib().lt_accel_build(gpu_ctx._h, C.byref(o), C.byref(s)) == 1, field
    # launch shapes out of range
    for kw in (dict(wf_sort=7), dict(xcd_bands=5), dict(regen_min=65), dict(path_blocks_per_cu=65)):
        with pytest.raises(LtError, match="LT_ERR_INVALID_ARG"):
            gpu_ctx.set_tuning(**kw)
    # integrator configurations out of range
    for make, code in ((lambda: la.Path(sc, gpu_ctx, 0, 16), "LT_ERR_INVALID_ARG"),
                       (lambda: la.BDPT(sc, gpu_ctx, 16, 16, path_length=40), "LT_ERR_UNSUPPORTED"),
                       (lambda: la.VCM(sc, gpu_ctx, 16, 16, radius_factor=-1.0), "LT_ERR_INVALID_ARG"),
                       (lambda: la.SPPM(sc, gpu_ctx, 16, 16, base_radius=float("nan")), "LT_ERR_INVALID_ARG")):
        with pytest.raises(LtError, match=code):  # the camera (image size) or the integrator refuses
            make().init()
    # a scene with a non-finite vertex is refused before anything is freed
    from scenes import DescScene, quad
    s = DescScene()
    m = s.material(bsdf_type=1, bsdf_props=9, albedo=(0.5, 0.5, 0.5))
    s.mesh(quad([-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]), m)
    d = s.build()
    d.vertices[1].pos[2] = float("nan")
    poisoned = la.Scene.__new__(la.Scene)
    poisoned.desc = d
    with pytest.raises(LtError, match="LT_ERR_INVALID_ARG"):
        gpu_ctx.upload(poisoned)
    # the Cornell box uploaded by init() is still the context's scene
    gpu_ctx.build_accel()
    vcm = la.VCM(sc, gpu_ctx, 16, 16)
    vcm.init()
    vcm.create_accel()
    with pytest.raises(LtError, match="LT_ERR_UNSUPPORTED"):
        vcm.set_graph(True)
    vcm.destroy()
    # after all of this a frame renders and replays the oracle
    it.seed = 5
    it.create_accel()
    it.render()
    g = it.output_tex
    o, _, _, _ = oracle.Accel(sc.desc).render_path(it.camera, 24, 16, it.path_length, sky_col=tuple(it.sky_col),
                                                    frames=1, seed=5)
    d = np.abs(g[..., :3] - o[..., :3])
    ok = np.all(d <= 1e-3 * (1.0 + np.abs(o[..., :3])), axis=-1)
    print(f"[gpu] frame after the refusals: pixels matching {ok.mean():.4f}", flush=True)
    assert np.all(np.isfinite(g)) and ok.mean() >= 0.995
    it.destroy()
