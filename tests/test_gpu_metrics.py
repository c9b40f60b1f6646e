"""lt_image_sq_error (the GPU reduction behind both RMSE definitions) against
numpy.  The reference's RMSE (rmse/calc_rmse.comp:20-41, reduce_rmse.comp,
output_rmse.comp:23) is sqrt(S) / (3 W H) with S = sum |gt.rgb - out.rgb|^2;
the standard RMSE is sqrt(S / (3 W H))."""
import ctypes as C

import numpy as np
import pytest

import lumen_amd as la

pytestmark = pytest.mark.gpu


def _numpy_S(a, b):
    d = a[..., :3] - b[..., :3]  # float32 differences and squares, as the kernel forms them
    return float(np.sum((d * d).astype(np.float64)))


@pytest.mark.parametrize("H,W", [(1, 1), (37, 53), (1080, 1920)])
def test_sq_error_matches_numpy(gpu_ctx, H, W):
    print(f"[gpu] sq_error {W}x{H}", flush=True)
    rng = np.random.default_rng(H + W)
    a = rng.random((H, W, 4)).astype(np.float32) * 4
    b = rng.random((H, W, 4)).astype(np.float32) * 4
    a[..., 3] = 100.0  # alpha is not part of the error
    S = gpu_ctx.sq_error(a, b)
    assert S == pytest.approx(_numpy_S(a, b), rel=1e-9)
    # device pointers give the same sum
    da, db = gpu_ctx.alloc(a.nbytes), gpu_ctx.alloc(b.nbytes)
    gpu_ctx.h2d(da, a)
    gpu_ctx.h2d(db, b)
    out = C.c_double()
    la.check(la.lib().lt_image_sq_error(gpu_ctx.handle, da, db, H * W, la.LT_PTR_DEVICE, C.byref(out)),
             gpu_ctx.handle, "lt_image_sq_error")
    assert out.value == pytest.approx(S, rel=1e-12)
    gpu_ctx.free(da)
    gpu_ctx.free(db)


def test_rmse_definitions_closed_form(gpu_ctx):
    """An image off by 0.5 in every colour channel: S = 0.75 N, Lumen's RMSE
    sqrt(0.75 N) / (3 N), standard RMSE 0.5."""
    H, W = 64, 48
    b = np.random.default_rng(1).random((H, W, 4)).astype(np.float32)
    a = b.copy()
    a[..., :3] += np.float32(0.5)
    lum, std = gpu_ctx.rmse(a, b)
    n = H * W
    assert std == pytest.approx(0.5, rel=1e-6)
    assert lum == pytest.approx(np.sqrt(0.75 * n) / (3 * n), rel=1e-6)
    assert gpu_ctx.sq_error(b, b) == 0.0
    assert gpu_ctx.sq_error(np.zeros((0, 4), np.float32), np.zeros((0, 4), np.float32)) == 0.0
