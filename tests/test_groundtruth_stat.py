"""The block z statistic of test_gpu_groundtruth, checked on the CPU with
the oracle on both sides: two independent seed sets pass it, and the same
images with a planted 5 % bias (one side scaled by 1.05) or a planted bias in
one region (16x16 pixels, four of the 64 blocks, scaled by 1.5) fail it."""
import numpy as np

import lumen_amd as la
import oracle
from test_gpu_groundtruth import _accept, _z_blocks


def _images(acc, cam, W, H, seeds):
    out = np.zeros((len(seeds), H, W), np.float64)
    for k, s in enumerate(seeds):
        o, _, _, _ = acc.render_path(cam, W, H, 5, frames=1, seed=s)
        out[k] = o[..., :3].sum(-1)
    return out


def test_block_z_statistic_accepts_equal_and_rejects_biased_estimates():
    W, H = 64, 64
    sc = la.Scene.procedural("cornell", seed=1)
    cam = sc.camera(W, H)
    acc = oracle.Accel(sc.desc)
    a = _images(acc, cam, W, H, range(100, 196))
    b = _images(acc, cam, W, H, range(900, 964))
    z = _z_blocks(a, b)
    assert len(z) >= 48
    assert _accept(z)
    assert not _accept(_z_blocks(a * 1.05, b))
    c = a.copy()
    c[:, 24:40, 24:40] *= 1.5
    assert not _accept(_z_blocks(c, b))
