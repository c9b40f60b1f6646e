"""Seeded ray sets for the traversal parity tests (host numpy)."""
import numpy as np

from lumen_amd import RAY_DTYPE


def random_rays(tris: np.ndarray, n: int, seed: int, tmin=0.0, tmax=1e30, scale=1.2) -> np.ndarray:
    """Origins uniform in the (scaled) scene box, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    c, h = (lo + hi) / 2, (hi - lo) / 2 * scale
    o = c + (rng.random((n, 3)) * 2 - 1) * h
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, RAY_DTYPE)
    r["org"], r["dir"], r["tmin"], r["tmax"] = o, d, tmin, tmax
    return r


def aimed_rays(tris: np.ndarray, n: int, seed: int, tmin=0.0, tmax=1e30, kind="mixed") -> np.ndarray:
    """Rays aimed exactly at triangle vertices, edge midpoints and centroids
    (the grazing / shared-edge / shared-vertex cases of a watertight test)."""
    rng = np.random.default_rng(seed)
    lo, hi = tris.reshape(-1, 3).min(0), tris.reshape(-1, 3).max(0)
    c, h = (lo + hi) / 2, (hi - lo) / 2 * 1.5
    o = (c + (rng.random((n, 3)) * 2 - 1) * h).astype(np.float32)
    ti = rng.integers(0, len(tris), n)
    k = rng.integers(0, 3, n)
    sel = rng.integers(0, 3, n)
    v = tris[ti, k]
    e = (tris[ti, k] + tris[ti, (k + 1) % 3]) * np.float32(0.5)
    m = tris[ti].mean(1)
    tgt = np.where(sel[:, None] == 0, v, np.where(sel[:, None] == 1, e, m)).astype(np.float32)
    d = (tgt - o).astype(np.float32)
    r = np.zeros(n, RAY_DTYPE)
    r["org"], r["dir"], r["tmin"], r["tmax"] = o, d, tmin, tmax
    return r
