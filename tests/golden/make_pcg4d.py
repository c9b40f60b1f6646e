"""Writes tests/golden/pcg4d.json: pcg4d (Jarzynski & Olano, "Hash Functions
for GPU Rendering", JCGT 9(3), 2020, listing of pcg4d) evaluated by this
independent numpy restatement of the published listing on seeded inputs and
on the reference's seeding pattern (pixel x, pixel y, frame, counter) of
src/shaders/utils.glsl:138-148.  Run: python tests/golden/make_pcg4d.py
"""
import json
from pathlib import Path

import numpy as np


def pcg4d(v):
    v = np.array(v, dtype=np.uint32).copy()
    with np.errstate(over="ignore"):
        v = v * np.uint32(1664525) + np.uint32(1013904223)
        v[..., 0] += v[..., 1] * v[..., 3]
        v[..., 1] += v[..., 2] * v[..., 0]
        v[..., 2] += v[..., 0] * v[..., 1]
        v[..., 3] += v[..., 1] * v[..., 2]
        v ^= v >> np.uint32(16)
        v[..., 0] += v[..., 1] * v[..., 3]
        v[..., 1] += v[..., 2] * v[..., 0]
        v[..., 2] += v[..., 0] * v[..., 1]
        v[..., 3] += v[..., 1] * v[..., 2]
    return v


def main():
    rng = np.random.default_rng(20200914)
    inputs = np.concatenate([
        np.zeros((1, 4), np.uint32),
        np.array([[0xFFFFFFFF] * 4, [1, 2, 3, 4], [255, 127, 0, 1]], np.uint32),
        rng.integers(0, 2**32, size=(60, 4), dtype=np.uint64).astype(np.uint32),
        np.array([[x, y, f, w] for x in (0, 17, 1919) for y in (0, 1079) for f in (0, 5) for w in (1, 2)], np.uint32),
    ])
    out = pcg4d(inputs)
    # rand(): counter pre-incremented, x lane of pcg4d, 23-bit mantissa (utils.glsl:136,145-148)
    seq = []
    for (x, y, f) in [(0, 0, 0), (3, 7, 1), (1919, 1079, 255)]:
        s = np.array([x, y, f, 0], np.uint32)
        vals = []
        for _ in range(16):
            s[3] += 1
            h = pcg4d(s)[0]
            vals.append(float(np.frombuffer(np.uint32(0x3F800000 | (int(h) >> 9)).tobytes(), np.float32)[0] - 1.0))
        seq.append({"x": x, "y": y, "frame": f, "rand": vals})
    Path(__file__).with_name("pcg4d.json").write_text(json.dumps(
        {"inputs": inputs.tolist(), "outputs": out.tolist(), "rand_sequences": seq}, indent=0))


if __name__ == "__main__":
    main()
