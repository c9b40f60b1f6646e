"""Launch-shape autotuner for the persistent kernels.

Sweeps the workgroups-per-CU of the Path kernel (and of the raw queries) on a
given scene, times a few frames per candidate with the kernels' own HIP-event
timing, and returns / stores the best value keyed by (GPU arch, kernel, scene
class) in a JSON table (lumen_amd/tuning.json by default).
"""
from __future__ import annotations

import json
from pathlib import Path

TABLE = Path(__file__).resolve().parent / "tuning.json"


def tune_path(integrator, candidates=(2, 4, 6, 8, 10, 12, 16), frames: int = 4) -> dict:
    """Times `frames` 1-spp frames per candidate; returns {candidate: avg ms} and sets the best."""
    ctx = integrator.ctx
    results = {}
    for c in candidates:
        ctx.set_tuning(path_blocks_per_cu=c)
        integrator.render()
        integrator.update()  # warm-up at this shape
        ms = []
        for _ in range(frames):
            integrator.render()
            ms.append(integrator.stats().kernel_ms)
            integrator.update()
        results[c] = sum(ms) / len(ms)
    best = min(results, key=results.get)
    ctx.set_tuning(path_blocks_per_cu=best)
    return {"best": best, "ms": results}


def tune_field(integrator, field: str, candidates, frames: int = 4) -> dict:
    """Times `frames` 1-spp frames per value of one lt_tuning field
    (e.g. regen_min of the persistent Path kernel, refill_min of the
    wavefront trace stages); returns {value: avg ms} and sets the best."""
    ctx = integrator.ctx
    results = {}
    for c in candidates:
        ctx.set_tuning(**{field: c})
        integrator.render()
        integrator.update()
        ms = []
        for _ in range(frames):
            integrator.render()
            ms.append(integrator.stats().kernel_ms)
            integrator.update()
        results[c] = sum(ms) / len(ms)
    best = min(results, key=results.get)
    ctx.set_tuning(**{field: best})
    return {"best": best, "ms": results}


def load(arch: str, kernel: str, scene_class: str, table: Path = TABLE):
    try:
        return json.loads(table.read_text())[arch][kernel][scene_class]
    except Exception:
        return None


def store(arch: str, kernel: str, scene_class: str, value, table: Path = TABLE) -> None:
    d = json.loads(table.read_text()) if table.exists() else {}
    d.setdefault(arch, {}).setdefault(kernel, {})[scene_class] = value
    table.write_text(json.dumps(d, indent=1, sort_keys=True) + "\n")
