"""Autotuner of the launch shapes and of the traversal's build-time sizes.

Two kinds of parameters, both measured on the device with the kernels' own
HIP-event timing of whole frames:

- run-time fields of lt_tuning (workgroups per CU of the persistent kernels,
  the idle-lane thresholds of path regeneration and of the wavefront trace
  stages, the BDPT pixel-chunk size, the wavefront material sort): searched
  by coordinate descent, one field at a time over its candidate values with
  the others held at their current best, for a few rounds (`tune`);
- build-time sizes (the LDS traversal stack, the per-lane LDS leaf queue and
  how many queued leaves still let a lane take its next node): each
  combination is compiled into a variant library and timed in a child
  process (`tune_build`).

Results are stored in a JSON table (lumen_amd/tuning.json) keyed by GPU
architecture, workload and scene class; `apply` sets a stored choice on a
context.  Run as

    python -m lumen_amd.autotune --workload path-persistent --scene terrain [--store]
    python -m lumen_amd.autotune --build --workload path-persistent --scene terrain
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

TABLE = Path(__file__).resolve().parent / "tuning.json"

# candidate values per lt_tuning field and workload
SPACE = {
    "path-persistent": {"path_blocks_per_cu": [6, 8, 10, 12, 16], "regen_min": [1, 2, 4, 8, 16, 32],
                        "xcd_bands": [1, 2]},
    "path-megakernel": {"path_blocks_per_cu": [4, 6, 8, 10, 12, 16]},
    "path-wavefront": {"path_blocks_per_cu": [6, 8, 10, 12, 16], "refill_min": [8, 16, 32, 48, 64],
                       "wf_sort": [1, 2]},
    "bdpt": {"bdpt_blocks_per_cu": [2, 3, 4, 6, 8], "regen_min": [1, 2, 4, 8, 16]},
    "vcm-corrected": {"bdpt_blocks_per_cu": [2, 3, 4, 6, 8]},
    "sppm": {"bdpt_blocks_per_cu": [2, 3, 4, 6, 8]},
}

# build-time sizes (preprocessor definitions of bvh_layout.h / dev_trace.h)
BUILD_SPACE = {"LT_LDS_STACK": [12, 16, 24], "LT_LEAF_QUEUE": [16, 24], "LT_NODE_PEND_MAX": [4, 8, 16]}


def _time_frames(integrator, frames: int) -> float:
    integrator.render()
    integrator.update()  # warm-up at this shape
    ms = []
    for _ in range(frames):
        integrator.render()
        ms.append(integrator.stats().kernel_ms)
        integrator.update()
    return sum(ms) / len(ms)


def tune_field(integrator, field: str, candidates, frames: int = 4) -> dict:
    """Times `frames` 1-spp frames per value of one lt_tuning field (the
    candidates and the current value); returns {"best": value, "ms": {value:
    avg ms}} and sets the best."""
    ctx = integrator.ctx
    results = {}
    for c in sorted(set(candidates) | {ctx.get_tuning()[field]}):
        ctx.set_tuning(**{field: c})
        results[c] = _time_frames(integrator, frames)
    best = min(results, key=results.get)
    ctx.set_tuning(**{field: best})
    return {"best": best, "ms": results}


def tune(integrator, space: dict, frames: int = 4, rounds: int = 2) -> dict:
    """Coordinate descent over the fields of `space` (field -> candidates):
    each round sweeps every field in turn with the others at their current
    best.  Returns {"best": {field: value}, "ms": best avg ms, "trace": [...]}
    and leaves the best setting on the context."""
    ctx = integrator.ctx
    cur = {f: ctx.get_tuning()[f] for f in space}
    trace = []
    best_ms = None
    for _ in range(rounds):
        changed = False
        for field, cands in space.items():
            r = tune_field(integrator, field, cands, frames)
            trace.append({"field": field, "ms": r["ms"]})
            if r["best"] != cur[field]:
                changed = True
            cur[field] = r["best"]
            best_ms = r["ms"][r["best"]]
        if not changed:
            break
    ctx.set_tuning(**cur)
    return {"best": cur, "ms": best_ms, "trace": trace}


def tune_build(variants, time_cmd, build=None) -> dict:
    """Builds one library per combination of preprocessor definitions in
    `variants` (a list of {name: value} dicts) and runs `time_cmd` (an argv
    list that prints the frame time in ms as its last output line) with
    LT_LIBRARY pointing at it.  Returns {definitions: ms}."""
    if build is None:
        from .build import build
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for v in variants:
            defs = tuple(f"{k}={val}" for k, val in sorted(v.items()))
            lib = Path(d) / ("liblt_" + "_".join(x.replace("=", "") for x in defs) + ".so")
            build(defines=defs, lib_path=lib)
            env = dict(os.environ, LT_LIBRARY=str(lib))
            r = subprocess.run(time_cmd, capture_output=True, text=True, env=env)
            lines = [x for x in r.stdout.strip().splitlines() if x.strip()]
            out[" ".join(defs)] = float(lines[-1]) if r.returncode == 0 and lines else None
    return out


def build_variants(space: dict = BUILD_SPACE) -> list:
    """Every combination of the build-time sizes that fits: a node step may
    queue 8 leaves, so LT_NODE_PEND_MAX + 8 <= LT_LEAF_QUEUE."""
    keys = sorted(space)
    out = []
    for vals in itertools.product(*(space[k] for k in keys)):
        v = dict(zip(keys, vals))
        if v.get("LT_NODE_PEND_MAX", 0) + 8 <= v.get("LT_LEAF_QUEUE", 16):
            out.append(v)
    return out


def load(arch: str, workload: str, scene_class: str, table: Path = TABLE):
    try:
        return json.loads(table.read_text())[arch][workload][scene_class]
    except Exception:
        return None


def store(arch: str, workload: str, scene_class: str, value, table: Path = TABLE) -> None:
    d = json.loads(table.read_text()) if table.exists() else {}
    d.setdefault(arch, {}).setdefault(workload, {})[scene_class] = value
    table.write_text(json.dumps(d, indent=1, sort_keys=True) + "\n")


def apply(ctx, workload: str, scene_class: str, table: Path = TABLE) -> dict | None:
    """Sets the stored run-time choice of (this device's arch, workload,
    scene class) on ctx; returns it, or None if the table has none."""
    _, arch = ctx.device_info()
    entry = load(arch.split(":")[0], workload, scene_class, table)
    if not entry or "runtime" not in entry:
        return None
    ctx.set_tuning(**entry["runtime"])
    return entry["runtime"]


def _make(workload: str, scene: str, ctx, width: int, height: int):
    import lumen_amd as la

    sc = la.Scene.procedural(scene, seed=7 if scene == "caustics" else 1)
    if workload.startswith("path-"):
        it = la.Path(sc, ctx, width, height, strategy=workload.split("-", 1)[1])
    elif workload == "bdpt":
        it = la.BDPT(sc, ctx, width, height, path_length=8, mis="power")
    elif workload == "vcm-corrected":
        it = la.VCM(sc, ctx, width, height, path_length=8, radius_factor=1.0, alpha=0.75, enable_vm=True,
                    estimator="corrected")
    elif workload == "sppm":
        it = la.SPPM(sc, ctx, width, height, path_length=8)
    else:
        raise ValueError(workload)
    it.init()
    it.create_accel()
    return it


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="path-persistent", choices=sorted(SPACE))
    ap.add_argument("--scene", default="terrain")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=4)
    ap.add_argument("--build", action="store_true", help="sweep the build-time sizes instead")
    ap.add_argument("--time-only", action="store_true", help="print the frame time of the current library")
    ap.add_argument("--store", action="store_true")
    args = ap.parse_args()
    import lumen_amd as la

    if args.build:
        cmd = [sys.executable, "-m", "lumen_amd.autotune", "--time-only", "--workload", args.workload, "--scene",
               args.scene, "--width", str(args.width), "--height", str(args.height), "--frames", str(args.frames)]
        res = tune_build(build_variants(), cmd)
        print(json.dumps({"build": res}), flush=True)
        if args.store:
            ctx = la.Context(0)
            arch = ctx.device_info()[1].split(":")[0]
            ctx.close()
            ok = {k: v for k, v in res.items() if v is not None}
            entry = load(arch, args.workload, args.scene) or {}
            entry["build"] = {"ms": res, "best": min(ok, key=ok.get) if ok else None}
            store(arch, args.workload, args.scene, entry)
        return 0
    ctx = la.Context(0)
    it = _make(args.workload, args.scene, ctx, args.width, args.height)
    if args.time_only:
        print(f"{_time_frames(it, args.frames):.5f}", flush=True)
        return 0
    res = tune(it, SPACE[args.workload], frames=args.frames)
    print(json.dumps(res), flush=True)
    if args.store:
        arch = ctx.device_info()[1].split(":")[0]
        entry = load(arch, args.workload, args.scene) or {}
        entry["runtime"] = res["best"]
        entry["runtime_ms"] = res["ms"]
        store(arch, args.workload, args.scene, entry)
    it.destroy()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
