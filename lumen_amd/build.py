"""Ahead-of-time build of the HIP core into lumen_amd/liblt.so (gfx950).

Every translation unit under lumen_amd/csrc is compiled by hipcc in parallel
and linked into one shared library; the C ABI is include/lt.h.  Run as
``python -m lumen_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "liblt.so"

ARCH = os.environ.get("LT_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"-I{ROOT / 'include'}",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-result",
]
# Packed FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build lumen_amd)")


_REMARK = re.compile(r"remark:\s+(Function Name|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]"
                     r"|LDS Size \[bytes/block\]):\s*(\S+)")
_KEYS = {"VGPRs": "vgprs", "AGPRs": "agprs", "ScratchSize [bytes/lane]": "scratch_bytes",
         "Occupancy [waves/SIMD]": "waves_per_simd", "LDS Size [bytes/block]": "lds_bytes"}


def _resources(remarks: str) -> dict:
    """Per-kernel register, scratch, LDS and occupancy figures from the
    compiler's kernel-resource-usage remarks."""
    out, cur = {}, None
    for m in _REMARK.finditer(remarks):
        key, val = m.group(1), m.group(2)
        if key == "Function Name":
            cur = out.setdefault(val, {})
        elif cur is not None:
            cur[_KEYS[key]] = int(val)
    return out


def _compile(src: Path, obj: Path = OBJ, defines=()):
    out = obj / (src.name + ".o")
    cmd = [_hipcc(), *COMMON, *[f"-D{d}" for d in defines]]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *DEVICE, "-Rpass-analysis=kernel-resource-usage"]
    else:
        # host-only C++ that calls the HIP runtime API
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    cmd += ["-c", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out, _resources(r.stderr) if src.suffix == ".hip" else {}


def build(verbose: bool = False, defines=(), lib_path: Path | None = None) -> Path:
    """Compiles every source and links lumen_amd/liblt.so; returns its path.
    defines (e.g. "LT_LDS_STACK=24") and lib_path build a variant library
    elsewhere (the autotuner's build-time parameters); the default build
    also links the headless CLI."""
    variant = lib_path is not None
    obj = OBJ / ("variant_" + "_".join(d.replace("=", "") for d in defines)) if variant else OBJ
    obj.mkdir(parents=True, exist_ok=True)
    srcs = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])
    jobs = min(len(srcs), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda p: _compile(p, obj, defines), srcs))
    objs = [o for o, _ in results]
    # the device kernels' footprints (tests/test_kernel_footprint.py)
    res = {}
    for _, r in results:
        res.update(r)
    (obj / "kernel_resources.json").write_text(json.dumps(res, indent=1, sort_keys=True))
    lib = Path(lib_path) if variant else LIB
    tmp = lib.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    if variant:
        return lib
    # headless C++ driver over the C ABI (host C++ only, links liblt.so)
    cli = PKG / "lt_render_cli"
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(PKG / "host" / "lt_render_cli.cpp"),
           f"-L{PKG}", "-llt", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(cli) + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(str(cli) + ".tmp", cli)
    if verbose:
        print(f"built {LIB} and {cli}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
