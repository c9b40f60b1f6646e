"""Register, scratch and occupancy footprint of the kernels whose occupancy is
capped with amdgpu_waves_per_eu (CPU only: read from the compiler's
kernel-resource-usage remarks that lumen_amd/build.py records).

Three of these kernels once produced wrong images at tighter caps (the Path
megakernel at 5 waves per SIMD, the VCM light pass and the adjoint at 4), a
result that depended on how the compiler allocated registers and spilled.
The shipped caps and the current footprints are the ones the GPU replay
tests validate, so any change of a capped kernel's registers, scratch or
occupancy fails here: after such a change, run the GPU replay tests (also
with the caps that once failed) and then update
tests/golden/kernel_footprint.json."""
import json
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
RES = ROOT / "lumen_amd" / "_build" / "kernel_resources.json"
GOLDEN = Path(__file__).resolve().parent / "golden" / "kernel_footprint.json"


@pytest.fixture(scope="module")
def resources():
    if not RES.exists():
        pytest.skip("lumen_amd/_build/kernel_resources.json is written by lumen_amd.build")
    return json.loads(RES.read_text())


def _find(resources, tag):
    hits = [v for name, v in resources.items() if tag in name]
    assert len(hits) == 1, f"{tag}: {len(hits)} kernels"
    return hits[0]


def test_capped_kernels_keep_their_footprint(resources):
    golden = json.loads(GOLDEN.read_text())
    assert len(golden) >= 20
    changed = []
    for kernel, want in golden.items():
        got = _find(resources, want["mangled_tag"])
        for key in ("vgprs", "scratch_bytes", "waves_per_simd"):
            if got[key] != want[key]:
                changed.append(f"{kernel}.{key}: {want[key]} -> {got[key]}")
    assert not changed, "capped kernel footprints changed: " + "; ".join(changed)


def test_trace_kernels_fit_the_lds_budget(resources):
    """Every traversal kernel holds kLdsWords (32) words per thread of
    128-thread workgroups in LDS: 16 KiB per workgroup, so that ten
    workgroups fit the 160 KiB of a CU."""
    for name, v in resources.items():
        if any(t in name for t in ("k_path_regen", "k_path_mega", "k_bdpt_walk", "k_vcm_light", "k_sppm_eye")):
            assert v["lds_bytes"] == 16384, name
