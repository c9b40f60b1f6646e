"""The C++ host adapter (lumen_amd/host/hip_integrator.hpp, Lumen's
Integrator contract over the C ABI) driven by the headless CLI renders the
same image as the Python mirror of the same interface."""
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import lumen_amd as la

pytestmark = pytest.mark.gpu
CLI = Path(la.__file__).resolve().parent / "lt_render_cli"


def read_pfm(p):
    data = Path(p).read_bytes()
    parts = data.split(b"\n", 3)
    w, h = map(int, parts[1].split())
    img = np.frombuffer(parts[3], "<f4").reshape(h, w, 3)
    return img[::-1]


def test_cli_matches_python_integrator(gpu_ctx):
    print("[gpu] cli cornell 64x48 x4", flush=True)
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "img.pfm"
        r = subprocess.run([str(CLI), "--scene", "cornell", "--width", "64", "--height", "48", "--spp", "4",
                            "--out", str(out)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        cli = read_pfm(out)
    sc = la.Scene.procedural("cornell")
    it = la.Path(sc, gpu_ctx, 64, 48)
    it.init()
    it.create_accel()
    for _ in range(4):
        it.render()
        it.update()
    py = it.output_tex[..., :3]
    it.destroy()
    assert np.array_equal(cli, py)


def test_cli_writes_exr(gpu_ctx):
    """--out *.exr writes what ImageUtils::save_exr writes (HALF), --exr-float
    the FLOAT file; both read back as the PFM image of the same run."""
    print("[gpu] cli cornell 64x48 x2 exr", flush=True)
    with tempfile.TemporaryDirectory() as d:
        imgs = {}
        for name, extra in (("a.pfm", []), ("a.exr", []), ("f.exr", ["--exr-float"])):
            out = Path(d) / name
            r = subprocess.run([str(CLI), "--scene", "cornell", "--width", "64", "--height", "48", "--spp", "2",
                                "--out", str(out), *extra], capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            imgs[name] = read_pfm(out) if name.endswith(".pfm") else la.load_exr(out)[..., :3]
    ref = imgs["a.pfm"]
    assert np.array_equal(imgs["f.exr"], ref)
    assert np.array_equal(imgs["a.exr"], ref.astype(np.float16).astype(np.float32))


@pytest.mark.parametrize("name,extra,py_kw", [
    ("bdpt", ["--mis", "power"], dict(mis="power")),
    ("vcm", ["--estimator", "corrected"], dict(estimator="corrected")),
    ("vcm", [], dict(estimator="reference")),
    ("sppm", [], {}),
])
def test_cli_bdpt_vcm_match_python_integrators(gpu_ctx, name, extra, py_kw):
    """The C++ factory builds BDPT, VCM and SPPM (RayTracer::create_integrator,
    RayTracer.cpp:244-282) from the scene's integrator block; their frames equal
    those of the Python mirror with the same options and seed."""
    print(f"[gpu] cli {name} {extra}", flush=True)
    seed = 9
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "img.pfm"
        r = subprocess.run([str(CLI), "--scene", "caustics", "--integrator", name, "--width", "48", "--height", "40",
                            "--spp", "3", "--seed", str(seed), "--out", str(out), *extra],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        cli = read_pfm(out)
    sc = la.Scene.procedural("caustics")
    it = la.create_integrator(name, sc, gpu_ctx, 48, 40, **py_kw)
    it.seed = seed
    it.init()
    it.create_accel()
    for _ in range(3):
        it.render()
        it.update()
    py = it.output_tex[..., :3]
    it.destroy()
    assert np.any(py > 0)
    if name == "vcm":  # float atomics sum the splats in a run-dependent order
        assert np.allclose(cli, py, rtol=1e-4, atol=1e-6)
    else:
        assert np.allclose(cli, py, rtol=1e-5, atol=1e-7)
