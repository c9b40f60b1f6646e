"""The autotuner's search and table logic (lumen_amd/autotune.py) on a
stand-in integrator whose frame time is a known function of the tuning
fields.  CPU only."""
import json

import lumen_amd.autotune as at


class FakeCtx:
    def __init__(self):
        self.t = {"path_blocks_per_cu": 10, "regen_min": 4, "refill_min": 32}
        self.calls = 0

    def set_tuning(self, **kw):
        self.t.update(kw)

    def get_tuning(self):
        return dict(self.t)

    def device_info(self):
        return 256, "gfx950:sramecc+:xnack-"


class FakeStats:
    def __init__(self, ms):
        self.kernel_ms = ms


class FakeIntegrator:
    """Frame time = a separable bowl with its minimum at blocks 12, regen 2,
    plus a coupling term, so the descent needs more than one field sweep."""

    def __init__(self):
        self.ctx = FakeCtx()

    def render(self):
        self.ctx.calls += 1

    def update(self):
        pass

    def stats(self):
        b, r = self.ctx.t["path_blocks_per_cu"], self.ctx.t["regen_min"]
        return FakeStats(4.0 + 0.01 * (b - 12) ** 2 + 0.02 * (r - 2) ** 2 + 0.001 * (b - 12) * (r - 2))


def test_tune_field_picks_the_minimum():
    it = FakeIntegrator()
    r = at.tune_field(it, "regen_min", [1, 2, 4, 8], frames=2)
    assert r["best"] == 2 and it.ctx.t["regen_min"] == 2
    assert r["ms"][2] == min(r["ms"].values())


def test_coordinate_descent_reaches_the_optimum():
    it = FakeIntegrator()
    r = at.tune(it, {"path_blocks_per_cu": [6, 8, 10, 12, 16], "regen_min": [1, 2, 4, 8, 16]}, frames=1)
    assert r["best"] == {"path_blocks_per_cu": 12, "regen_min": 2}
    assert it.ctx.t["path_blocks_per_cu"] == 12 and it.ctx.t["regen_min"] == 2
    assert abs(r["ms"] - 4.0) < 1e-9


def test_table_round_trip_and_apply(tmp_path):
    table = tmp_path / "tuning.json"
    at.store("gfx950", "path-persistent", "terrain", {"runtime": {"regen_min": 8}}, table)
    at.store("gfx950", "bdpt", "caustics", {"runtime": {"bdpt_blocks_per_cu": 3}}, table)
    d = json.loads(table.read_text())
    assert d["gfx950"]["path-persistent"]["terrain"]["runtime"] == {"regen_min": 8}
    ctx = FakeCtx()
    assert at.apply(ctx, "path-persistent", "terrain", table) == {"regen_min": 8}
    assert ctx.t["regen_min"] == 8
    assert at.apply(ctx, "path-persistent", "cornell", table) is None


def test_build_variants_respect_the_leaf_queue_bound():
    vs = at.build_variants()
    assert vs and all(v["LT_NODE_PEND_MAX"] + 8 <= v["LT_LEAF_QUEUE"] for v in vs)
    assert {"LT_LDS_STACK": 16, "LT_LEAF_QUEUE": 16, "LT_NODE_PEND_MAX": 8} in vs


def test_build_sweep_times_each_variant(tmp_path):
    built = []

    def fake_build(defines, lib_path):
        built.append(defines)
        lib_path.write_text("x")
        return lib_path

    cmd = ["python", "-c", "import os; print(len(os.environ['LT_LIBRARY']) / 1000.0)"]
    res = at.tune_build([{"LT_LDS_STACK": 12}, {"LT_LDS_STACK": 24}], cmd, build=fake_build)
    assert built == [("LT_LDS_STACK=12",), ("LT_LDS_STACK=24",)]
    assert set(res) == {"LT_LDS_STACK=12", "LT_LDS_STACK=24"} and all(v > 0 for v in res.values())


def test_stored_table_parses():
    if at.TABLE.exists():
        d = json.loads(at.TABLE.read_text())
        for arch, workloads in d.items():
            for wl, scenes in workloads.items():
                assert wl in at.SPACE
                for sc, entry in scenes.items():
                    for field in entry.get("runtime", {}):
                        assert field in at.SPACE[wl]
