"""Copies the reference's own scene files used by the tests into
tests/golden/scenes (data fixtures: the Cornell box of scenes/cornell_box with
its Path / BDPT / VCM / SPPM integrator blocks, the directional-light variant
with its wood texture and the "disney" variant; scenes/caustics and its
zoomed view; the occluded-light scenes occluded and occluded2; the material
test scene material_test/materials; the OBJ meshes gzip-compressed).  Run in a checkout that has the reference at
/root/reference."""
import gzip
import shutil
from pathlib import Path

SRC = Path("/root/reference/scenes")
DST = Path(__file__).resolve().parent / "scenes"

FILES = ["cornell_box/cornell_box_path.json", "cornell_box/cornell_box_bdpt.json", "cornell_box/cornell_box_vcm.json",
         "cornell_box/cornell_box_sppm.json", "cornell_box/cornell_box_dir.json", "cornell_box/wood1.jpg",
         "cornell_box/cornell_box_disney.json", "caustics.json", "caustics_zoomed.json", "occluded.json",
         "occluded2.json", "material_test/materials.json"]
MESHES = ["cornell_box/cornell_box.obj", "caustics.obj", "occluded.obj", "occluded2.obj",
          "material_test/materials.obj"]


def main():
    DST.mkdir(exist_ok=True)
    for f in FILES:
        shutil.copyfile(SRC / f, DST / Path(f).name)
    for f in MESHES:
        (DST / (Path(f).name + ".gz")).write_bytes(gzip.compress((SRC / f).read_bytes(), 9, mtime=0))


if __name__ == "__main__":
    main()
