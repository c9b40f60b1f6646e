"""The scene loader's JPEG texture decoder (lt_image_load_rgba8, CPU only).

Against two independent decoders: Pillow's libjpeg on files written by
Pillow in every chroma layout the loader meets (4:4:4, 4:2:2, 4:2:0, grey),
with restart markers, odd sizes and a one-block image; and, where the
reference tree is present, the reference's own stb_image
(oracle/_ref/libref_stbi.so, compiled from its libs/stb_image) on the
reference's texture scenes/cornell_box/wood1.jpg (a fixture under
tests/golden/scenes).  The IDCT and colour conversion round differently in
each decoder, so the bar is a difference of at most 4 in any 8-bit value and
a mean below 0.25; a wrong zig-zag table, quantiser order, chroma placement
or colour matrix gives differences of tens to hundreds.  Progressive files
are refused."""
import io
from pathlib import Path

import numpy as np
import pytest

import lumen_amd as la
import oracle
from lumen_amd import LtError

Image = pytest.importorskip("PIL.Image")

FIX = Path(__file__).resolve().parent / "golden" / "scenes"


def _photo(h, w, seed):
    """A smooth colour image with edges (what a texture looks like) plus noise."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.stack([128 + 100 * np.sin(x / 7.0 + seed), 128 + 90 * np.cos(y / 5.0),
                    128 + 80 * np.sin((x + y) / 11.0)], -1)
    img[(x // 9 + y // 13) % 2 == 0] *= 0.6
    img += rng.normal(0, 6, img.shape)
    return np.clip(img, 0, 255).astype(np.uint8)


def _close(got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    assert d.max() <= 4 and d.mean() < 0.25, (d.max(), d.mean())


@pytest.mark.parametrize("h,w,kw", [
    (64, 64, dict(subsampling=0, quality=90)),
    (37, 53, dict(subsampling=1, quality=85)),
    (50, 77, dict(subsampling=2, quality=75)),
    (50, 77, dict(subsampling=2, quality=95, restart_marker_blocks=3)),
    (33, 41, dict(subsampling=2, quality=60, restart_marker_rows=1)),
    (8, 8, dict(subsampling=0, quality=50)),
    (1, 1, dict(subsampling=2, quality=90)),
])
def test_decoder_matches_libjpeg(tmp_path, h, w, kw):
    src = Image.fromarray(_photo(h, w, h * w))
    path = tmp_path / "t.jpg"
    src.save(path, "JPEG", **kw)
    got = la.load_image_rgba8(path)
    want = np.asarray(Image.open(path).convert("RGBA"))
    assert got.shape == (h, w, 4)
    assert np.all(got[..., 3] == 255)
    _close(got, want)


def test_grey_jpeg(tmp_path):
    src = Image.fromarray(_photo(40, 30, 3)[..., 0], mode="L")
    path = tmp_path / "g.jpg"
    src.save(path, "JPEG", quality=80)
    got = la.load_image_rgba8(path)
    want = np.asarray(Image.open(path).convert("RGBA"))
    assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    _close(got, want)


def test_progressive_and_garbage_are_refused(tmp_path):
    path = tmp_path / "p.jpg"
    Image.fromarray(_photo(32, 32, 1)).save(path, "JPEG", progressive=True)
    with pytest.raises(LtError, match="LT_ERR_UNSUPPORTED"):
        la.load_image_rgba8(path)
    bad = tmp_path / "bad.jpg"
    bad.write_bytes(b"\xff\xd8\xff\xdb\x00\x43" + bytes(20))
    with pytest.raises(LtError, match="LT_ERR_PARSE"):
        la.load_image_rgba8(bad)
    bad.write_bytes(b"not a jpeg")
    with pytest.raises(LtError, match="LT_ERR_PARSE"):
        la.load_image_rgba8(bad)


def test_truncated_file_does_not_crash(tmp_path):
    path = tmp_path / "t.jpg"
    Image.fromarray(_photo(48, 48, 5)).save(path, "JPEG", quality=90)
    data = path.read_bytes()
    for cut in (len(data) // 3, len(data) // 2, len(data) - 3):
        path.write_bytes(data[:cut])
        try:
            img = la.load_image_rgba8(path)
            assert img.shape == (48, 48, 4)
        except LtError:
            pass


def test_reference_texture_against_libjpeg_and_stb_image():
    path = FIX / "wood1.jpg"
    got = la.load_image_rgba8(path)
    assert got.shape == (194, 259, 4)
    _close(got, np.asarray(Image.open(path).convert("RGBA")))
    ref = oracle.ref_stbi_load(str(path))
    if ref is None:
        pytest.skip("oracle/_ref (the reference's stb_image) is built only where the reference tree exists")
    _close(got, ref)
    assert (got == ref).mean() > 0.95
