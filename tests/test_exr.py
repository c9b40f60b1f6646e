"""OpenEXR output and ground-truth input (lt_exr_save / lt_exr_load), host
code only.  The writer replaces ImageUtils::save_exr
(src/Framework/ImageUtils.cpp:23-89: channels B, G, R, HALF storage, no
compression); the reader replaces ImageUtils::load_exr (ImageUtils.cpp:9-21).

Pinned by: the byte layout the OpenEXR file-format document fixes (magic,
version word, attribute list, offset table, chunk headers); numpy's float16
conversion (IEEE round to nearest even) for the HALF rounding; and the two
ZIP-compressed FLOAT images the reference ships (assets/kernels/*.exr, point
spread kernels: finite, non-negative, peaked at pixel (256, 256), nearly
symmetric about it), read where the reference tree is present."""
import ctypes as C
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import lumen_amd as la
from lumen_amd import LtError

REF_KERNELS = Path("/root/reference/assets/kernels")


def parse_header(b: bytes):
    assert struct.unpack("<I", b[:4])[0] == 20000630
    assert struct.unpack("<I", b[4:8])[0] == 2
    attrs, i = {}, 8
    while b[i] != 0:
        name = b[i:b.index(b"\0", i)].decode()
        i += len(name) + 1
        typ = b[i:b.index(b"\0", i)].decode()
        i += len(typ) + 1
        size = struct.unpack("<i", b[i:i + 4])[0]
        attrs[name] = (typ, b[i + 4:i + 4 + size])
        i += 4 + size
    return attrs, i + 1


def chlist(v: bytes):
    out, i = [], 0
    while v[i] != 0:
        name = v[i:v.index(b"\0", i)].decode()
        i += len(name) + 1
        ptype, lin, xs, ys = struct.unpack("<iIii", v[i:i + 16])
        out.append((name, ptype, lin, xs, ys))
        i += 16
    return out


def test_float_file_layout(tmp_path):
    img = np.zeros((2, 3, 4), np.float32)
    img[..., 0] = [[1, 2, 3], [4, 5, 6]]
    img[..., 1] = 10 * img[..., 0]
    img[..., 2] = -img[..., 0]
    img[..., 3] = 7.0  # alpha is not written
    p = tmp_path / "a.exr"
    la.save_exr(p, img, half=False)
    b = p.read_bytes()
    attrs, end = parse_header(b)
    assert set(attrs) == {"channels", "compression", "dataWindow", "displayWindow", "lineOrder", "pixelAspectRatio",
                          "screenWindowCenter", "screenWindowWidth"}
    assert attrs["channels"][0] == "chlist"
    assert chlist(attrs["channels"][1]) == [("B", 2, 0, 1, 1), ("G", 2, 0, 1, 1), ("R", 2, 0, 1, 1)]
    assert attrs["compression"] == ("compression", b"\x00")
    assert attrs["dataWindow"] == ("box2i", struct.pack("<4i", 0, 0, 2, 1))
    assert attrs["displayWindow"] == attrs["dataWindow"]
    assert attrs["lineOrder"] == ("lineOrder", b"\x00")
    assert attrs["pixelAspectRatio"] == ("float", struct.pack("<f", 1.0))
    offs = struct.unpack("<2Q", b[end:end + 16])
    line = 3 * 3 * 4
    assert offs == (end + 16, end + 16 + 8 + line)
    for y, o in enumerate(offs):
        assert struct.unpack("<iI", b[o:o + 8]) == (y, line)
        px = np.frombuffer(b[o + 8:o + 8 + line], np.float32).reshape(3, 3)  # channel, x
        np.testing.assert_array_equal(px[0], img[y, :, 2])
        np.testing.assert_array_equal(px[1], img[y, :, 1])
        np.testing.assert_array_equal(px[2], img[y, :, 0])
    assert len(b) == offs[1] + 8 + line


def half_values():
    rng = np.random.default_rng(5)
    special = np.array([0.0, -0.0, 1.0, -2.5, 65504.0, 65519.99, 65520.0, 1e9, -1e9, np.inf, -np.inf,
                        2.0**-14, 2.0**-24, 2.0**-25, 2.0**-25 * 1.0000001, 3 * 2.0**-25, 2.0**-26, 1e-30,
                        1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, 1.0 + 2.0**-11 + 2.0**-20, 2049.0, 2051.0],
                       np.float32)
    ties = (np.arange(1, 2048, dtype=np.float32) + 0.5).astype(np.float32) * np.float32(2.0**-10)  # exact half ulps
    rand = (rng.standard_normal(20000) * np.exp(rng.uniform(-30, 12, 20000))).astype(np.float32)
    return np.concatenate([special, ties, rand, -rand[:1000]]).astype(np.float32)


def test_half_rounding_matches_ieee_round_to_nearest_even(tmp_path):
    v = half_values()
    n = v.size
    img = np.zeros((1, n, 4), np.float32)
    img[0, :, 0] = v
    img[0, :, 1] = v[::-1]
    img[0, :, 2] = np.nan
    p = tmp_path / "h.exr"
    la.save_exr(p, img, half=True)
    b = p.read_bytes()
    attrs, end = parse_header(b)
    assert [c[1] for c in chlist(attrs["channels"][1])] == [1, 1, 1]
    (o,) = struct.unpack("<Q", b[end:end + 8])
    assert struct.unpack("<iI", b[o:o + 8]) == (0, 3 * n * 2)
    h = np.frombuffer(b[o + 8:o + 8 + 6 * n], np.uint16).reshape(3, n)
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(h[2], want)          # R
    np.testing.assert_array_equal(h[1], want[::-1])    # G
    assert np.isnan(h[0].view(np.float16)).all()       # B: NaN stays NaN


def test_round_trips(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.uniform(-3, 100, (37, 53, 4)).astype(np.float32)
    la.save_exr(tmp_path / "f.exr", img, half=False)
    back = la.load_exr(tmp_path / "f.exr")
    assert back.shape == img.shape
    np.testing.assert_array_equal(back[..., :3], img[..., :3])
    assert (back[..., 3] == 1.0).all()
    la.save_exr(tmp_path / "h.exr", img, half=True)
    back = la.load_exr(tmp_path / "h.exr")
    np.testing.assert_array_equal(back[..., :3], img[..., :3].astype(np.float16).astype(np.float32))


def zip_filter(raw: bytes) -> bytes:
    """The two ZIP filters of the OpenEXR format: even bytes then odd bytes,
    then each byte minus its predecessor plus 128 (mod 256)."""
    a = np.frombuffer(raw, np.uint8)
    s = np.concatenate([a[0::2], a[1::2]]).astype(np.int64)
    d = s.copy()
    d[1:] = (s[1:] - s[:-1] + 128) % 256
    return d.astype(np.uint8).tobytes()


def recompress(src: bytes, lines_per_chunk: int, comp: int) -> bytes:
    """Rewrites an uncompressed scanline file written by lt_exr_save as ZIPS
    (1 line per chunk) or ZIP (16 lines per chunk)."""
    attrs, end = parse_header(src)
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    H = y1 - y0 + 1
    offs = struct.unpack(f"<{H}Q", src[end:end + 8 * H])
    lines = []
    for o in offs:
        y, n = struct.unpack("<iI", src[o:o + 8])
        lines.append(src[o + 8:o + 8 + n])
    head = bytearray(src[:end - 1])
    i = head.index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    head[i] = comp
    head += b"\0"
    chunks = []
    for c0 in range(0, H, lines_per_chunk):
        raw = b"".join(lines[c0:c0 + lines_per_chunk])
        z = zlib.compress(zip_filter(raw))
        if len(z) >= len(raw):
            z = raw  # an incompressible chunk is stored as is
        chunks.append(struct.pack("<iI", y0 + c0, len(z)) + z)
    table, at = [], len(head) + 8 * len(chunks)
    for c in chunks:
        table.append(at)
        at += len(c)
    return bytes(head) + struct.pack(f"<{len(table)}Q", *table) + b"".join(chunks)


@pytest.mark.parametrize("lines,comp", [(1, 2), (16, 3)])
def test_zip_and_zips_chunks(tmp_path, lines, comp):
    y, x = np.mgrid[0:45, 0:29].astype(np.float32)
    img = np.stack([np.sin(x / 5) + y, x * y, np.cos(y / 3), np.ones_like(x)], -1).astype(np.float32)
    img[3, 4, :3] = [1e30, -7, 0.125]
    la.save_exr(tmp_path / "n.exr", img, half=(comp == 2))
    z = recompress((tmp_path / "n.exr").read_bytes(), lines, comp)
    (tmp_path / "z.exr").write_bytes(z)
    want = la.load_exr(tmp_path / "n.exr")
    np.testing.assert_array_equal(la.load_exr(tmp_path / "z.exr"), want)


@pytest.mark.skipif(not REF_KERNELS.is_dir(), reason="the reference tree is not present on this machine")
@pytest.mark.parametrize("name", ["Pupil512.exr", "Octagonal512.exr"])
def test_reference_zip_images(name):
    b = (REF_KERNELS / name).read_bytes()
    attrs, _ = parse_header(b)
    assert attrs["compression"][1] == b"\x03"  # ZIP, 16 lines per chunk
    assert [c[:2] for c in chlist(attrs["channels"][1])] == [("B", 2), ("G", 2), ("R", 2)]
    img = la.load_exr(REF_KERNELS / name)
    assert img.shape == (512, 512, 4)
    rgb = img[..., :3]
    assert np.isfinite(rgb).all() and (rgb > 0).all()
    assert (img[..., 3] == 1).all()
    for c in range(3):
        assert np.unravel_index(np.argmax(rgb[..., c]), (512, 512)) == (256, 256)
    core = rgb[1:, 1:]  # rows and columns 1..511 mirror about 256
    peak = rgb.max()
    assert np.abs(core - core[:, ::-1]).max() < 0.1 * peak
    assert np.abs(core - core[::-1]).max() < 0.1 * peak
    # the energy sits near the centre: a mis-decoded chunk scatters it
    r = np.hypot(*np.mgrid[-256:256, -256:256])
    assert rgb[r < 64].sum() > 0.5 * rgb.sum()


def test_error_paths(tmp_path):
    lib = la.lib()
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(256)
    assert lib.lt_exr_load(str(tmp_path / "missing.exr").encode(), None, 0, C.byref(w), C.byref(h), err, 256) == 1
    (tmp_path / "junk.exr").write_bytes(b"not an exr file at all")
    assert lib.lt_exr_load(str(tmp_path / "junk.exr").encode(), None, 0, C.byref(w), C.byref(h), err, 256) == 5
    assert b"not an OpenEXR" in err.value
    img = np.ones((4, 5, 4), np.float32)
    la.save_exr(tmp_path / "s.exr", img)
    buf = np.zeros(4 * 5 * 4 - 1, np.float32)
    assert lib.lt_exr_load(str(tmp_path / "s.exr").encode(), buf.ctypes.data, buf.size, C.byref(w), C.byref(h),
                           err, 256) == 1
    tiled = bytearray((tmp_path / "s.exr").read_bytes())
    tiled[5] |= 0x02  # version flag bit 9: tiled
    (tmp_path / "t.exr").write_bytes(bytes(tiled))
    assert lib.lt_exr_load(str(tmp_path / "t.exr").encode(), None, 0, C.byref(w), C.byref(h), err, 256) == 6
    trunc = (tmp_path / "s.exr").read_bytes()[:60]
    (tmp_path / "u.exr").write_bytes(trunc)
    assert lib.lt_exr_load(str(tmp_path / "u.exr").encode(), None, 0, C.byref(w), C.byref(h), err, 256) == 5
    with pytest.raises(LtError):
        la.save_exr(tmp_path / "no_such_dir" / "x.exr", img)
    assert lib.lt_exr_save(str(tmp_path / "x.exr").encode(), img.ctypes.data, 5, 4, 3, err, 256) == 1
    assert lib.lt_exr_save(str(tmp_path / "x.exr").encode(), None, 5, 4, 1, err, 256) == 1


def _load_status(path):
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(256)
    buf = np.zeros(4 * 64 * 64, np.float32)
    st = la.lib().lt_exr_load(str(path).encode(), buf.ctypes.data, buf.size, C.byref(w), C.byref(h), err, 256)
    return st, err.value.decode()


def test_scanline_coverage_is_checked(tmp_path):
    """Every scanline must come from exactly one chunk: a file whose offset
    table points twice at the same chunk (one scanline given twice, another
    missing) is rejected."""
    img = np.random.default_rng(0).random((4, 5, 4)).astype(np.float32)
    la.save_exr(tmp_path / "s.exr", img, half=False)
    b = bytearray((tmp_path / "s.exr").read_bytes())
    _, end = parse_header(bytes(b))
    offs = list(struct.unpack("<4Q", b[end:end + 32]))
    offs[2] = offs[1]  # chunk of scanline 1 twice, scanline 2 never
    b[end:end + 32] = struct.pack("<4Q", *offs)
    (tmp_path / "dup.exr").write_bytes(bytes(b))
    st, msg = _load_status(tmp_path / "dup.exr")
    assert st == 5 and "given twice" in msg
    # a chunk header naming another scanline than its slot leaves one uncovered
    b = bytearray((tmp_path / "s.exr").read_bytes())
    o3 = offs[3]
    b[o3:o3 + 4] = struct.pack("<i", 0)
    (tmp_path / "gap.exr").write_bytes(bytes(b))
    st, msg = _load_status(tmp_path / "gap.exr")
    assert st == 5 and "given twice" in msg


def test_channel_count_is_limited(tmp_path):
    """A channel list longer than 64 entries is refused before any pixel
    buffer is sized from it."""
    img = np.ones((2, 2, 4), np.float32)
    la.save_exr(tmp_path / "s.exr", img, half=False)
    b = (tmp_path / "s.exr").read_bytes()
    attrs, end = parse_header(b)
    entry = lambda n: n.encode() + b"\0" + struct.pack("<iIii", 2, 0, 1, 1)
    chl = b"".join(entry(f"X{i}") for i in range(65)) + b"\0"
    hdr = struct.pack("<II", 20000630, 2)
    for name, (typ, val) in attrs.items():
        if name == "channels":
            val = chl
        hdr += name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(val)) + val
    hdr += b"\0"
    (tmp_path / "many.exr").write_bytes(hdr + b[end:])
    st, msg = _load_status(tmp_path / "many.exr")
    assert st == 6 and "channels" in msg
