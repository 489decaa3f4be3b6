"""Image files read as bitmap textures by the scene loader (host side, CPU only).

The reference reads every bitmap through stb_image (src/imageio.cpp:6-21): paths ending in "hdr"
as float RGBA (src/texture.cu:49-55), everything else as 8-bit RGBA (src/texture.cu:57-63).
The files here are written by the encoders of tests/image_files.py, from the PNG specification (scanline
filters, sample packing, Adam7) and the Radiance RGBE format (adaptive run-length scanlines),
with Python's zlib for deflate. Expected texels are formed from the source samples directly,
never by decoding, so a mistake in the loader's unfiltering, unpacking, de-interlacing or RGBE
scaling fails a case here.
"""
import json
import struct
import zlib

import numpy as np

import illume_amd
from image_files import _chunk, hdr_bytes, paeth, png_bytes, rgbe_decode, rgbe_encode


def test_paeth_predictor_known_answers():
    assert paeth(10, 20, 15) == 15       # p = 15: c exact
    assert paeth(10, 20, 10) == 20       # p = 20: b exact
    assert paeth(20, 10, 10) == 20       # p = 20: a exact
    assert paeth(5, 5, 5) == 5
    assert paeth(0, 4, 2) == 2           # p = 2, pa = 2, pb = 2, pc = 0 -> c
    assert paeth(3, 7, 5) == 5           # p = 5, pa = pb = 2, pc = 0 -> c
    assert paeth(4, 8, 0) == 8           # p = 12, pa = 8, pb = 4, pc = 12 -> b
    assert paeth(8, 4, 0) == 8           # p = 12, pa = 4, pb = 8 -> a
    assert paeth(6, 2, 4) == 4           # p = 4, pa = pb = 2, pc = 0 -> c
    assert paeth(2, 4, 6) == 2           # p = 0, pa = 2, pb = 4, pc = 6 -> a
    assert paeth(9, 3, 7) == 3           # p = 5, pa = 4, pb = 2, pc = 2 -> b (ties with c go to b)


# ---------------------------------------------------------------- loading through the scene JSON

def _load_textures(tmp_path, files):
    """Load a scene whose materials use each file as a bitmap albedo; returns the loaded texture
    descriptors (kept alive by the returned loader)."""
    mats = [{"name": "m%d" % i, "type": "lambert", "albedo": {"type": "bitmap", "file": f}} for i, f in enumerate(files)]
    js = {"materials": mats, "primitives": [{"type": "sphere", "radius": 1, "center": [0, 0, 0], "material": "m0"}],
          "camera": {"transform": {"translation": [0, 0, -4]}, "fov": 40},
          "render_settings": {"resolution": [8, 8], "spp": 1, "max_depth": 2}}
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(js))
    sc = illume_amd.load_json(str(p))
    d = sc.desc
    return sc, [d.materials[i].albedo for i in range(d.num_materials)]


def _texels(t):
    n = t.width * t.height * 4
    if t.is_hdr:
        return np.ctypeslib.as_array(t.texels_f32, (n,)).reshape(t.height, t.width, 4).copy()
    return np.ctypeslib.as_array(t.texels_u8, (n,)).reshape(t.height, t.width, 4).copy()


def _rgba8(rgb, alpha=None):
    h, w, _ = rgb.shape
    out = np.full((h, w, 4), 255, np.uint8)
    out[..., :3] = rgb
    if alpha is not None:
        out[..., 3] = alpha
    return out


def test_png_colour_types_depths_filters_and_interlace(tmp_path):
    rng = np.random.default_rng(11)
    cases = []   # (file name, png bytes, expected RGBA8)
    for w, h, il in ((13, 7, 0), (11, 9, 1), (3, 2, 1), (1, 1, 0), (17, 5, 1)):
        rgb = rng.integers(0, 256, (h, w, 3))
        cases.append(("rgb8_%dx%d_%d.png" % (w, h, il), png_bytes(rgb, 2, 8, il), _rgba8(rgb)))
        rgba = rng.integers(0, 256, (h, w, 4))
        cases.append(("rgba8_%dx%d_%d.png" % (w, h, il), png_bytes(rgba, 6, 8, il), rgba.astype(np.uint8)))
        rgb16 = rng.integers(0, 65536, (h, w, 3))
        cases.append(("rgb16_%dx%d_%d.png" % (w, h, il), png_bytes(rgb16, 2, 16, il), _rgba8(rgb16 >> 8)))
        rgba16 = rng.integers(0, 65536, (h, w, 4))
        cases.append(("rgba16_%dx%d_%d.png" % (w, h, il), png_bytes(rgba16, 6, 16, il), (rgba16 >> 8).astype(np.uint8)))
        ga = rng.integers(0, 256, (h, w, 2))
        cases.append(("ga8_%dx%d_%d.png" % (w, h, il), png_bytes(ga, 4, 8, il), _rgba8(np.repeat(ga[..., :1], 3, 2), ga[..., 1])))
        ga16 = rng.integers(0, 65536, (h, w, 2))
        cases.append(("ga16_%dx%d_%d.png" % (w, h, il), png_bytes(ga16, 4, 16, il),
                      _rgba8(np.repeat(ga16[..., :1] >> 8, 3, 2), ga16[..., 1] >> 8)))
        for depth, mul in ((1, 255), (2, 85), (4, 17), (8, 1)):
            g = rng.integers(0, 1 << depth, (h, w, 1))
            cases.append(("g%d_%dx%d_%d.png" % (depth, w, h, il), png_bytes(g, 0, depth, il), _rgba8(np.repeat(g * mul, 3, 2))))
            pal = rng.integers(0, 256, (1 << depth, 3))
            idx = rng.integers(0, 1 << depth, (h, w, 1))
            cases.append(("p%d_%dx%d_%d.png" % (depth, w, h, il), png_bytes(idx, 3, depth, il, palette=pal), _rgba8(pal[idx[..., 0]])))
        g16 = rng.integers(0, 65536, (h, w, 1))
        cases.append(("g16_%dx%d_%d.png" % (w, h, il), png_bytes(g16, 0, 16, il), _rgba8(np.repeat(g16 >> 8, 3, 2))))
    # the image data split over several IDAT chunks, ancillary chunks before it, one filter per file
    rgb = rng.integers(0, 256, (6, 10, 3))
    for ft in range(5):
        cases.append(("split_f%d.png" % ft, png_bytes(rgb, 2, 8, filters=(ft,), idat_parts=3,
                                                      extra=[(b"tEXt", b"Comment\x00test"), (b"gAMA", struct.pack(">I", 45455))]),
                      _rgba8(rgb)))
    for name, data, _ in cases:
        (tmp_path / name).write_bytes(data)
    sc, texs = _load_textures(tmp_path, [c[0] for c in cases])
    for (name, _, want), t in zip(cases, texs):
        assert t.type == 2 and t.is_hdr == 0, name
        assert (t.width, t.height) == (want.shape[1], want.shape[0]), name
        np.testing.assert_array_equal(_texels(t), want, err_msg=name)
    sc.free()


def test_png_and_ppm_of_one_image_give_the_same_texels(tmp_path):
    rng = np.random.default_rng(12)
    rgb = rng.integers(0, 256, (9, 14, 3)).astype(np.uint8)
    (tmp_path / "a.png").write_bytes(png_bytes(rgb, 2, 8, 1))
    (tmp_path / "a.ppm").write_bytes(b"P6\n# comment line\n14 9\n255\n" + rgb.tobytes())
    g = rgb[..., :1]
    (tmp_path / "g.pgm").write_bytes(b"P5 14 9 255\n" + g.tobytes())
    sc, (tp, tq, tg) = _load_textures(tmp_path, ["a.png", "a.ppm", "g.pgm"])
    np.testing.assert_array_equal(_texels(tp), _texels(tq))
    np.testing.assert_array_equal(_texels(tg), _rgba8(np.repeat(g, 3, 2)))
    sc.free()


def test_hdr_run_length_and_flat(tmp_path):
    rng = np.random.default_rng(13)
    cases = []
    for w, h in ((37, 5), (8, 3), (300, 2)):
        rgb = rng.lognormal(0.0, 2.0, (h, w, 3)).astype(np.float32)
        rgb[0, : w // 2] = rgb[0, 0]            # long runs
        rgb[1, 3] = 0.0                          # a zero texel (exponent 0)
        e = rgbe_encode(rgb)
        want = rgbe_decode(e)
        # RGBE keeps 8 bits of the largest component: within 2^-7 of it
        assert np.all(np.abs(want - rgb) <= rgb.max(axis=2, keepdims=True) * 2.0 ** -7 + 1e-30)
        cases.append(("rle_%d.hdr" % w, hdr_bytes(e), want))
        cases.append(("flat_%d.hdr" % w, hdr_bytes(e, rle=False), want))   # first scanline not run-length
    small = rgbe_encode(rng.lognormal(0.0, 1.0, (4, 5, 3)))
    cases.append(("narrow.hdr", hdr_bytes(small, rle=False, magic=b"#?RGBE"), rgbe_decode(small)))   # w < 8: flat
    # exponents across the float range: 2^-135 (a subnormal scale) .. 2^119
    ex = np.zeros((1, 8, 4), np.uint8)
    ex[0, :, :3] = [[1, 2, 255], [128, 0, 3], [255, 255, 255], [7, 9, 11], [1, 1, 1], [200, 100, 50], [0, 0, 0], [5, 6, 7]]
    ex[0, :, 3] = [1, 2, 100, 128, 136, 200, 255, 255]
    cases.append(("exp.hdr", hdr_bytes(ex), rgbe_decode(ex)))
    for name, data, _ in cases:
        (tmp_path / name).write_bytes(data)
    sc, texs = _load_textures(tmp_path, [c[0] for c in cases])
    for (name, _, want), t in zip(cases, texs):
        assert t.type == 2 and t.is_hdr == 1, name
        got = _texels(t)
        np.testing.assert_array_equal(got[..., :3], want, err_msg=name)
        assert np.all(got[..., 3] == 1.0), name
    assert _texels(texs[-1])[0, 0, 0] == np.float32(2.0 ** -135)
    sc.free()


def test_unreadable_images_fall_back_to_black(tmp_path, capfd):
    rgb = np.zeros((4, 4, 3), np.int64)
    good = png_bytes(rgb, 2, 8)
    bad = {
        "trunc.png": good[:40],
        "crit.png": good[:33] + _chunk(b"ZZZZ", b"x") + good[33:],     # unknown critical chunk
        "badfilter.png": png_bytes(rgb, 2, 8, filters=(0,))[:0] + b"\x89PNG\r\n\x1a\n"
        + _chunk(b"IHDR", struct.pack(">IIBBBBB", 4, 4, 8, 2, 0, 0, 0))
        + _chunk(b"IDAT", zlib.compress(bytes([7] + [0] * 12) * 4)) + _chunk(b"IEND", b""),
        "palidx.png": png_bytes(np.full((2, 2, 1), 3), 3, 8, palette=[[1, 2, 3]]),
        "xyze.hdr": hdr_bytes(rgbe_encode(np.ones((2, 9, 3))), fmt=b"FORMAT=32-bit_rle_xyze"),
        "orient.hdr": hdr_bytes(rgbe_encode(np.ones((2, 9, 3))), orient=b"+Y 2 +X 9\n"),
        "shortrle.hdr": hdr_bytes(rgbe_encode(np.ones((2, 9, 3))))[:-3],
        "image.jpg": b"\xff\xd8\xff\xe0" + b"\x00" * 32,
        "missing.png": None,
    }
    for name, data in bad.items():
        if data is not None:
            (tmp_path / name).write_bytes(data)
    sc, texs = _load_textures(tmp_path, list(bad))
    for name, t in zip(bad, texs):
        assert t.type == 0 and tuple(t.color) == (0.0, 0.0, 0.0), name
    err = capfd.readouterr().err
    for name in bad:
        assert name in err
    sc.free()
