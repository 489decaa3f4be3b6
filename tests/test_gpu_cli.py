"""The `illume` command line (src/main.cu:12-43): renders a reference-schema scene JSON and saves
<save>-<W>x<H>-<spp>spp-<maxdepth>md.png (and .pfm). The PNG's pixels must equal, byte for byte, a render of
the same scene through the Python API with the scene's own render_settings; the .pfm written
beside it must hold the API's accumulated radiance divided by spp, bit for bit."""
import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

import illume_amd

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(illume_amd.LIB_PATH), "illume")

OBJ_BOX = """v -1 -1 -1
v 1 -1 -1
v 1 1 -1
v -1 1 -1
v -1 -1 1
v 1 -1 1
v 1 1 1
v -1 1 1
f 1 2 3 4
f 5 8 7 6
f 1 5 6 2
f 4 3 7 8
f 2 6 7 3
f 1 4 8 5
"""


def _scene(tmp_path, width, height):
    (tmp_path / "box.obj").write_text(OBJ_BOX)
    js = {
        "materials": [
            {"name": "white", "type": "lambert", "albedo": [0.7, 0.7, 0.7]},
            {"name": "red", "type": "lambert", "albedo": [0.7, 0.1, 0.1]},
            {"name": "light", "type": "emissive", "albedo": 6},
            {"name": "metal", "type": "roughconductor", "material": "Au", "roughness": 0.3},
        ],
        "meshes": [{"name": "box", "file": "box.obj", "flip_normals": True}],
        "primitives": [
            {"type": "mesh_instance", "mesh": "box", "material": "white",
             "transform": {"scale": [2, 2, 2]}},
            {"type": "mesh_instance", "mesh": "box", "material": "red",
             "transform": {"translation": [0.6, -1.5, 0.4], "scale": [0.4, 0.5, 0.4], "rotation": [0, 30, 0]}},
            {"type": "sphere", "radius": 0.35, "center": [0, 1.6, 0], "material": "light"},
            {"type": "sphere", "radius": 0.5, "center": [-0.7, -1.5, -0.3], "material": "metal"},
        ],
        # no depth_of_field or aperture_radius: a pinhole focused at distance 1
        "camera": {"transform": {"translation": [0, 0, -1.9], "lookat": [0, 0, 0]}, "fov": 60},
        "render_settings": {"resolution": [width, height], "ray_bias": 1e-3, "tonemapper": "Filmic"},
    }
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(js))
    return str(p)


def _read_png(path):
    """RGB8 pixels of a non-interlaced PNG whose rows all use filter 0 (what the CLI writes)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xffffffff
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            assert (depth, ctype, interlace) == (8, 2, 0)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 3)


def _read_pfm(path):
    """Float RGB PFM (top row first in the returned array)."""
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = (int(x) for x in parts[1].split())
    scale = float(parts[2])
    px = np.frombuffer(parts[3], "<f4" if scale < 0 else ">f4").reshape(h, w, 3)
    return px[::-1].astype(np.float32)


@pytest.mark.parametrize("width,height,spp,max_depth", [(48, 32, 4, 5), (37, 23, 1, 2)])
def test_cli_png_equals_api_render(tmp_path, width, height, spp, max_depth):
    assert os.path.exists(CLI), "illume CLI not built"
    scene = _scene(tmp_path, width, height)
    r = subprocess.run([CLI, "out", scene, str(spp), str(max_depth)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    name = tmp_path / ("out-%dx%d-%dspp-%dmd.png" % (width, height, spp, max_depth))
    assert "Saved to: %s" % name.name in r.stdout
    assert "Rays per sec:" in r.stdout
    png = _read_png(str(name))

    pfm_name = tmp_path / ("out-%dx%d-%dspp-%dmd.pfm" % (width, height, spp, max_depth))
    assert "Saved to: %s" % pfm_name.name in r.stdout
    pfm = _read_pfm(str(pfm_name))

    js = illume_amd.load_json(scene)
    p = js.settings()
    p.spp, p.max_depth, p.first_sample = spp, max_depth, 0
    with illume_amd.Context(0) as ctx:
        acc, pix, info = ctx.render(js.upload(ctx), p)
    assert info["segments"] > 0
    assert pix.max() <= 255
    assert png.shape == (height, width, 3)
    assert np.array_equal(png, pix.astype(np.uint8))
    assert png.std() > 0   # not a blank frame
    # the PFM holds the mean radiance: the API's sums over spp, in float
    assert pfm.shape == (height, width, 3)
    assert np.array_equal(pfm, acc / np.float32(spp))


def test_cli_usage_and_errors(tmp_path):
    assert os.path.exists(CLI), "illume CLI not built"
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage" in r.stdout
    r = subprocess.run([CLI, "out", str(tmp_path / "missing.json"), "1", "1"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "illume:" in r.stderr
    assert not any(f.endswith(".png") for f in os.listdir(str(tmp_path)))
