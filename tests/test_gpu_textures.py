"""GPU parity of the camera and texture paths the configuration scenes do not reach:

* a thin-lens camera (aperture > 0: the lens sample of src/renderer.cu:97-102, src/math/sample.cu
  sample_circle) on the Cornell box, per-path radiance and a whole render against the oracle;
* a scene read by the scene-JSON loader whose textures and meshes come from files: an 8-bit PPM
  or PNG albedo (point sampled, src/texture.cu:112,156) on a textured OBJ quad and on a sphere (sphere
  uv, src/primitives/sphere.cu:74-75), a float PFM or Radiance HDR environment (bilinear, env uv
  src/renderer.cu:380-381), a generated 8,000-triangle OBJ with vertex normals (smooth shading)
  in rough gold, and a thin-lens camera; device and oracle render the same loaded descriptor.
"""
import json

import numpy as np
import pytest

import illume_amd
import illume_amd.scenes as S
from illume_amd._abi import render_params
from oracle.pyoracle import OracleScene
from image_files import hdr_bytes, png_bytes, rgbe_encode

pytestmark = pytest.mark.gpu


def _p(st, **kw):
    d = dict(st)
    d.update(kw)
    return render_params(**d)


def _compare_render(ctx, sc, orc, p, label):
    acc, px, info = ctx.render(sc, p)
    oacc, opx, osegs, _ = orc.render(p, kahan=True)
    oacc = oacc.reshape(acc.shape)
    rel = np.abs(acc - oacc) / np.maximum(np.abs(oacc), 1e-6)
    frac = np.mean(np.all(rel <= 1e-4, axis=2))
    lit = float(np.mean(acc.max(axis=2) > 0))
    print("%s: segments gpu %d oracle %d, pixels within 1e-4 %.5f, lit %.3f" % (label, info["segments"], osegs, frac, lit))
    assert info["segments"] == osegs
    assert frac >= 0.999 and lit > 0.2
    return acc


def _path_parity(ctx, sc, orc, p, n, seed, label):
    rng = np.random.default_rng(seed)
    px = rng.integers(0, p.width * p.height, n)
    pairs = np.array([[q, s] for q in px for s in range(4)], np.uint32)
    g, gseg = ctx.path_radiance(sc, p, pairs)
    o, oseg = orc.path_radiance(p, pairs)
    rel = np.abs(g - o) / np.maximum(np.abs(o), 1e-6)
    ok = np.all(rel <= 1e-4, axis=1)
    print("%s: %d paths, within 1e-4 %.5f, bit-identical %.5f" % (label, len(pairs), ok.mean(), np.all(g == o, axis=1).mean()))
    assert ok.mean() >= 0.999
    assert np.mean(gseg == oseg) >= 0.999


def test_thin_lens_camera_matches_oracle(ctx):
    spec, st = S.cornell()
    spec = dict(spec, camera=dict(spec["camera"], aperture=0.06, dof=2.4))   # focused on the back wall
    sc = ctx.scene(spec)
    orc = OracleScene(spec)
    _path_parity(ctx, sc, orc, _p(st), 3000, 21, "thin lens paths")
    acc = _compare_render(ctx, sc, orc, _p(st, width=64, height=64, spp=32), "thin lens render")
    # the lens blurs: a pinhole render of the same scene differs
    pin_scene = ctx.scene(S.cornell()[0])
    pin, _, _ = ctx.render(pin_scene, _p(st, width=64, height=64, spp=32))
    pin_scene.free()
    assert not np.array_equal(pin, acc)
    sc.free()
    orc.free()


OBJ_QUAD = """v -0.5 0.5 0.0
v 0.5 0.5 0.0
v 0.5 -0.5 0.0
v -0.5 -0.5 0.0
vt 0.0 0.0
vt 1.0 0.0
vt 1.0 1.0
vt 0.0 1.0
f 1/1 2/2 3/3
f 3/3 4/4 1/1
"""


def _write_ppm(path, rgb):
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(rgb.astype(np.uint8).tobytes())


def _write_pfm(path, rgb):
    """Little-endian PFM (negative scale), rows bottom to top."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        f.write(np.ascontiguousarray(rgb[::-1], dtype="<f4").tobytes())


def _textured_json(tmp_path, png_hdr=False):
    """png_hdr: the albedo as an interlaced PNG and the environment as a run-length Radiance HDR
    file (the formats the reference's own scenes use) instead of PPM and PFM."""
    rng = np.random.default_rng(5)
    albedo = rng.integers(20, 250, (7, 13, 3))
    env = S.sky_env(48, 24, sun_radiance=60.0, sun_diam_deg=12.0)[..., :3]
    if png_hdr:
        (tmp_path / "albedo.png").write_bytes(png_bytes(albedo, 2, 8, interlace=1))
        (tmp_path / "sky.hdr").write_bytes(hdr_bytes(rgbe_encode(env)))
    else:
        _write_ppm(tmp_path / "albedo.ppm", albedo)
        _write_pfm(tmp_path / "sky.pfm", env)
    alb, sky = ("albedo.png", "sky.hdr") if png_hdr else ("albedo.ppm", "sky.pfm")
    (tmp_path / "quad.obj").write_text(OBJ_QUAD)
    ball = S.displaced_icosphere(20, 11)
    ball["normals"] = ball["positions"] / np.linalg.norm(ball["positions"], axis=1, keepdims=True)
    S.write_obj(ball, str(tmp_path / "ball.obj"), normals=True)
    js = {
        "environment_map": {"type": "bitmap", "file": sky},
        "materials": [
            {"name": "tex", "type": "lambert", "albedo": {"type": "bitmap", "file": alb, "scale": [3, 2]}},
            {"name": "ball", "type": "lambert", "albedo": {"type": "bitmap", "file": alb}},
            {"name": "glossy", "type": "plastic", "ior": 1.5, "albedo": {"type": "bitmap", "file": alb}},
            {"name": "metal", "type": "roughconductor", "material": "Au", "roughness": 0.25},
        ],
        "meshes": [{"name": "quad", "file": "quad.obj", "has_texcoords": True},
                   {"name": "ball", "file": "ball.obj", "face_normals": False}],
        "primitives": [
            {"type": "mesh_instance", "mesh": "quad", "material": "tex",
             "transform": {"translation": [0, 0, 0], "scale": [6, 6, 1], "rotation": [90, 0, 0]}},
            {"type": "mesh_instance", "mesh": "quad", "material": "glossy",
             "transform": {"translation": [0, 1, 2], "scale": [3, 2, 1], "rotation": [0, 0, 0]}},
            {"type": "sphere", "radius": 0.6, "center": [-0.4, 0.6, 0.3], "material": "ball"},
            {"type": "sphere", "radius": 0.4, "center": [0.9, 0.4, -0.2], "material": "ball"},
            {"type": "mesh_instance", "mesh": "ball", "material": "metal",
             "transform": {"translation": [0.2, 0.45, -0.9], "scale": [0.4, 0.4, 0.4]}},
        ],
        "camera": {"transform": {"translation": [0.3, 1.4, -3.2], "lookat": [0, 0.5, 0.5]}, "fov": 55,
                   "depth_of_field": 3.4, "aperture_radius": 0.02},
        "render_settings": {"resolution": [64, 48], "spp": 8, "max_depth": 6, "ray_bias": 0.0001},
    }
    path = tmp_path / "textured.json"
    path.write_text(json.dumps(js))
    return str(path)


@pytest.mark.parametrize("png_hdr", [False, True], ids=["ppm_pfm", "png_hdr"])
def test_file_textures_and_env_from_json_match_oracle(ctx, tmp_path, png_hdr):
    js = illume_amd.load_json(_textured_json(tmp_path, png_hdr))
    d = js.desc
    m = [d.materials[i] for i in range(d.num_materials)]
    assert m[0].albedo.type == 2 and m[0].albedo.is_hdr == 0 and (m[0].albedo.width, m[0].albedo.height) == (13, 7)
    assert d.has_environment == 1 and d.environment.is_hdr == 1 and (d.environment.width, d.environment.height) == (48, 24)
    assert d.camera.aperture == pytest.approx(0.02)
    st = js.settings()
    sc = js.upload(ctx)
    orc = OracleScene.from_desc(d, owner=js)
    p = render_params(width=st.width, height=st.height, spp=8, max_depth=6, ray_bias=st.ray_bias, tonemap=st.tonemap)
    label = "png/hdr" if png_hdr else "ppm/pfm"
    _path_parity(ctx, sc, orc, p, 1500, 22, label + " file-textured paths")
    _compare_render(ctx, sc, orc, p, label + " file-textured render")
    sc.free()
    orc.free()
    js.free()
