"""Reference-schema scene JSON + OBJ loader (host side, CPU only)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import illume_amd

OBJ_QUAD = """# quad
v -0.5 0.5 0.0
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

OBJ_SMOOTH = """v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vn 0 0 2
vn 0 0 1
f 1//1 2//1 3//2
f 2//1 4//2 3//2
"""


def _scene(tmp_path):
    (tmp_path / "quad.obj").write_text(OBJ_QUAD)
    (tmp_path / "smooth.obj").write_text(OBJ_SMOOTH)
    js = {
        "mediums": [{"name": "m1", "absorption": [0.95, 0.6, -0.2], "scattering": 2.0, "g": 0.7}],
        "materials": [
            {"name": "white", "type": "lambert", "albedo": [0.7, 0.6, 0.5]},
            {"name": "light", "type": "emissive", "albedo": 17},
            {"name": "glass", "type": "refractive", "ior": 1.5, "albedo": [1, 1, 1], "medium": "m1"},
            {"name": "gold", "type": "roughconductor", "material": "Au", "roughness": 0.2},
            {"name": "bad", "type": "diffuse", "color": [1, 0, 0]},
            {"name": "unk", "type": "conductor", "material": "Xx"},
            {"name": "ggx", "type": "roughconductor", "eta": [0.18, 0.42, 1.37], "k": [3.42, 2.35, 1.77],
             "roughness": 0.3, "distribution": "ggx"},
            {"name": "check", "type": "lambert",
             "albedo": {"type": "checkerboard", "on_color": [1, 1, 1], "off_color": [0, 0, 0], "scale": [4, 4]}},
        ],
        "meshes": [{"name": "quad", "file": "quad.obj", "has_texcoords": True},
                   {"name": "sm", "file": "smooth.obj", "face_normals": False, "flip_normals": True}],
        "primitives": [
            {"type": "mesh_instance", "mesh": "quad", "material": "white",
             "transform": {"translation": [1, 2, 3], "scale": [2, 2, 2], "rotation": [90, 0, 0]}},
            {"type": "mesh_instance", "mesh": "sm", "material": "missing-material"},
            {"type": "sphere", "radius": 0.5, "center": [0, 1, 0], "material": "glass"},
        ],
        "camera": {"transform": {"translation": [0, 1, -3], "lookat": [0, 0, 0]}, "fov": 39, "depth_of_field": 1,
                   "aperture_radius": 0},
        "render_settings": {"resolution": [320, 200], "ray_bias": 0.01, "tonemapper": "Filmic"},
    }
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(js))
    return str(p)


def test_loader_reads_reference_schema(tmp_path):
    js = illume_amd.load_json(_scene(tmp_path))
    d = js.desc
    st = js.settings()
    assert (st.width, st.height, st.tonemap) == (320, 200, 2) and st.ray_bias == pytest.approx(0.01)
    assert d.num_meshes == 2 and d.num_instances == 2 and d.num_spheres == 1
    # mediums store 1 - absorption (src/medium.cu:9)
    assert list(d.mediums[0].sigma_a) == pytest.approx([0.05, 0.4, 1.2], abs=1e-6)
    m = [d.materials[i] for i in range(d.num_materials)]
    assert m[0].type == 1 and list(m[0].albedo.color) == pytest.approx([0.7, 0.6, 0.5])
    assert m[1].type == 0 and list(m[1].albedo.color) == pytest.approx([17, 17, 17])
    assert m[2].type == 3 and m[2].medium == 0 and m[2].ior == pytest.approx(1.5)
    assert m[3].type == 7 and list(m[3].eta) == pytest.approx([0.14311, 0.37495, 1.44247])
    assert m[3].roughness.color[0] == pytest.approx(0.2)
    assert m[4].type == 1 and list(m[4].albedo.color) == [0, 0, 0]          # stale schema -> black diffuse
    assert list(m[5].eta) == pytest.approx([0.21249, 0.97909, 1.3343])     # unknown conductor -> Cu
    assert m[6].distribution == 1 and list(m[6].k) == pytest.approx([3.42, 2.35, 1.77])
    assert m[7].albedo.type == 1 and list(m[7].albedo.scale) == [4, 4]
    assert m[0].roughness.color[0] == pytest.approx(1e-4)                  # default roughness
    # unknown material name on a primitive -> black diffuse appended
    assert d.num_materials == 9 and d.instances[1].material == 8 and d.materials[8].type == 1
    q = d.meshes[0]
    assert q.num_vertices == 4 and q.num_triangles == 2 and q.has_texcoords == 1
    idx = np.ctypeslib.as_array(q.indices, shape=(6,))
    assert list(idx) == [0, 1, 2, 2, 3, 0]
    s = d.meshes[1]
    assert s.face_normals == 0 and s.num_triangles == 2
    nrm = np.ctypeslib.as_array(s.normals, shape=(s.num_vertices * 3,)).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0)                  # normalised on load
    assert np.all(nrm[:, 2] < 0)                                           # flip_normals negates
    t = d.instances[0].transform
    assert list(t.translation) == [1, 2, 3] and list(t.rotation) == [90, 0, 0]
    assert d.camera.transform.has_lookat == 1 and d.camera.fov == pytest.approx(39)


def test_loader_errors(tmp_path):
    with pytest.raises(illume_amd.IlmError) as e:
        illume_amd.load_json(str(tmp_path / "missing.json"))
    assert e.value.code == -4
    bad = tmp_path / "bad.json"
    bad.write_text('{"materials": [ {"name": "a", ')
    with pytest.raises(illume_amd.IlmError) as e:
        illume_amd.load_json(str(bad))
    assert e.value.code == -5
    nomesh = tmp_path / "nomesh.json"
    nomesh.write_text(json.dumps({"meshes": [{"name": "x", "file": "nope.obj"}]}))
    with pytest.raises(illume_amd.IlmError) as e:
        illume_amd.load_json(str(nomesh))
    assert e.value.code == -4


def test_loader_quads_and_negative_indices(tmp_path):
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf -4 -3 -2 -1\n")
    (tmp_path / "s.json").write_text(json.dumps({"meshes": [{"name": "q", "file": "q.obj"}], "camera": {"fov": 40}}))
    d = illume_amd.load_json(str(tmp_path / "s.json")).desc
    assert d.meshes[0].num_triangles == 2 and d.meshes[0].num_vertices == 4
    assert d.camera.dof == 1.0 and d.camera.aperture == 0.0   # no depth_of_field: focus at 1


REF_CORNELL = "/root/reference/res/scenes/cornell_box.json"


@pytest.mark.skipif(not os.path.exists(REF_CORNELL), reason="reference checkout not present")
def test_loader_reads_the_reference_cornell_scene():
    js = illume_amd.load_json(REF_CORNELL)
    d = js.desc
    st = js.settings()
    assert (st.width, st.height) == (1024, 1024) and st.ray_bias == pytest.approx(0.01)
    assert d.num_meshes == 2 and d.num_instances == 8
    assert d.meshes[0].num_triangles == 12 and d.meshes[1].num_triangles == 2
    assert d.materials[3].type == 0   # "light" emissive


def test_loader_reads_generator_obj_exactly(tmp_path):
    """A generated mesh written as OBJ (shortest round-trip float text) reads back with the same
    triangles bit for bit; vertices are renumbered in first-occurrence order (as
    src/primitives/mesh.cu:264-340 merges them), so positions are compared per triangle. With
    vertex normals (f a//a ...) the smooth-normal path reads normalised normals."""
    import illume_amd.scenes as S
    m = S.displaced_icosphere(30, 5)
    S.write_obj(m, str(tmp_path / "m.obj"))
    (tmp_path / "s.json").write_text(json.dumps({"meshes": [{"name": "m", "file": "m.obj"}], "camera": {"fov": 40}}))
    js = illume_amd.load_json(str(tmp_path / "s.json"))
    d = js.desc.meshes[0]
    pos = np.ctypeslib.as_array(d.positions, shape=(d.num_vertices * 3,)).reshape(-1, 3)
    idx = np.ctypeslib.as_array(d.indices, shape=(d.num_triangles * 3,)).reshape(-1, 3)
    assert d.num_vertices == len(m["positions"]) and d.num_triangles == len(m["indices"])
    assert np.array_equal(pos[idx], m["positions"][m["indices"]])
    parse_s, merge_s = js.load_times()
    assert parse_s > 0 and merge_s > 0
    nm = dict(m, normals=m["positions"] / np.linalg.norm(m["positions"], axis=1, keepdims=True) * 2.0)
    S.write_obj(nm, str(tmp_path / "n.obj"), normals=True)
    (tmp_path / "n.json").write_text(json.dumps({"meshes": [{"name": "n", "file": "n.obj", "face_normals": False}],
                                                 "camera": {"fov": 40}}))
    d = illume_amd.load_json(str(tmp_path / "n.json")).desc.meshes[0]
    nrm = np.ctypeslib.as_array(d.normals, shape=(d.num_vertices * 3,)).reshape(-1, 3)
    assert d.face_normals == 0 and np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-6)
