"""ctypes mirror of include/illume_c.h and the conversion of a Python scene spec to an
ilm_scene_desc.

A scene spec is a plain dict (see illume_amd.scenes for generators):
    meshes:      [{positions (N,3) f32, indices (M,3) i32, normals, texcoords, face_normals,
                   has_texcoords, leaf_size}]
    instances:   [{mesh, material, transform}]
    spheres:     [{center, radius, material}]
    materials:   [{type, albedo, roughness, ior, medium, eta, k, distribution}]
    mediums:     [{sigma_a, sigma_s, g}]
    environment: texture or None
    camera:      {fov, dof, aperture, transform}
A transform is {translation, scale, rotation (degrees), lookat} or {matrix (4,4)}.
A texture is a number, an RGB triple, {"type": "checkerboard", on, off, scale} or
{"type": "bitmap", texels (H,W,4) float32 (HDR, bilinear) or uint8 (LDR, point), scale}.
"""
import ctypes as C

import numpy as np

ILM_OK = 0
ERRORS = {-1: "INVALID", -2: "NOMEM", -3: "DEVICE", -4: "IO", -5: "PARSE", -6: "BUSY",
          -7: "CANCELLED", -8: "TIMEOUT", -9: "UNSUPPORTED", -10: "STATE"}

TEX = {"constant": 0, "checkerboard": 1, "bitmap": 2}
MAT = {"emissive": 0, "lambert": 1, "diffuse": 1, "reflective": 2, "specular": 2, "refractive": 3,
       "roughreflective": 4, "roughrefractive": 5, "conductor": 6, "roughconductor": 7, "plastic": 8}
DIST = {"beckmann": 0, "ggx": 1}
TONEMAP = {"linear": 0, "reinhard": 1, "filmic": 2, "uncharted": 3}
BVH = {"fp32": 0, "quant8": 1}
TRI = {"mt": 0, "watertight": 1}
REQ_STATE = {0: "queued", 1: "running", 2: "done", 3: "cancelled", 4: "failed"}

F3 = C.c_float * 3


class TextureDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", F3), ("on", F3), ("off", F3), ("scale", C.c_float * 2),
                ("width", C.c_int32), ("height", C.c_int32), ("is_hdr", C.c_int32),
                ("texels_f32", C.POINTER(C.c_float)), ("texels_u8", C.POINTER(C.c_uint8))]


class MediumDesc(C.Structure):
    _fields_ = [("sigma_a", F3), ("sigma_s", C.c_float), ("g", C.c_float)]


class MaterialDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("albedo", TextureDesc), ("roughness", TextureDesc), ("ior", C.c_float),
                ("medium", C.c_int32), ("eta", F3), ("k", F3), ("distribution", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("texcoords", C.POINTER(C.c_float)), ("num_triangles", C.c_int32), ("indices", C.POINTER(C.c_int32)),
                ("face_normals", C.c_int32), ("has_texcoords", C.c_int32), ("leaf_size", C.c_int32)]


class TransformDesc(C.Structure):
    _fields_ = [("translation", F3), ("scale", F3), ("rotation", F3), ("has_lookat", C.c_int32), ("lookat", F3),
                ("has_matrix", C.c_int32), ("matrix", C.c_float * 16)]


class InstanceDesc(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("material", C.c_int32), ("transform", TransformDesc)]


class SphereDesc(C.Structure):
    _fields_ = [("center", F3), ("radius", C.c_float), ("material", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("fov", C.c_float), ("dof", C.c_float), ("aperture", C.c_float), ("transform", TransformDesc)]


class SceneDesc(C.Structure):
    _fields_ = [("num_meshes", C.c_int32), ("meshes", C.POINTER(MeshDesc)),
                ("num_instances", C.c_int32), ("instances", C.POINTER(InstanceDesc)),
                ("num_spheres", C.c_int32), ("spheres", C.POINTER(SphereDesc)),
                ("num_materials", C.c_int32), ("materials", C.POINTER(MaterialDesc)),
                ("num_mediums", C.c_int32), ("mediums", C.POINTER(MediumDesc)),
                ("has_environment", C.c_int32), ("environment", TextureDesc), ("camera", CameraDesc)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
                ("ray_bias", C.c_float), ("tonemap", C.c_int32), ("first_sample", C.c_int32)]


class Pixel(C.Structure):
    _fields_ = [("r", C.c_uint64), ("g", C.c_uint64), ("b", C.c_uint64)]


class Opts(C.Structure):
    _fields_ = [("bvh_layout", C.c_int32), ("tri_test", C.c_int32), ("paths_in_flight", C.c_int64),
                ("hbm_budget_bytes", C.c_int64), ("use_graphs", C.c_int32), ("max_queued", C.c_int32),
                ("extend_mode", C.c_int32), ("profile", C.c_int32), ("count_tests", C.c_int32), ("verbose", C.c_int32),
                ("sample_group", C.c_int32), ("tile_rows", C.c_int32), ("chunk_passes", C.c_int32),
                ("bvh_cache_dir", C.c_char_p), ("refill_min", C.c_int32), ("ext_fetch", C.c_int32),
                ("leaf_wait_pct", C.c_int32), ("shade_sort", C.c_int32),
                ("sched_fair", C.c_int32), ("camera_packets", C.c_int32), ("pipes", C.c_int32),
                ("lone_split", C.c_int32)]


class RenderInfo(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("device_ms", C.c_double), ("extend_ms", C.c_double),
                ("extend_launches", C.c_int32), ("chunks", C.c_int32), ("paths_per_chunk", C.c_int64),
                ("box_tests", C.c_uint64), ("tri_tests", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("requests_done", C.c_uint64), ("requests_rejected", C.c_uint64),
                ("requests_cancelled", C.c_uint64), ("busy_ms", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("hbm_budget_bytes", C.c_int64), ("hbm_in_use_bytes", C.c_int64),
                ("bvh_cache_hits", C.c_uint64), ("bvh_cache_misses", C.c_uint64)]


class ReqStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("samples_done", C.c_int32), ("segments", C.c_uint64),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("end_ms", C.c_double)]


class BvhInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_leaves", C.c_int32), ("max_depth", C.c_int32),
                ("max_leaf_size", C.c_int32), ("sah_cost", C.c_double)]


def struct_to_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _u8ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _vec3(v, default=(0.0, 0.0, 0.0)):
    if v is None:
        v = default
    if np.isscalar(v):
        v = (v, v, v)
    return F3(*[float(x) for x in v])


def texture_desc(t, keep):
    d = TextureDesc()
    d.scale[0] = d.scale[1] = 1.0
    if t is None:
        t = 0.0
    if np.isscalar(t):
        t = (t, t, t)
    if isinstance(t, (list, tuple, np.ndarray)) and not isinstance(t, dict):
        d.type = 0
        d.color = _vec3(t)
        return d
    kind = t.get("type", "constant")
    d.type = TEX[kind]
    if "scale" in t:
        d.scale[0], d.scale[1] = float(t["scale"][0]), float(t["scale"][1])
    if kind == "constant":
        d.color = _vec3(t.get("color", 0.0))
    elif kind == "checkerboard":
        d.on = _vec3(t["on"])
        d.off = _vec3(t["off"])
    else:
        tex = t["texels"]
        h, w = tex.shape[:2]
        assert tex.shape == (h, w, 4), "bitmap texels must be (H, W, 4)"
        d.width, d.height = w, h
        if tex.dtype == np.uint8:
            arr = np.ascontiguousarray(tex)
            d.is_hdr = 0
            d.texels_u8 = _u8ptr(arr)
        else:
            arr = np.ascontiguousarray(tex, dtype=np.float32)
            d.is_hdr = 1
            d.texels_f32 = _fptr(arr)
        keep.append(arr)
    return d


def transform_desc(t):
    d = TransformDesc()
    t = t or {}
    d.translation = _vec3(t.get("translation"))
    d.scale = _vec3(t.get("scale"), (1.0, 1.0, 1.0))
    d.rotation = _vec3(t.get("rotation"))
    if t.get("lookat") is not None:
        d.has_lookat = 1
        d.lookat = _vec3(t["lookat"])
    if t.get("matrix") is not None:
        d.has_matrix = 1
        m = np.asarray(t["matrix"], dtype=np.float32).reshape(16)
        for i in range(16):
            d.matrix[i] = float(m[i])
    return d


def scene_desc(spec):
    """Return (SceneDesc, keepalive) for a scene spec dict."""
    keep = []
    meshes = spec.get("meshes", [])
    md = (MeshDesc * max(1, len(meshes)))()
    for i, m in enumerate(meshes):
        pos = np.ascontiguousarray(m["positions"], dtype=np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(m["indices"], dtype=np.int32).reshape(-1, 3)
        keep += [pos, idx]
        md[i].num_vertices = pos.shape[0]
        md[i].positions = _fptr(pos)
        md[i].num_triangles = idx.shape[0]
        md[i].indices = _iptr(idx)
        md[i].face_normals = int(m.get("face_normals", True))
        md[i].has_texcoords = int(m.get("has_texcoords", False))
        md[i].leaf_size = int(m.get("leaf_size", 4))
        if m.get("normals") is not None:
            n = np.ascontiguousarray(m["normals"], dtype=np.float32).reshape(-1, 3)
            keep.append(n)
            md[i].normals = _fptr(n)
        if m.get("texcoords") is not None:
            uv = np.ascontiguousarray(m["texcoords"], dtype=np.float32).reshape(-1, 2)
            keep.append(uv)
            md[i].texcoords = _fptr(uv)
    insts = spec.get("instances", [])
    idesc = (InstanceDesc * max(1, len(insts)))()
    for i, inst in enumerate(insts):
        idesc[i].mesh = int(inst["mesh"])
        idesc[i].material = int(inst["material"])
        idesc[i].transform = transform_desc(inst.get("transform"))
    sph = spec.get("spheres", [])
    sdesc = (SphereDesc * max(1, len(sph)))()
    for i, s in enumerate(sph):
        sdesc[i].center = _vec3(s["center"])
        sdesc[i].radius = float(s["radius"])
        sdesc[i].material = int(s["material"])
    mats = spec.get("materials", [])
    mdesc = (MaterialDesc * max(1, len(mats)))()
    for i, m in enumerate(mats):
        t = m["type"]
        mdesc[i].type = MAT[t] if isinstance(t, str) else int(t)
        mdesc[i].albedo = texture_desc(m.get("albedo", 0.0), keep)
        mdesc[i].roughness = texture_desc(m.get("roughness", 1e-4), keep)
        mdesc[i].ior = float(m.get("ior", 0.0))
        mdesc[i].medium = int(m.get("medium", -1))
        mdesc[i].eta = _vec3(m.get("eta"))
        mdesc[i].k = _vec3(m.get("k"))
        dist = m.get("distribution", "beckmann")
        mdesc[i].distribution = DIST[dist] if isinstance(dist, str) else int(dist)
    meds = spec.get("mediums", [])
    medesc = (MediumDesc * max(1, len(meds)))()
    for i, m in enumerate(meds):
        medesc[i].sigma_a = _vec3(m["sigma_a"])
        medesc[i].sigma_s = float(m["sigma_s"])
        medesc[i].g = float(m.get("g", 0.0))
    d = SceneDesc()
    d.num_meshes, d.meshes = len(meshes), md
    d.num_instances, d.instances = len(insts), idesc
    d.num_spheres, d.spheres = len(sph), sdesc
    d.num_materials, d.materials = len(mats), mdesc
    d.num_mediums, d.mediums = len(meds), medesc
    if spec.get("environment") is not None:
        d.has_environment = 1
        d.environment = texture_desc(spec["environment"], keep)
    cam = spec["camera"]
    d.camera.fov = float(cam["fov"])
    d.camera.dof = float(cam.get("dof", 1.0))
    d.camera.aperture = float(cam.get("aperture", 0.0))
    d.camera.transform = transform_desc(cam.get("transform"))
    keep += [md, idesc, sdesc, mdesc, medesc]
    return d, keep


def render_params(width, height, spp, max_depth, ray_bias=1e-4, tonemap="linear", first_sample=0):
    p = RenderParams()
    p.width, p.height, p.spp, p.max_depth = int(width), int(height), int(spp), int(max_depth)
    p.ray_bias = float(ray_bias)
    p.tonemap = TONEMAP[tonemap.lower()] if isinstance(tonemap, str) else int(tonemap)
    p.first_sample = int(first_sample)
    return p
