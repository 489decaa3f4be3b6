"""illume_amd - MI355X (gfx950) wavefront path tracer behind a C ABI (include/illume_c.h).

The compute path is libillume_amd.so (HIP kernels + C++ runtime). This module is a thin
ctypes layer over that ABI: it never renders anything itself, and it raises if the library is
missing. Build it with `python -m illume_amd.build`.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (BVH, TONEMAP, TRI, BvhInfo, MeshDesc, Opts, Pixel, RenderInfo, RenderParams, ReqStatus,
                   SceneDesc, Stats, render_params, scene_desc, struct_to_dict)

__all__ = ["lib", "Context", "Scene", "Request", "IlmError", "render_params", "bvh_build", "load_json"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libillume_amd.so")
_lib = None


class IlmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_abi.ERRORS.get(code, "?"), code, msg))
        self.code = code


def lib():
    """Load libillume_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libillume_amd.so not built; run `python -m illume_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ilm_version": (C.c_int, []),
        "ilm_global_error": (C.c_char_p, []),
        "ilm_open": (C.c_int, [C.c_int, P(Opts), P(vp)]),
        "ilm_close": (None, [vp]),
        "ilm_last_error": (C.c_char_p, [vp]),
        "ilm_stats_get": (C.c_int, [vp, P(Stats)]),
        "ilm_tune": (C.c_int, [vp, P(Opts)]),
        "ilm_scene_upload": (C.c_int, [vp, P(SceneDesc), P(vp)]),
        "ilm_scene_free": (None, [vp]),
        "ilm_scene_device_bytes": (C.c_int64, [vp]),
        "ilm_desc_load_json": (C.c_int, [C.c_char_p, P(vp)]),
        "ilm_desc_scene": (P(SceneDesc), [vp]),
        "ilm_desc_settings": (C.c_int, [vp, P(RenderParams)]),
        "ilm_desc_free": (None, [vp]),
        "ilm_desc_load_times": (C.c_int, [vp, P(C.c_double), P(C.c_double)]),
        "ilm_loader_error": (C.c_char_p, []),
        "ilm_render": (C.c_int, [vp, vp, P(RenderParams), P(C.c_float), P(Pixel), P(RenderInfo)]),
        "ilm_render_noalloc": (C.c_int, [vp, vp, P(RenderParams), P(C.c_float), P(Pixel), P(RenderInfo)]),
        "ilm_render_batch": (C.c_int, [vp, C.c_int32, P(vp), P(RenderParams), P(P(C.c_float)), P(P(Pixel)), P(RenderInfo)]),
        "ilm_submit": (C.c_int, [vp, vp, P(RenderParams), P(vp)]),
        "ilm_poll": (C.c_int, [vp, P(ReqStatus)]),
        "ilm_wait": (C.c_int, [vp, C.c_int32]),
        "ilm_result": (C.c_int, [vp, P(C.c_float), P(Pixel)]),
        "ilm_cancel": (C.c_int, [vp]),
        "ilm_req_free": (None, [vp]),
        "ilm_bvh_build": (C.c_int, [P(MeshDesc), P(vp)]),
        "ilm_bvh_get_info": (C.c_int, [vp, P(BvhInfo)]),
        "ilm_bvh_export": (C.c_int, [vp, P(C.c_float), P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
        "ilm_bvh_quantized_size": (C.c_int, [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
        "ilm_bvh_export_quantized": (C.c_int, [vp, P(C.c_float), P(C.c_float), P(C.c_int32)]),
        "ilm_bvh_free": (None, [vp]),
        "ilm_debug_closest_hit": (C.c_int, [vp, vp, P(C.c_float), C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_float)]),
        "ilm_debug_path_radiance": (C.c_int, [vp, vp, P(RenderParams), P(C.c_uint32), C.c_int32, P(C.c_float), P(C.c_int32)]),
        "ilm_debug_instance_matrices": (C.c_int, [vp, C.c_int32, P(C.c_float)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _pxp(a):
    """(n, 3) uint64 array as ilm_pixel* (same layout: three size_t per pixel), or None."""
    return None if a is None else a.ctypes.data_as(C.POINTER(Pixel))


def _check(rc, ctx=None):
    if rc != 0:
        L = lib()
        msg = (L.ilm_last_error(ctx) if ctx else L.ilm_global_error()) or b""
        raise IlmError(rc, msg.decode(errors="replace"))


def _pixels_np(n):
    return np.zeros((n, 3), dtype=np.uint64)


class Context:
    """One GPU. opts: bvh_layout ('fp32'|'quant8'), tri_test ('mt'|'watertight'),
    paths_in_flight, hbm_budget_bytes, use_graphs, max_queued, profile, count_tests."""

    def __init__(self, device=0, **opts):
        L = lib()
        o = Opts()
        o.use_graphs = 1
        for k, v in opts.items():
            if k == "bvh_cache_dir":
                setattr(o, k, None if v is None else os.fsencode(v))
                continue
            if k == "bvh_layout" and isinstance(v, str):
                v = BVH[v]
            if k == "tri_test" and isinstance(v, str):
                v = TRI[v]
            setattr(o, k, int(v))
        h = C.c_void_p()
        _check(L.ilm_open(int(device), C.byref(o), C.byref(h)))
        self.h = h
        self.opts = o

    TUNABLE = ("bvh_layout", "extend_mode", "refill_min", "ext_fetch", "leaf_wait_pct", "shade_sort", "sched_fair",
               "camera_packets",
               "sample_group", "tile_rows", "chunk_passes", "paths_in_flight", "lone_split")

    def tune(self, **kw):
        """Change the tunable options (TUNABLE) of an idle context; others keep their values."""
        o = Opts()
        C.memmove(C.byref(o), C.byref(self.opts), C.sizeof(Opts))
        for k, v in kw.items():
            if k not in self.TUNABLE:
                raise ValueError("not tunable: %s" % k)
            if k == "bvh_layout" and isinstance(v, str):
                v = BVH[v]
            setattr(o, k, int(v))
        _check(lib().ilm_tune(self.h, C.byref(o)), self.h)
        self.opts = o

    def settings(self):
        return {k: getattr(self.opts, k) for k in self.TUNABLE}

    def close(self):
        if self.h:
            lib().ilm_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scene(self, spec):
        return Scene(self, spec)

    def render(self, scene, params, want_accum=True, want_pixels=True):
        """Plain synchronous render. Returns (accum (H,W,3) f32 sum over spp, pixels (H,W,3) u64, info)."""
        W, H = params.width, params.height
        acc = np.empty((H * W, 3), dtype=np.float32) if want_accum else None
        pix = np.empty((H * W, 3), dtype=np.uint64) if want_pixels else None
        info = RenderInfo()
        _check(lib().ilm_render(self.h, scene.h, C.byref(params), _fp(acc) if acc is not None else None,
                                _pxp(pix), C.byref(info)), self.h)
        return (acc.reshape(H, W, 3) if acc is not None else None), \
            (pix.reshape(H, W, 3) if pix is not None else None), struct_to_dict(info)

    def render_noalloc(self, scene, params, want_accum=True, want_pixels=True):
        """ilm_render_noalloc: the plain render with every device buffer carved from the library's
        static scratch array (no device allocation). Same return values as render()."""
        W, H = params.width, params.height
        acc = np.empty((H * W, 3), dtype=np.float32) if want_accum else None
        pix = np.empty((H * W, 3), dtype=np.uint64) if want_pixels else None
        info = RenderInfo()
        _check(lib().ilm_render_noalloc(self.h, scene.h, C.byref(params), _fp(acc) if acc is not None else None,
                                        _pxp(pix), C.byref(info)), self.h)
        return (acc.reshape(H, W, 3) if acc is not None else None), \
            (pix.reshape(H, W, 3) if pix is not None else None), struct_to_dict(info)

    def render_batch(self, scenes, params, want_accum=True):
        n = len(scenes)
        hs = (C.c_void_p * n)(*[s.h.value for s in scenes])
        ps = (RenderParams * n)(*params)
        accs = [np.empty((p.height * p.width, 3), dtype=np.float32) for p in params]
        pxs = [np.empty((p.height * p.width, 3), dtype=np.uint64) for p in params]
        acc_ptrs = (C.POINTER(C.c_float) * n)(*[_fp(a) for a in accs])
        px_ptrs = (C.POINTER(Pixel) * n)(*[_pxp(p) for p in pxs])
        info = RenderInfo()
        _check(lib().ilm_render_batch(self.h, n, hs, ps, acc_ptrs if want_accum else None, px_ptrs, C.byref(info)), self.h)
        out = []
        for p, a, px in zip(params, accs, pxs):
            out.append((a.reshape(p.height, p.width, 3), px.reshape(p.height, p.width, 3)))
        return out, struct_to_dict(info)

    def submit(self, scene, params):
        return Request(self, scene, params)

    def stats(self):
        s = Stats()
        _check(lib().ilm_stats_get(self.h, C.byref(s)), self.h)
        return struct_to_dict(s)

    def closest_hit(self, scene, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = rays.shape[0]
        inst = np.empty(n, np.int32)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        _check(lib().ilm_debug_closest_hit(self.h, scene.h, _fp(rays), n, _ip(inst), _ip(prim), _fp(t)), self.h)
        return inst, prim, t

    def path_radiance(self, scene, params, pixel_pass):
        pp = np.ascontiguousarray(pixel_pass, dtype=np.uint32).reshape(-1, 2)
        n = pp.shape[0]
        rgb = np.empty((n, 3), np.float32)
        seg = np.empty(n, np.int32)
        _check(lib().ilm_debug_path_radiance(self.h, scene.h, C.byref(params),
                                             pp.ctypes.data_as(C.POINTER(C.c_uint32)), n, _fp(rgb), _ip(seg)), self.h)
        return rgb, seg


class Scene:
    """A device-resident scene uploaded from a spec dict (or a loaded JSON descriptor)."""

    def __init__(self, ctx, spec=None, desc=None):
        self.ctx = ctx
        h = C.c_void_p()
        if desc is None:
            desc, keep = scene_desc(spec)
        _check(lib().ilm_scene_upload(ctx.h, C.byref(desc), C.byref(h)), ctx.h)
        self.h = h

    @property
    def device_bytes(self):
        return lib().ilm_scene_device_bytes(self.h)

    def instance_matrices(self, i):
        out = np.empty(48, np.float32)
        _check(lib().ilm_debug_instance_matrices(self.h, int(i), _fp(out)))
        return out.reshape(3, 4, 4)

    def free(self):
        if self.h:
            lib().ilm_scene_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Request:
    def __init__(self, ctx, scene, params):
        self.ctx, self.params = ctx, params
        h = C.c_void_p()
        _check(lib().ilm_submit(ctx.h, scene.h, C.byref(params), C.byref(h)), ctx.h)
        self.h = h

    def poll(self):
        s = ReqStatus()
        _check(lib().ilm_poll(self.h, C.byref(s)))
        d = struct_to_dict(s)
        d["state"] = _abi.REQ_STATE[d["state"]]
        return d

    def wait(self, timeout_ms=-1):
        _check(lib().ilm_wait(self.h, int(timeout_ms)), self.ctx.h)

    def result(self):
        p = self.params
        acc = np.empty((p.height * p.width, 3), dtype=np.float32)
        px = np.empty((p.height * p.width, 3), dtype=np.uint64)
        _check(lib().ilm_result(self.h, _fp(acc), _pxp(px)), self.ctx.h)
        return acc.reshape(p.height, p.width, 3), px.reshape(p.height, p.width, 3)

    def cancel(self):
        _check(lib().ilm_cancel(self.h))

    def free(self):
        if self.h:
            lib().ilm_req_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bvh_device_layouts(positions, indices, cache_dir=None):
    """The device BVH layouts ilm_scene_upload uses for a mesh, through the on-disk cache in
    cache_dir. Returns (nodes bytes, qnodes bytes, leaf order, (root, qroot, depth, qdepth),
    from_cache)."""
    L = lib()
    f = L.ilm_bvh_device_layouts
    f.restype = C.c_int
    f.argtypes = [C.POINTER(MeshDesc), C.c_char_p, C.POINTER(C.c_int32), C.c_void_p, C.POINTER(C.c_int64),
                  C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
    m = MeshDesc()
    m.num_vertices, m.positions = pos.shape[0], _fp(pos)
    m.num_triangles, m.indices = idx.shape[0], _ip(idx)
    m.face_normals = 1
    d = os.fsencode(cache_dir) if cache_dir else None
    nb, qb = C.c_int64(0), C.c_int64(0)
    hit = C.c_int32(0)
    _check(f(C.byref(m), None, None, None, C.byref(nb), None, C.byref(qb), None, None))   # sizes
    nodes = np.zeros(nb.value, np.uint8)
    qnodes = np.zeros(qb.value, np.uint8)
    order = np.zeros(idx.shape[0], np.int32)
    info = np.zeros(4, np.int32)
    _check(f(C.byref(m), d, C.byref(hit), nodes.ctypes.data, C.byref(nb), qnodes.ctypes.data, C.byref(qb),
             _ip(order), _ip(info)))
    return nodes, qnodes, order, tuple(int(x) for x in info), bool(hit.value)


def bvh_build(positions, indices):
    """Build the product BVH on the host; returns a dict of numpy arrays (tree + quantised layout)."""
    L = lib()
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.int32).reshape(-1, 3)
    m = MeshDesc()
    m.num_vertices, m.positions = pos.shape[0], _fp(pos)
    m.num_triangles, m.indices = idx.shape[0], _ip(idx)
    m.face_normals, m.leaf_size = 1, 4
    h = C.c_void_p()
    _check(L.ilm_bvh_build(C.byref(m), C.byref(h)))
    try:
        info = BvhInfo()
        L.ilm_bvh_get_info(h, C.byref(info))
        n = info.num_nodes
        boxes = np.empty((n, 6), np.float32)
        children = np.empty((n, 2), np.int32)
        first = np.empty(n, np.int32)
        count = np.empty(n, np.int32)
        order = np.empty(idx.shape[0], np.int32)
        L.ilm_bvh_export(h, _fp(boxes), _ip(children), _ip(first), _ip(count), _ip(order))
        nq, root, qd = C.c_int32(), C.c_int32(), C.c_int32()
        L.ilm_bvh_quantized_size(h, C.byref(nq), C.byref(root), C.byref(qd))
        qb = np.empty((max(nq.value, 1), 4, 6), np.float32)
        ex = np.empty((max(nq.value, 1), 4, 6), np.float32)
        ch = np.empty((max(nq.value, 1), 4), np.int32)
        L.ilm_bvh_export_quantized(h, _fp(qb), _fp(ex), _ip(ch))
        return dict(info=struct_to_dict(info), boxes=boxes, children=children, leaf_first=first, leaf_count=count,
                    order=order, q_boxes=qb[:nq.value], q_exact=ex[:nq.value], q_child=ch[:nq.value],
                    q_root=root.value, q_depth=qd.value)
    finally:
        L.ilm_bvh_free(h)


class JsonScene:
    """A reference-schema scene JSON loaded by the C++ loader (host side)."""

    def __init__(self, path):
        L = lib()
        h = C.c_void_p()
        rc = L.ilm_desc_load_json(path.encode(), C.byref(h))
        if rc != 0:
            raise IlmError(rc, (L.ilm_loader_error() or b"").decode())
        self.h = h

    @property
    def desc(self):
        d = lib().ilm_desc_scene(self.h).contents
        d._owner = self   # the struct points into this loader's memory: keep it alive
        return d

    def settings(self):
        p = RenderParams()
        _check(lib().ilm_desc_settings(self.h, C.byref(p)))
        return p

    def upload(self, ctx):
        return Scene(ctx, desc=self.desc)

    def load_times(self):
        """(parse seconds, merge seconds) the load spent on OBJ files."""
        a, b = C.c_double(), C.c_double()
        _check(lib().ilm_desc_load_times(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def free(self):
        if self.h:
            lib().ilm_desc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def load_json(path):
    return JsonScene(path)
