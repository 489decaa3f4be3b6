"""Edge cases on the GPU, checked against the oracle (brute force for closest hits): empty and
sphere-only scenes, degenerate triangles, axis-parallel and grazing rays, rays through vertices
and edges, ragged image sizes, one-sample and one-bounce renders, pixel-range chunking."""
import numpy as np
import pytest

import illume_amd
import illume_amd.scenes as S
from illume_amd._abi import render_params
from oracle.pyoracle import OracleScene
from conftest import compare_hits

pytestmark = pytest.mark.gpu


def _p(st, **kw):
    d = dict(st)
    d.update(kw)
    return render_params(**d)


def _env_only():
    spec = dict(meshes=[], instances=[], spheres=[], materials=[dict(type="lambert", albedo=[0.5, 0.5, 0.5])],
                mediums=[], environment=[0.3, 0.6, 0.9],
                camera=dict(fov=45.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 0, -3])))
    return spec, dict(width=17, height=9, spp=3, max_depth=4, ray_bias=1e-4, tonemap="linear")


def _spheres_only():
    spec = dict(meshes=[], instances=[],
                spheres=[dict(center=[0, 0, 0], radius=1.0, material=0), dict(center=[1.5, 0, 0.5], radius=0.5, material=1)],
                materials=[dict(type="lambert", albedo=[0.7, 0.2, 0.2]), dict(type="emissive", albedo=[3, 3, 3])],
                mediums=[], environment=[0.5, 0.5, 0.5],
                camera=dict(fov=50.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 0, -4])))
    return spec, dict(width=40, height=30, spp=4, max_depth=6, ray_bias=1e-4, tonemap="linear")


def _degenerate_mesh_scene():
    """A cube whose triangle list also holds zero-area triangles (repeated and collinear
    vertices), seen axis-aligned so many rays run parallel to faces."""
    cube = S.cube_mesh()
    pos = np.asarray(cube["positions"], np.float32)
    idx = np.asarray(cube["indices"], np.int32)
    extra_v = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25], [0.5, 0.5, 0.5]], np.float32)
    n0 = pos.shape[0]
    pos = np.concatenate([pos, extra_v])
    deg = np.array([[0, 0, 1], [n0, n0 + 1, n0 + 2], [2, 2, 2]], np.int32)
    idx = np.concatenate([idx, deg])
    mesh = dict(positions=pos, indices=idx, face_normals=True)
    spec = dict(meshes=[mesh], instances=[dict(mesh=0, material=0, transform=dict(translation=[0, 0, 0]))],
                spheres=[], materials=[dict(type="lambert", albedo=[0.8, 0.8, 0.8])], mediums=[],
                environment=[1.0, 1.0, 1.0],
                camera=dict(fov=40.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 0, -5])))
    return spec, dict(width=32, height=32, spp=2, max_depth=3, ray_bias=1e-4, tonemap="linear")


def _adversarial_rays(pos, idx, n=4000, seed=5):
    """Rays aimed at vertices, edge midpoints and face centres from random origins, plus
    axis-parallel and grazing rays."""
    rng = np.random.default_rng(seed)
    tri = pos[idx]                          # (T, 3, 3)
    targets = np.concatenate([tri.reshape(-1, 3),
                              0.5 * (tri[:, 0] + tri[:, 1]), 0.5 * (tri[:, 1] + tri[:, 2]),
                              tri.mean(axis=1)])
    t = targets[rng.integers(0, len(targets), n)]
    o = rng.uniform(-4, 4, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = [np.concatenate([o, d], 1)]
    # axis-parallel: origins on a grid, directions exactly +-x, +-y, +-z
    m = n // 4
    o2 = rng.uniform(-1.2, 1.2, (m, 3))
    axis = rng.integers(0, 3, m)
    sign = rng.choice([-1.0, 1.0], m)
    o2[np.arange(m), axis] = -3.0 * sign
    d2 = np.zeros((m, 3))
    d2[np.arange(m), axis] = sign
    rays.append(np.concatenate([o2, d2], 1))
    # grazing: nearly parallel to a face plane, through the cube
    o3 = rng.uniform(-0.9, 0.9, (m, 3))
    o3[:, 2] = -3.0
    d3 = np.stack([rng.uniform(-1e-4, 1e-4, m), rng.uniform(-1e-4, 1e-4, m), np.ones(m)], 1)
    d3 /= np.linalg.norm(d3, axis=1, keepdims=True)
    rays.append(np.concatenate([o3, d3], 1))
    return np.concatenate(rays).astype(np.float32)


def test_closest_hit_degenerate_and_axis_parallel(ctx):
    spec, st = _degenerate_mesh_scene()
    sc = ctx.scene(spec)
    orc = OracleScene(spec)
    mesh = spec["meshes"][0]
    rays = _adversarial_rays(np.asarray(mesh["positions"], np.float64), np.asarray(mesh["indices"]))
    g = ctx.closest_hit(sc, rays)
    o = orc.closest_hit(rays, brute=True)
    exact, tie, bad = compare_hits(g, o)
    print("adversarial rays: %d, same %d, ties %d, mismatches %d, hits %d"
          % (len(rays), exact, tie, len(bad), int((o[0] != -1).sum())))
    for i in bad[:5]:
        print("  ray", i, rays[i], "gpu", g[0][i], g[1][i], g[2][i], "oracle", o[0][i], o[1][i], o[2][i])
    assert len(bad) == 0
    sc.free()
    orc.free()


@pytest.mark.parametrize("maker", [_env_only, _spheres_only, _degenerate_mesh_scene])
def test_small_scenes_render_like_oracle(ctx, maker):
    spec, st = maker()
    sc = ctx.scene(spec)
    orc = OracleScene(spec)
    p = _p(st)
    acc, px, info = ctx.render(sc, p)
    oacc, opx, osegs, _ = orc.render(p, kahan=True)
    oacc = oacc.reshape(acc.shape)
    rel = np.abs(acc - oacc) / np.maximum(np.abs(oacc), 1e-6)
    print("%s: segments gpu %d oracle %d, max rel %.3g" % (maker.__name__, info["segments"], osegs, rel.max()))
    assert info["segments"] == osegs
    assert np.mean(rel <= 1e-4) >= 0.999
    sc.free()
    orc.free()


@pytest.mark.parametrize("w,h,spp,depth", [(1, 1, 1, 1), (1, 7, 3, 2), (13, 1, 2, 5), (33, 19, 1, 1)])
def test_ragged_sizes_and_one_bounce(ctx, w, h, spp, depth):
    spec, st = S.cornell()
    sc = ctx.scene(spec)
    orc = OracleScene(spec)
    p = _p(st, width=w, height=h, spp=spp, max_depth=depth)
    acc, px, info = ctx.render(sc, p)
    oacc, opx, osegs, _ = orc.render(p, kahan=True)
    assert info["segments"] == osegs
    rel = np.abs(acc.reshape(-1, 3) - oacc) / np.maximum(np.abs(oacc), 1e-6)
    assert np.mean(rel <= 1e-4) >= 0.999
    sc.free()
    orc.free()


def test_pixel_range_chunks_equal_whole_image(ctx):
    """A chunk smaller than one pass splits each pass into pixel ranges; results are identical."""
    spec, st = S.tiny_mesh_scene()
    p = _p(st, width=150, height=110, spp=3)
    sc = ctx.scene(spec)
    ref, refpx, ri = ctx.render(sc, p)
    with illume_amd.Context(0, paths_in_flight=4096) as c:
        s2 = c.scene(spec)
        got, px, i = c.render(s2, p)
        assert i["chunks"] > 3 * (150 * 110 // 4096)
        assert np.array_equal(got, ref) and np.array_equal(px, refpx)
        s2.free()
    sc.free()


def test_bvh_cache_upload_hits(ctx, tmp_path):
    spec, st = S.gold_spheres(n=20)
    with illume_amd.Context(0, bvh_cache_dir=str(tmp_path)) as c:
        s1 = c.scene(spec)
        a, _, _ = c.render(s1, _p(st, width=48, height=48, spp=2))
        s2 = c.scene(spec)
        b, _, _ = c.render(s2, _p(st, width=48, height=48, spp=2))
        stats = c.stats()
        assert stats["bvh_cache_misses"] == len(spec["meshes"]) and stats["bvh_cache_hits"] == len(spec["meshes"])
        assert np.array_equal(a, b)
        s1.free()
        s2.free()


def test_scene_extent_beyond_8bit_quantum_refused():
    """8-bit child boxes have a quantum of at most 2^24 units, so that the node-level product
    quantum x 1/d stays finite for any direction: a mesh spread over 2e10 units (root quantum
    2^27) is refused in that layout at upload and after a live switch to it, and renders in the
    FP32 layout before and after."""
    xs = np.linspace(-1e10, 1e10, 16)
    pos, idx = [], []
    for i, x in enumerate(xs):
        pos += [[x, -1.0, 5.0], [x + 1e6, -1.0, 5.0], [x, 1.0, 5.0]]
        idx.append([3 * i, 3 * i + 1, 3 * i + 2])
    mesh = dict(positions=np.array(pos, np.float32), indices=np.array(idx, np.int32), face_normals=True)
    spec = dict(meshes=[mesh], instances=[dict(mesh=0, material=0, transform=dict(translation=[0, 0, 0]))],
                spheres=[], materials=[dict(type="lambert", albedo=[0.5, 0.5, 0.5])], mediums=[],
                environment=[1.0, 1.0, 1.0],
                camera=dict(fov=40.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 0, -5])))
    p = render_params(width=16, height=16, spp=1, max_depth=2, ray_bias=1e-4, tonemap="linear")
    with illume_amd.Context(0, bvh_layout="quant8") as c:
        with pytest.raises(illume_amd.IlmError) as e:
            c.scene(spec)
        assert e.value.code == -9
    with illume_amd.Context(0) as c:
        s = c.scene(spec)
        a, _, _ = c.render(s, p)
        c.tune(bvh_layout="quant8")
        with pytest.raises(illume_amd.IlmError) as e:
            c.render(s, p)
        assert e.value.code == -9
        c.tune(bvh_layout="fp32")
        b, _, _ = c.render(s, p)
        s.free()
    assert np.array_equal(a, b)


def _translated(spec, T):
    """The scene moved by T: every instance and the camera translated (object spaces unchanged)."""
    out = dict(spec)
    out["instances"] = []
    for inst in spec["instances"]:
        tr = dict(inst["transform"])
        tr["translation"] = list(np.asarray(tr.get("translation", [0, 0, 0]), np.float64) + T)
        out["instances"].append(dict(inst, transform=tr))
    cam = dict(spec["camera"])
    ct = dict(cam["transform"])
    ct["translation"] = list(np.asarray(ct.get("translation", [0, 0, 0]), np.float64) + T)
    out["camera"] = dict(cam, transform=ct)
    return out


@pytest.mark.parametrize("offset", [1e3, 1e4])
@pytest.mark.parametrize("layout", ["fp32", "quant8"])
def test_closest_hit_far_from_the_world_origin(offset, layout):
    """The Cornell box moved 1e3 and 1e4 from the world origin (the box tests form slab distances
    from differences to the ray origin, so their error does not grow with the distance from the
    world origin): camera rays and rays starting inside the box, against brute force over every
    triangle of every instance."""
    from conftest import camera_rays, random_rays
    spec, st = S.cornell()
    T = np.array([offset, 0.37 * offset, -0.61 * offset])
    spec = _translated(spec, T)
    rays = np.concatenate([camera_rays(spec, st, 20000, seed=8),
                           random_rays(T + [-0.45, 0.05, -0.45], T + [0.45, 0.95, 0.45], 20000, seed=9)])
    with illume_amd.Context(0, bvh_layout=layout) as c:
        sc = c.scene(spec)
        g = c.closest_hit(sc, rays)
        sc.free()
    orc = OracleScene(spec)
    o = orc.closest_hit(rays, brute=True)
    orc.free()
    exact, tie, bad = compare_hits(g, o)
    print("offset %g [%s]: %d rays, same %d, ties %d, mismatches %d, hits %d"
          % (offset, layout, len(rays), exact, tie, len(bad), int((o[0] != -1).sum())))
    for i in bad[:5]:
        print("  ray", i, rays[i], "gpu", g[0][i], g[1][i], g[2][i], "oracle", o[0][i], o[1][i], o[2][i])
    assert len(bad) == 0
    assert (o[0] != -1).sum() > 30000
