"""Synthetic scenes of the benchmark configurations (all seeded, generated in memory).

    cornell()            256x256, 16 spp, depth 8: 5 wall quads + ceiling light + 2 boxes
    gold_spheres()       1024x1024, 64 spp, depth 16: displaced icosphere (n=71, 100,820 tris)
                         as rough gold (GGX a=0.3, eta+ik given), rough dielectric sphere
                         (Beckmann roughness 0.2, IOR 1.5) on a checkerboard plane
    subsurface()         960x720, 32 spp, depth 45: displaced icosphere (n=162, 524,880 tris)
                         holding a medium (sigma_a=(0.05,0.4,1.2), sigma_s=2, g=0.7) inside a
                         smooth dielectric (IOR 1.5), 2048x1024 FP32 sky + 0.5 deg sun (5e4)
    instanced()          12 instances of a displaced icosphere (n=190, 722,000 tris) on a 4x3 grid

Geometry is built from first principles: a geodesic icosphere with 20*n^2 triangles
(frequency-n subdivision of the icosahedron, vertices shared exactly between faces) whose
vertices are pushed out radially by 3-octave value noise (lattice values from an integer hash
of (x, y, z, seed), smoothstep-trilinear interpolation, octaves f, 2f, 4f with weights 1, 1/2,
1/4 normalised), amplitude 0.1.
Each function returns (spec, settings): spec for illume_amd._abi.scene_desc, settings a dict
with width, height, spp, max_depth, ray_bias, tonemap.
"""
import math

import numpy as np


def quad_mesh():
    """Unit square in the z=0 plane, two triangles, texcoords 0..1."""
    pos = np.array([[-0.5, 0.5, 0.0], [0.5, 0.5, 0.0], [0.5, -0.5, 0.0], [-0.5, -0.5, 0.0]], np.float32)
    idx = np.array([[0, 1, 2], [2, 3, 0]], np.int32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return dict(positions=pos, indices=idx, texcoords=uv, has_texcoords=True, face_normals=True, leaf_size=4)


def cube_mesh():
    """Unit cube centred at the origin, 12 triangles, outward winding."""
    p = np.array([[x, y, z] for z in (-0.5, 0.5) for y in (-0.5, 0.5) for x in (-0.5, 0.5)], np.float32)
    faces = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    idx = []
    for a, b, c, d in faces:
        idx += [[a, b, c], [a, c, d]]
    return dict(positions=p, indices=np.array(idx, np.int32), face_normals=True, leaf_size=12)


def icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def icosphere(n):
    """Geodesic sphere with 20*n*n triangles. Vertices on shared edges are identical: a vertex
    is keyed by its barycentric weights over the icosahedron's corner indices."""
    corners, faces = icosahedron()
    keys = {}
    verts = []
    tris = []

    def vid(wmap):
        key = tuple(sorted((c, w) for c, w in wmap.items() if w))
        if key not in keys:
            p = np.zeros(3)
            for c, w in key:
                p += corners[c] * (w / n)
            keys[key] = len(verts)
            verts.append(p / np.linalg.norm(p))
        return keys[key]

    for a, b, c in faces:
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                k = n - i - j
                wm = {}
                for cc, w in ((a, k), (b, i), (c, j)):
                    wm[cc] = wm.get(cc, 0) + w
                grid[(i, j)] = vid(wm)
        for i in range(n):
            for j in range(n - i):
                tris.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j < n - 1:
                    tris.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    return np.array(verts, np.float64), np.array(tris, np.int32)


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iy.astype(np.uint32) * np.uint32(0xD8163841)) \
        ^ (iz.astype(np.uint32) * np.uint32(0xCB1AB31F)) ^ np.uint32((seed * 0x9E3779B9) & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7FEB352D)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846CA68B)
    h ^= h >> np.uint32(16)
    return h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0


def value_noise(p, seed):
    """Smoothstep-trilinear value noise in [-1, 1] at points p (N,3)."""
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    s = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0])
    with np.errstate(over="ignore"):
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1.0 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1.0 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1.0 - s[:, 2]
                    out += wx * wy * wz * _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def displaced_icosphere(n, seed, amplitude=0.1, freq=3.0):
    v, t = icosphere(n)
    noise = (value_noise(v * freq, seed) + 0.5 * value_noise(v * 2 * freq, seed + 101)
             + 0.25 * value_noise(v * 4 * freq, seed + 202)) / 1.75
    v = v * (1.0 + amplitude * noise)[:, None]
    return dict(positions=v.astype(np.float32), indices=t, face_normals=True, leaf_size=4)


def sky_env(width=2048, height=1024, sun_elev_deg=35.0, sun_azim_deg=60.0, sun_radiance=5e4, sun_diam_deg=0.5):
    """Equirect FP32 env in the renderer's (u, v) convention (u = atan2(z, x)/2pi + 1/2,
    v = 1/2 - asin(y)/pi): sky gradient, dark ground, sun disk of the given angular diameter."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi = (u - 0.5) * 2.0 * np.pi
    lat = (0.5 - v) * np.pi
    cy = np.sin(lat)[:, None]
    cl = np.cos(lat)[:, None]
    d = np.stack([np.broadcast_to(cl * np.cos(phi)[None, :], (height, width)),
                  np.broadcast_to(cy, (height, width)),
                  np.broadcast_to(cl * np.sin(phi)[None, :], (height, width))], axis=-1)
    zen = np.array([0.25, 0.45, 1.0])
    hor = np.array([1.0, 0.95, 0.9])
    gnd = np.array([0.18, 0.16, 0.14])
    y = d[..., 1:2]
    a = np.clip(y, 0.0, 1.0) ** 0.5
    sky = np.where(y >= 0, hor * (1 - a) + zen * a, gnd)
    el, az = math.radians(sun_elev_deg), math.radians(sun_azim_deg)
    sd = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    cosang = (d @ sd)
    sun = cosang >= math.cos(math.radians(sun_diam_deg / 2.0))
    rgb = np.where(sun[..., None], sun_radiance, sky)
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = rgb
    return out


def _settings(w, h, spp, depth, tonemap="linear", ray_bias=1e-4):
    return dict(width=w, height=h, spp=spp, max_depth=depth, ray_bias=ray_bias, tonemap=tonemap)


def cornell():
    white, red, green, light, box = 0, 1, 2, 3, 4
    mats = [dict(type="lambert", albedo=[0.73, 0.73, 0.73]), dict(type="lambert", albedo=[0.65, 0.05, 0.05]),
            dict(type="lambert", albedo=[0.12, 0.45, 0.15]), dict(type="emissive", albedo=[17.0, 12.0, 4.0]),
            dict(type="lambert", albedo=[0.73, 0.73, 0.73])]
    Q, B = 0, 1
    inst = [
        dict(mesh=Q, material=white, transform=dict(translation=[0, 0, 0], rotation=[90, 0, 0])),      # floor
        dict(mesh=Q, material=white, transform=dict(translation=[0, 1, 0], rotation=[90, 0, 0])),      # ceiling
        dict(mesh=Q, material=white, transform=dict(translation=[0, 0.5, 0.5])),                       # back
        dict(mesh=Q, material=red, transform=dict(translation=[-0.5, 0.5, 0], rotation=[0, 90, 0])),   # left
        dict(mesh=Q, material=green, transform=dict(translation=[0.5, 0.5, 0], rotation=[0, 90, 0])),  # right
        dict(mesh=Q, material=light, transform=dict(translation=[0, 0.998, 0], scale=[0.25, 0.25, 1], rotation=[90, 0, 0])),
        dict(mesh=B, material=box, transform=dict(translation=[0.18, 0.15, -0.12], scale=[0.3, 0.3, 0.3], rotation=[0, 30, 0])),
        dict(mesh=B, material=box, transform=dict(translation=[-0.17, 0.3, 0.15], scale=[0.3, 0.6, 0.3], rotation=[0, -30, 0])),
    ]
    spec = dict(meshes=[quad_mesh(), cube_mesh()], instances=inst, spheres=[], materials=mats, mediums=[],
                environment=None,
                camera=dict(fov=40.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 0.5, -1.9])))
    return spec, _settings(256, 256, 16, 8)


def gold_spheres(n=71, seed=1):
    gold = dict(type="roughconductor", distribution="ggx", roughness=0.3,
                eta=[0.18, 0.42, 1.37], k=[3.42, 2.35, 1.77])
    glass = dict(type="roughrefractive", roughness=0.2, ior=1.5, albedo=[1.0, 1.0, 1.0])
    floor = dict(type="lambert", albedo=dict(type="checkerboard", on=[0.8, 0.8, 0.8], off=[0.15, 0.15, 0.15], scale=[20, 20]))
    mesh = displaced_icosphere(n, seed)
    inst = [
        dict(mesh=0, material=0, transform=dict(translation=[-1.15, 1.0, 0.0])),
        dict(mesh=1, material=2, transform=dict(translation=[0, 0, 0], scale=[20, 20, 1], rotation=[90, 0, 0])),
    ]
    spec = dict(meshes=[mesh, quad_mesh()], instances=inst,
                spheres=[dict(center=[1.25, 0.8, 0.3], radius=0.8, material=1)],
                materials=[gold, glass, floor], mediums=[], environment=[0.9, 0.95, 1.0],
                camera=dict(fov=40.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 1.7, -5.5], lookat=[0, 0.8, 0])))
    return spec, _settings(1024, 1024, 64, 16)


def subsurface(n=162, seed=2, env_size=(2048, 1024)):
    mesh = displaced_icosphere(n, seed)
    mat = dict(type="refractive", ior=1.5, albedo=[1.0, 1.0, 1.0], medium=0)
    medium = dict(sigma_a=[0.05, 0.4, 1.2], sigma_s=2.0, g=0.7)
    spec = dict(meshes=[mesh], instances=[dict(mesh=0, material=0, transform=dict(translation=[0, 0, 0]))],
                spheres=[], materials=[mat], mediums=[medium],
                environment=dict(type="bitmap", texels=sky_env(*env_size)),
                camera=dict(fov=40.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 0.6, -4.0], lookat=[0, 0, 0])))
    return spec, _settings(960, 720, 32, 45, tonemap="filmic")


def instanced(n=190, seed=5):
    mesh = displaced_icosphere(n, seed)
    mats = [dict(type="lambert", albedo=[0.7, 0.3, 0.2]), dict(type="plastic", albedo=[0.2, 0.4, 0.8], ior=1.5),
            dict(type="roughconductor", roughness=0.25, eta=[0.14311, 0.37495, 1.44247], k=[3.9831, 2.3857, 1.6032]),
            dict(type="lambert", albedo=dict(type="checkerboard", on=[0.75, 0.75, 0.75], off=[0.2, 0.2, 0.2], scale=[30, 30]))]
    inst = []
    for r in range(3):
        for c in range(4):
            inst.append(dict(mesh=0, material=(r * 4 + c) % 3,
                             transform=dict(translation=[(c - 1.5) * 2.3, 1.0, r * 2.3], rotation=[0, 30 * (r * 4 + c), 0])))
    inst.append(dict(mesh=1, material=3, transform=dict(translation=[0, 0, 2.3], scale=[30, 30, 1], rotation=[90, 0, 0])))
    spec = dict(meshes=[mesh, quad_mesh()], instances=inst, spheres=[], materials=mats, mediums=[],
                environment=[0.8, 0.85, 1.0],
                camera=dict(fov=50.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 4.5, -7.0], lookat=[0, 0.5, 2.3])))
    return spec, _settings(1280, 720, 64, 16)


def tiny_mesh_scene(n=4, seed=7):
    """Small scene exercising every material type, a medium, texcoords and smooth normals."""
    v, t = icosphere(n)
    sm = dict(positions=v.astype(np.float32), indices=t, normals=v.astype(np.float32), face_normals=False, leaf_size=4)
    mats = [dict(type="lambert", albedo=[0.7, 0.7, 0.7]),
            dict(type="emissive", albedo=[4.0, 4.0, 4.0]),
            dict(type="reflective", albedo=[1, 1, 1]),
            dict(type="refractive", ior=1.4, albedo=[0.95, 0.95, 0.95], medium=0),
            dict(type="roughreflective", ior=1.5, roughness=0.3, albedo=[0.9, 0.8, 0.7]),
            dict(type="roughrefractive", ior=1.5, roughness=0.2, albedo=[1, 1, 1], medium=0),
            dict(type="conductor", eta=[0.21249, 0.97909, 1.3343], k=[4.1005, 2.3691, 2.2965]),
            dict(type="roughconductor", roughness=0.2, eta=[0.16761, 0.14032, 0.13529], k=[4.2868, 2.8794, 2.2916]),
            dict(type="plastic", ior=1.5, albedo=[0.3, 0.6, 0.3]),
            dict(type="lambert", albedo=dict(type="checkerboard", on=[0.9, 0.9, 0.9], off=[0.1, 0.1, 0.1], scale=[8, 8])),
            dict(type="roughconductor", distribution="ggx", roughness=0.3, eta=[0.18, 0.42, 1.37], k=[3.42, 2.35, 1.77])]
    inst = []
    pos = [(-2.4, 0.6, 0), (-1.2, 0.6, 0.2), (0.0, 0.6, 0), (1.2, 0.6, 0.3), (2.4, 0.6, 0), (-1.8, 0.6, 1.8),
           (-0.6, 0.6, 1.6), (0.6, 0.6, 1.7), (1.8, 0.6, 1.9)]
    for i, p in enumerate(pos):
        inst.append(dict(mesh=0, material=[0, 2, 3, 4, 5, 6, 7, 8, 10][i], transform=dict(translation=list(p), scale=[0.5, 0.5, 0.5],
                                                                                        rotation=[10 * i, 20 * i, 5 * i])))
    inst.append(dict(mesh=1, material=9, transform=dict(translation=[0, 0, 1], scale=[8, 8, 1], rotation=[90, 0, 0])))
    inst.append(dict(mesh=1, material=1, transform=dict(translation=[0, 3.0, 1], scale=[1.5, 1.5, 1], rotation=[90, 0, 0])))
    env = dict(type="bitmap", texels=sky_env(64, 32, sun_radiance=200.0, sun_diam_deg=8.0))
    spec = dict(meshes=[sm, quad_mesh()], instances=inst,
                spheres=[dict(center=[0.0, 1.6, 2.8], radius=0.5, material=5), dict(center=[-2.5, 1.5, 2.5], radius=0.4, material=8)],
                materials=mats, mediums=[dict(sigma_a=[0.05, 0.4, 1.2], sigma_s=2.0, g=0.7)], environment=env,
                camera=dict(fov=60.0, dof=1.0, aperture=0.0, transform=dict(translation=[0, 2.2, -4.0], lookat=[0, 0.6, 1.0])))
    return spec, _settings(64, 48, 4, 12)


CONFIGS = {0: cornell, 1: gold_spheres, 2: subsurface}

# Scenes of the request mixes: 0-2 are the configurations above, 3 the instanced mesh.
MIX_SCENES = {0: cornell, 1: gold_spheres, 2: subsurface, 3: instanced}


def batch_requests(seed=3, n=64):
    """Batched workload: n requests, scene uniform over {0,1,2}, square side uniform over
    {128,256,512,1024}, spp uniform in 4..256; depth and tonemap from the scene."""
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n):
        sc = int(rng.randint(3))
        side = [128, 256, 512, 1024][rng.randint(4)]
        out.append(dict(scene=sc, width=side, height=side, spp=int(rng.randint(4, 257))))
    return out


def serve_trace(seed=4, rate=20.0, duration=60.0):
    """Served workload: Poisson arrivals (rate per second) over duration seconds; scene uniform
    over MIX_SCENES, resolution uniform over {256^2, 512^2, 1280x720, 1920x1080}, spp uniform in
    8..512, depth uniform in 8..32. Arrival times t are seconds from the start."""
    rng = np.random.RandomState(seed)
    res = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    out = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        w, h = res[rng.randint(4)]
        out.append(dict(t=t, scene=int(rng.randint(4)), width=w, height=h, spp=int(rng.randint(8, 513)),
                        max_depth=int(rng.randint(8, 33))))
    return out


def write_obj(mesh, path, normals=False):
    """Write a mesh dict (positions, indices; vertex normals when normals=True) as a Wavefront
    OBJ: `v x y z` lines with shortest round-trip float text, then `f a b c` (or `f a//a ...`)
    with 1-based indices."""
    pos = np.asarray(mesh["positions"], np.float32)
    idx = np.asarray(mesh["indices"], np.int64) + 1
    with open(path, "w") as f:
        f.write("".join("v %r %r %r\n" % (float(x), float(y), float(z)) for x, y, z in pos))
        if normals:
            nrm = np.asarray(mesh["normals"], np.float32)
            f.write("".join("vn %r %r %r\n" % (float(x), float(y), float(z)) for x, y, z in nrm))
            f.write("".join("f %d//%d %d//%d %d//%d\n" % (a, a, b, b, c, c) for a, b, c in idx))
        else:
            f.write("".join("f %d %d %d\n" % (a, b, c) for a, b, c in idx))
