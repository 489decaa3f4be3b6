import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ctx():
    import illume_amd
    c = illume_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_quant():
    import illume_amd
    c = illume_amd.Context(0, bvh_layout="quant8")
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_wt():
    import illume_amd
    c = illume_amd.Context(0, tri_test="watertight")
    yield c
    c.close()


def camera_rays(spec, settings, n, seed=0):
    """World-space camera rays through random pixels (pinhole), for closest-hit batches."""
    from oracle.pyoracle import lib, vec
    import ctypes as C
    from illume_amd._abi import transform_desc
    rng = np.random.default_rng(seed)
    W, H = settings["width"], settings["height"]
    cam = spec["camera"]
    td = transform_desc(cam.get("transform"))
    m = np.empty(48, np.float32)
    lib().orc_transform_matrices(C.byref(td), m.ctypes.data_as(C.POINTER(C.c_float)), None)
    M = m[:16].reshape(4, 4).astype(np.float64)
    tanh = np.tan(np.pi * cam["fov"] / 360.0)
    dof = cam.get("dof", 1.0)
    ps = 2 * dof * tanh / W
    x = rng.uniform(0, W, n)
    y = rng.uniform(0, H, n)
    rx = -dof * tanh + ps * x
    ry = (H / W) * dof * tanh - ps * y
    p = np.stack([rx, ry, np.full(n, dof), np.ones(n)], 1) @ M.T
    o = M[:3, 3]
    d = p[:, :3] - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([np.broadcast_to(o, (n, 3)), d], 1).astype(np.float32)


def random_rays(lo, hi, n, seed=1):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def compare_hits(a, b, tol=1e-6):
    """a, b: (inst, prim, t). Returns (n_exact, n_tie, list of mismatch indices). Exact means the
    same primitive; a tie is a different primitive at a distance equal within tol (relative)."""
    ia, pa, ta = a
    ib, pb, tb = b
    same = (ia == ib) & (pa == pb)
    both_miss = (ia == -1) & (ib == -1)
    same = same | both_miss
    rel = np.abs(ta.astype(np.float64) - tb) / np.maximum(np.abs(tb.astype(np.float64)), 1e-30)
    tie = (~same) & (ia != -1) & (ib != -1) & (rel <= tol)
    bad = np.nonzero(~same & ~tie)[0]
    return int(same.sum()), int(tie.sum()), bad
