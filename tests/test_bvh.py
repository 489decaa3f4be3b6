"""BVH builder invariants (CPU), and the oracle's reference traversal against brute force.

The product builder (binned SAH) is inspected through ilm_bvh_build/ilm_bvh_export; the
reference's sweep builder lives in the oracle. Closest hit does not depend on the tree except
for ties, which is what lets the device use its own tree and still match the reference.
"""
import numpy as np
import pytest

import illume_amd
import illume_amd.scenes as S
from oracle.pyoracle import OracleScene
from conftest import camera_rays, compare_hits, random_rays

DEPTH_CAP_PLUS_MEDIAN = 64


def _tri_boxes(pos, idx):
    t = pos[idx]
    return t.min(axis=1), t.max(axis=1)


@pytest.mark.parametrize("n,seed", [(3, 1), (12, 2), (40, 3)])
def test_product_bvh_invariants(n, seed):
    m = S.displaced_icosphere(n, seed)
    pos, idx = m["positions"], m["indices"]
    b = illume_amd.bvh_build(pos, idx)
    info = b["info"]
    boxes, ch, first, count, order = b["boxes"], b["children"], b["leaf_first"], b["leaf_count"], b["order"]
    nt = idx.shape[0]
    # every triangle in exactly one leaf
    assert np.array_equal(np.sort(order), np.arange(nt))
    leaves = np.nonzero(ch[:, 0] < 0)[0]
    covered = np.zeros(nt, np.int64)
    for l in leaves:
        covered[first[l]:first[l] + count[l]] += 1
    assert np.all(covered == 1)
    assert info["max_leaf_size"] <= 16 and info["num_leaves"] == len(leaves)
    assert info["max_depth"] <= DEPTH_CAP_PLUS_MEDIAN
    # containment: leaf box holds its triangles; parent box holds both children
    lo, hi = _tri_boxes(pos, idx)
    for l in leaves:
        tr = order[first[l]:first[l] + count[l]]
        assert np.all(boxes[l, :3] <= lo[tr]) and np.all(boxes[l, 3:] >= hi[tr])
    inner = np.nonzero(ch[:, 0] >= 0)[0]
    for p in inner:
        for c in ch[p]:
            assert np.all(boxes[p, :3] <= boxes[c, :3]) and np.all(boxes[p, 3:] >= boxes[c, 3:])
    # quantised layout: every dequantised child box contains the exact FP32 box
    qb, ex, qc = b["q_boxes"], b["q_exact"], b["q_child"]
    used = qc != 0x7FFFFFFF
    assert np.all(qb[..., :3][used] <= ex[..., :3][used])
    assert np.all(qb[..., 3:][used] >= ex[..., 3:][used])
    # and is not needlessly loose: on average within a few grid steps of the exact box
    exu, qbu = ex[used], qb[used]
    ext = exu[:, 3:] - exu[:, :3]
    slack = (exu[:, :3] - qbu[:, :3]) + (qbu[:, 3:] - exu[:, 3:])
    assert np.median(slack / np.maximum(ext, 1e-6)) < 0.1


def test_sah_cost_beats_reference_sweep_builder():
    """The binned-SAH tree against the reference's sweep-built tree on the same mesh, by SAH cost
    (traversal 1, intersection 1 per triangle, areas relative to the root)."""
    m = S.displaced_icosphere(16, 4)
    pos, idx = m["positions"], m["indices"]
    b = illume_amd.bvh_build(pos, idx)
    spec = dict(meshes=[dict(positions=pos, indices=idx, face_normals=True, leaf_size=4)],
                instances=[dict(mesh=0, material=0, transform={})], spheres=[],
                materials=[dict(type="lambert", albedo=0.5)], mediums=[], environment=None,
                camera=dict(fov=40, transform={}))
    orc = OracleScene(spec)
    boxes, ab, ti = orc.ref_bvh(0)

    def ha(bx):
        d = np.maximum(bx[..., 3:] - bx[..., :3], 0).astype(np.float64)
        return d[..., 0] * d[..., 1] + d[..., 1] * d[..., 2] + d[..., 2] * d[..., 0]
    root = ha(boxes[0])
    leaf = (ab[:, 0] & 0x80000000) != 0
    ref_cost = ha(boxes[~leaf]).sum() / root + (ha(boxes[leaf]) * (ab[leaf, 0] & 0x7FFFFFFF)).sum() / root
    print("SAH cost: binned %.2f, reference sweep %.2f" % (b["info"]["sah_cost"], ref_cost))
    assert b["info"]["sah_cost"] < ref_cost * 1.1


def _mesh_spec(n=20, seed=3, instances=None):
    m = S.displaced_icosphere(n, seed)
    inst = instances or [dict(mesh=0, material=0, transform=dict(translation=[0, 0, 0]))]
    return dict(meshes=[m], instances=inst, spheres=[dict(center=[2.5, 0, 0], radius=0.7, material=0)],
                materials=[dict(type="lambert", albedo=0.5)], mediums=[], environment=None,
                camera=dict(fov=40, transform=dict(translation=[0, 0, -5])))


def test_reference_traversal_equals_brute_force():
    spec = _mesh_spec(instances=[dict(mesh=0, material=0, transform=dict(translation=[0, 0, 0])),
                                 dict(mesh=0, material=0, transform=dict(translation=[-2.2, 0.3, 0.5], scale=[0.6, 0.8, 0.7],
                                                                         rotation=[20, 40, 10]))])
    orc = OracleScene(spec)
    rays = random_rays([-3, -1.5, -3], [3, 1.5, 3], 20000, seed=11)
    bvh = orc.closest_hit(rays)
    brute = orc.closest_hit(rays, brute=True)
    exact, tie, bad = compare_hits(bvh, brute)
    print("reference traversal vs brute force: %d same, %d ties, %d other" % (exact, tie, len(bad)))
    assert len(bad) == 0


def test_bvh_cache_round_trip_and_rejection(tmp_path):
    """A cached entry gives the same device layouts as a fresh build; a corrupted or foreign
    entry is rejected and rebuilt; a different mesh gets its own entry."""
    from illume_amd import bvh_device_layouts
    m = S.displaced_icosphere(12, 3)
    fresh = bvh_device_layouts(m["positions"], m["indices"])
    assert not fresh[4]
    first = bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))
    assert not first[4]                     # miss: built and stored
    files = list(tmp_path.glob("*.ilmbvh"))
    assert len(files) == 1
    second = bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))
    assert second[4]                        # hit
    for a, b in zip(fresh[:3], second[:3]):
        assert np.array_equal(a, b)
    assert fresh[3] == second[3]
    # flip one payload byte: the payload hash rejects the entry, which is rebuilt and rewritten
    raw = bytearray(files[0].read_bytes())
    raw[len(raw) // 2] ^= 0x40
    files[0].write_bytes(bytes(raw))
    third = bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))
    assert not third[4] and np.array_equal(third[0], fresh[0])
    assert bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))[4]
    # truncated file
    files[0].write_bytes(files[0].read_bytes()[:100])
    assert not bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))[4]
    # a moved vertex is a different key
    p2 = np.array(m["positions"], copy=True)
    p2[0, 0] += 1e-3
    other = bvh_device_layouts(p2, m["indices"], cache_dir=str(tmp_path))
    assert not other[4]
    assert len(list(tmp_path.glob("*.ilmbvh"))) == 2


class _CacheHash:
    """The cache's payload hash (illume_amd/csrc/bvh_cache.cpp Hasher), to forge entries whose
    checksum matches while their tree is wrong."""
    M = (1 << 64) - 1

    def __init__(self):
        self.h = 0x243f6a8885a308d3

    def mix(self, w):
        h = self.h
        h ^= (w + 0x9e3779b97f4a7c15 + ((h << 6) & self.M) + (h >> 2)) & self.M
        h = (h * 0xff51afd7ed558ccd) & self.M
        h ^= h >> 33
        self.h = h

    def bytes(self, b):
        n, i = len(b), 0
        while i + 8 <= n:
            self.mix(int.from_bytes(b[i:i + 8], "little"))
            i += 8
        self.mix(int.from_bytes(b[i:] + bytes(8 - (n - i)), "little") ^ (((n - i) << 56) & self.M))
        self.mix(n)


def _forge(path, edit):
    """Rewrite a cache entry after edit(header dict, nodes, qnodes, order), with the payload hash
    recomputed so that only the tree checks can reject it."""
    import struct
    raw = bytearray(path.read_bytes())
    fmt = "<8sIIQQiiiiQQQII"
    hsz = struct.calcsize(fmt)
    f = list(struct.unpack(fmt, raw[:hsz]))
    hdr = dict(zip(["magic", "version", "hb", "key", "hash", "root", "qroot", "depth", "qdepth", "nn", "nq", "no",
                    "sb", "sq"], f))
    nb, qb = hdr["nn"] * 64, hdr["nq"] * 64
    nodes = np.frombuffer(bytes(raw[hsz:hsz + nb]), np.int32).reshape(-1, 16).copy()
    qnodes = np.frombuffer(bytes(raw[hsz + nb:hsz + nb + qb]), np.int32).reshape(-1, 16).copy()
    order = np.frombuffer(bytes(raw[hsz + nb + qb:]), np.int32).copy()
    edit(hdr, nodes, qnodes, order)
    H = _CacheHash()
    for a in (nodes, qnodes, order):
        H.bytes(a.tobytes())
    hdr["hash"] = H.h
    head = struct.pack(fmt, *[hdr[k] for k in ["magic", "version", "hb", "key", "hash", "root", "qroot", "depth",
                                               "qdepth", "nn", "nq", "no", "sb", "sq"]])
    path.write_bytes(head + nodes.tobytes() + qnodes.tobytes() + order.tobytes())


def test_bvh_cache_rejects_entries_with_a_matching_hash_but_a_bad_tree(tmp_path):
    """The payload hash only catches accidents: an entry whose checksum is recomputed after its
    tree was changed must still be refused (and rebuilt) when a child index is out of range, a
    child points back up the tree, a leaf range runs past the triangles, an 8-bit node claims
    fewer than two children, or the header understates a depth that bounds the device
    traversal stack."""
    from illume_amd import bvh_device_layouts
    m = S.displaced_icosphere(4, 3)
    fresh = bvh_device_layouts(m["positions"], m["indices"])
    assert not bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))[4]
    path = next(tmp_path.glob("*.ilmbvh"))
    good = path.read_bytes()
    _forge(path, lambda h, n, q, o: None)    # unchanged tree, recomputed hash: still accepted
    assert bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))[4]

    def bad_child(h, n, q, o):
        n[1, 12] = h["nn"] + 5              # BNode child0 (bytes 48..51) out of range

    def cycle(h, n, q, o):
        n[1, 13] = 0                         # child1 back to the root

    def leaf_past_end(h, n, q, o):
        for i in range(q.shape[0]):         # first 8-bit leaf child: push its range past the end
            for k in range(4):
                c = int(q[i, 12 + k])
                if c < 0 and c != 0x7fffffff:
                    q[i, 12 + k] = np.int32(np.uint32(0x80000000 | (3 << 27) | (h["no"] - 2)).view(np.int32))
                    return

    def shallow(h, n, q, o):
        h["depth"] = 1                       # understated FP32 depth

    def one_child(h, n, q, o):
        # an 8-bit node claiming a single child (the kernel's box test assumes two or more)
        w = int(np.int32(q[0, 3]).view(np.uint32))
        q[0, 3] = np.uint32((w & 0x00ffffff) | (1 << 24)).view(np.int32)
        q[0, 13:16] = 0x7fffffff

    def qshallow(h, n, q, o):
        h["qdepth"] = 0

    for edit in (bad_child, cycle, leaf_past_end, shallow, qshallow, one_child):
        path.write_bytes(good)
        _forge(path, edit)
        got = bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))
        assert not got[4], edit.__name__     # refused, rebuilt and rewritten
        for a, b in zip(fresh[:3], got[:3]):
            assert np.array_equal(a, b), edit.__name__
        assert got[3] == fresh[3]
        assert bvh_device_layouts(m["positions"], m["indices"], cache_dir=str(tmp_path))[4]


def test_tpruned_traversal_same_hits_fewer_tests():
    """The oracle's second traversal (the build's order: children tested together, nearer first,
    subtrees beyond the closest hit skipped), used for the t-pruned CPU baseline figure, finds
    the reference walk's closest hits (ties aside) with fewer box and triangle tests, and renders
    the same image within the parity bar."""
    from illume_amd._abi import render_params
    spec, st = S.gold_spheres(n=24)
    orc = OracleScene(spec)
    rays = np.concatenate([camera_rays(spec, st, 20000, seed=5), random_rays([-2, 0, -2], [2, 2, 2], 20000, seed=6)])
    ref = orc.closest_hit(rays, counts=True)
    orc.set_tpruned(True)
    got = orc.closest_hit(rays, counts=True)
    exact, tie, bad = compare_hits(got[:3], ref[:3])
    assert len(bad) == 0
    assert got[3][:, 0].sum() < ref[3][:, 0].sum() and got[3][:, 1].sum() < ref[3][:, 1].sum()
    p = render_params(**dict(st, width=48, height=32, spp=4, max_depth=6))
    a, _, segs_t, _ = orc.render(p)
    orc.set_tpruned(False)
    b, _, segs_r, _ = orc.render(p)
    orc.free()
    assert segs_t == segs_r
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
    assert np.mean(np.all(rel <= 1e-4, axis=1)) >= 0.999
