"""The batched and served workloads at their full definitions, on the GPU.

* Batched: all 64 seeded requests of `scenes.batch_requests()` (scenes cornell / gold / subsurface,
  square sides 128..1024, 4..256 spp, each scene's own depth and tonemap) in one
  `ilm_render_batch` call, once with 8-bit BVH4 nodes and once with FP32 BVH2 nodes, against the
  same requests rendered one at a time with `ilm_render`.
* Served: the first 8 requests of the Poisson trace `scenes.serve_trace()` (20 req/s; they span
  all four scenes, two of them on the 12-instance 722,000-triangle mesh, 256^2..1920x1080,
  8..393 spp, depth 10..29), submitted through `ilm_submit` at their arrival times, against plain
  `ilm_render` calls.

Each test requires: every image bit-equal to its plain render (accumulated sums and tonemapped
pixels); segment totals equal; pixels of every request (128 of its middle row, pooled over the
requests) within 1e-4 relative of the FP32 oracle for >= 99.9 % of them (the oracle keeps the
reference's first-met choice at exact distance ties, the device the smaller primitive); and the
relative RMSE of those pixels against the FP64 oracle (same RNG stream) <= 1 % per request.
The reference renders one scene per process (src/main.cu:14-43): these forms have no
counterpart there and are checked against its per-path routine through the oracle.
"""
import time

import numpy as np
import pytest

import illume_amd
import illume_amd.scenes as S
from illume_amd._abi import render_params
from oracle.pyoracle import OracleScene

pytestmark = pytest.mark.gpu

ORACLE_PX = 128
REL_TOL = 1e-4
MATCH_FRAC = 0.999
RMSE_MAX = 0.01


@pytest.fixture(scope="module")
def mix_specs():
    return {k: fn() for k, fn in S.MIX_SCENES.items()}


@pytest.fixture(scope="module")
def oracles(mix_specs):
    """FP32 + FP64 oracle copies of the mix scenes (median-split BVH: hits equal the reference
    tree's except at ties, and the build takes about a second on the 0.5-0.7M-triangle meshes)."""
    out = {k: OracleScene(spec, fp64=True, median_bvh=True) for k, (spec, _) in mix_specs.items()}
    yield out
    for o in out.values():
        o.free()


def _middle_row(p):
    n = min(ORACLE_PX, p.width)
    pb = (p.height // 2) * p.width + (p.width - n) // 2
    return pb, pb + n


def check_against_oracle(oracles, keys, params, accs, label):
    """Pooled FP32-oracle agreement and per-request FP64 RMSE over the middle-row pixels."""
    total, within = 0, 0
    worst = 0.0
    for k, p, acc in zip(keys, params, accs):
        pb, pe = _middle_row(p)
        g = np.asarray(acc, np.float32).reshape(-1, 3)[pb:pe]
        o32, _, _, _ = oracles[k].render(p, pb, pe, kahan=True, pixels=False)
        o64, _, _, _ = oracles[k].render(p, pb, pe, kahan=True, fp64=True, pixels=False)
        rel = np.abs(g - o32) / np.maximum(np.abs(o32), 1e-6)
        px_ok = np.all(rel <= REL_TOL, axis=1)
        total += px_ok.size
        within += int(px_ok.sum())
        ga, oa = g.astype(np.float64) / p.spp, np.asarray(o64, np.float64) / p.spp
        e = float(np.sqrt(np.mean((ga - oa) ** 2)) / max(np.mean(np.abs(oa)), 1e-12))
        worst = max(worst, e)
        assert e <= RMSE_MAX, (label, k, p.width, p.height, p.spp, e)
    frac = within / total
    print("%s: %d/%d oracle pixels within %.0e (%.5f), worst RMSE vs FP64 %.5f" % (label, within, total, REL_TOL, frac, worst))
    assert frac >= MATCH_FRAC, (label, frac)


def test_batched_64_requests_equal_plain_and_oracle(mix_specs, oracles):
    reqs = S.batch_requests()
    assert len(reqs) == 64
    params = [render_params(**dict(mix_specs[r["scene"]][1], width=r["width"], height=r["height"], spp=r["spp"]))
              for r in reqs]
    keys = [r["scene"] for r in reqs]
    results = {}
    for layout in ("quant8", "fp32"):
        with illume_amd.Context(0, bvh_layout=layout) as c:
            sc = {k: c.scene(mix_specs[k][0]) for k in sorted(set(keys))}
            t0 = time.perf_counter()
            batch, info = c.render_batch([sc[k] for k in keys], params)
            tb = time.perf_counter() - t0
            plain = None
            if layout == "quant8":
                t0 = time.perf_counter()
                plain, psegs = [], 0
                for k, p in zip(keys, params):
                    a, px, i = c.render(sc[k], p)
                    plain.append((a, px))
                    psegs += i["segments"]
                tp = time.perf_counter() - t0
                print("batched 64 requests [%s]: %d segments, batch %.2f s, plain %.2f s"
                      % (layout, info["segments"], tb, tp))
                assert info["segments"] == psegs
            for s in sc.values():
                s.free()
        results[layout] = (batch, info["segments"], plain)
    batch_q, segs_q, plain_q = results["quant8"]
    batch_f, segs_f, _ = results["fp32"]
    not_plain = [n for n, ((aq, pq), (ap, pp)) in enumerate(zip(batch_q, plain_q))
                 if not (np.array_equal(aq, ap) and np.array_equal(pq, pp))]
    not_layout = [n for n, ((af, pf), (ap, pp)) in enumerate(zip(batch_f, plain_q))
                  if not (np.array_equal(af, ap) and np.array_equal(pf, pp))]
    print("batched 64: %d images differ from plain, %d differ between FP32 and 8-bit nodes"
          % (len(not_plain), len(not_layout)))
    assert not not_plain, [(n, reqs[n]) for n in not_plain]
    assert not not_layout, [(n, reqs[n]) for n in not_layout]
    assert segs_q == segs_f
    check_against_oracle(oracles, keys, params, [a for a, _ in batch_q], "batched 64")


def test_served_trace_prefix_equals_plain_and_oracle(mix_specs, oracles):
    trace = S.serve_trace()[:8]
    assert sorted({r["scene"] for r in trace}) == [0, 1, 2, 3]
    assert sum(r["scene"] == 3 for r in trace) >= 2
    t0 = trace[0]["t"]
    keys = [r["scene"] for r in trace]
    params = [render_params(**dict(mix_specs[r["scene"]][1], width=r["width"], height=r["height"], spp=r["spp"],
                                   max_depth=r["max_depth"])) for r in trace]
    with illume_amd.Context(0, bvh_layout="quant8") as c:
        sc = {k: c.scene(mix_specs[k][0]) for k in sorted(set(keys))}
        s0 = c.stats()["segments"]
        handles = []
        start = time.perf_counter()
        for r, p in zip(trace, params):
            dt = start + (r["t"] - t0) - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            handles.append(c.submit(sc[r["scene"]], p))
        served = []
        for h in handles:
            h.wait(120000)
            st = h.poll()
            assert st["state"] == "done" and st["end_ms"] >= st["start_ms"] >= st["submit_ms"]
            served.append(h.result())
            h.free()
        span = time.perf_counter() - start
        served_segs = c.stats()["segments"] - s0
        plain_segs = 0
        for n, (k, p) in enumerate(zip(keys, params)):
            a, px, i = c.render(sc[k], p)
            plain_segs += i["segments"]
            assert np.array_equal(served[n][0], a) and np.array_equal(served[n][1], px), ("served vs plain", n, trace[n])
        print("served trace prefix: %d requests, %d segments in %.2f s" % (len(trace), served_segs, span))
        assert served_segs == plain_segs
        for s in sc.values():
            s.free()
    check_against_oracle(oracles, keys, params, [a for a, _ in served], "served 8")
