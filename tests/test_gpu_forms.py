"""Batched and served forms on the GPU: each request's result is the same alone and in a mix,
graph replays are identical to direct launches, and admission/cancellation behave."""
import time

import numpy as np
import pytest

import illume_amd
import illume_amd.scenes as S
from illume_amd._abi import render_params
from oracle.pyoracle import OracleScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(ctx):
    out = {}
    for name, fn in [("cornell", S.cornell), ("tiny", S.tiny_mesh_scene), ("gold", lambda: S.gold_spheres(n=20))]:
        spec, st = fn()
        out[name] = (spec, st, ctx.scene(spec))
    return out


def _p(st, **kw):
    d = dict(st)
    d.update(kw)
    return render_params(**d)


MIX = [("cornell", dict(width=64, height=64, spp=5)), ("tiny", dict(width=80, height=60, spp=3)),
       ("gold", dict(width=96, height=72, spp=4, max_depth=6)), ("cornell", dict(width=40, height=30, spp=7, first_sample=9)),
       ("tiny", dict(width=33, height=17, spp=2, max_depth=20)), ("gold", dict(width=128, height=128, spp=2))]


def test_graph_replays_match_direct_launches(scenes):
    spec, st, _ = scenes["tiny"]
    p = _p(st, width=64, height=48, spp=6)
    with illume_amd.Context(0, use_graphs=0) as c0, illume_amd.Context(0, use_graphs=1, paths_in_flight=8192) as c1:
        s0, s1 = c0.scene(spec), c1.scene(spec)
        ref, _, i0 = c0.render(s0, p)
        outs = []
        for k in range(3):   # replays with eager work in between
            a, _, i1 = c1.render(s1, p)
            outs.append(a)
            c0.render(s0, _p(st, width=16, height=16, spp=1))
        for a in outs:
            assert np.array_equal(a, ref)
        assert i1["segments"] == i0["segments"] and i1["chunks"] > 1


@pytest.mark.parametrize("order", [dict(sample_group=4, tile_rows=8), dict(sample_group=64, tile_rows=4),
                                   dict(sample_group=2, tile_rows=3, paths_in_flight=3000),
                                   dict(extend_mode=1), dict(extend_mode=2, refill_min=4), dict(extend_mode=3, refill_min=1),
                                   dict(extend_mode=3, refill_min=48, sample_group=16),
                                   dict(extend_mode=3, leaf_wait_pct=10), dict(extend_mode=3, leaf_wait_pct=5000),
                                   dict(shade_sort=2), dict(shade_sort=1, paths_in_flight=5000),
                                   dict(camera_packets=2), dict(camera_packets=1, sample_group=1, tile_rows=1),
                                   dict(sched_fair=1, paths_in_flight=3000)])
def test_path_order_leaves_results_unchanged(ctx, scenes, order):
    """Samples grouped per pixel and tiled pixel walks only reorder path ids: every pixel still
    receives the same samples, summed in pass order, so the images are bit-identical."""
    with illume_amd.Context(0, **order) as c:
        for name, kw in [("tiny", dict(width=33, height=17, spp=6)), ("cornell", dict(width=64, height=48, spp=8)),
                         ("gold", dict(width=40, height=40, spp=3, max_depth=5))]:
            spec, st, s_ref = scenes[name]
            p = _p(st, **kw)
            ref, px_ref, i_ref = ctx.render(s_ref, p)
            s = c.scene(spec)
            got, px, i = c.render(s, p)
            s.free()
            assert np.array_equal(got, ref), (name, order)
            assert np.array_equal(px, px_ref)
            assert i["segments"] == i_ref["segments"]


def test_two_pipelines_match_one(scenes):
    """Chunks alternating between two pipelines (two streams, accumulations kept in chunk order
    by events) give the same sums, pixels and segment counts as one pipeline: a render split into
    many chunks (some holding a pixel range of one pass, so consecutive chunks add into the same
    pixels), a batch of the mix, and served requests, each with and without captured graphs."""
    spec, st, _ = scenes["gold"]
    p = _p(st, width=96, height=64, spp=9, max_depth=7)   # 6144 px per pass: two chunks per pass
    specs = {k: v[0] for k, v in scenes.items()}
    results = {}
    for pipes in (1, 2):
        for graphs in (0, 1):
            with illume_amd.Context(0, pipes=pipes, use_graphs=graphs, paths_in_flight=4096) as c:
                sc = {k: c.scene(v) for k, v in specs.items()}
                a, px, info = c.render(sc["gold"], p)
                batch, binfo = c.render_batch([sc[k] for k, _ in MIX], [_p(scenes[k][1], **kw) for k, kw in MIX])
                hs = [c.submit(sc[k], _p(scenes[k][1], **kw)) for k, kw in MIX]
                served = []
                for h in hs:
                    h.wait(60000)
                    served.append(h.result())
                    h.free()
                for s_ in sc.values():
                    s_.free()
                results[(pipes, graphs)] = (a, px, info["segments"], info["chunks"], batch, binfo["segments"], served)
    ref = results[(1, 0)]
    assert ref[3] == 18
    for key, r in results.items():
        assert np.array_equal(r[0], ref[0]) and np.array_equal(r[1], ref[1]), key
        assert r[2] == ref[2] and r[5] == ref[5], key
        for (ba, bp), (ra, rp) in zip(r[4], ref[4]):
            assert np.array_equal(ba, ra) and np.array_equal(bp, rp), key
        for (sa, sp), (ra, rp) in zip(r[6], ref[4]):
            assert np.array_equal(sa, ra) and np.array_equal(sp, rp), key


def test_lone_render_split_in_two_matches_one_chunk(scenes):
    """A lone request of at least 2^22 paths can run as two chunks of half its pixels (all passes
    each) on the two pipelines (lone_split): sums, pixels and segment counts equal the one-chunk
    render, with and without captured graphs, and for a pass count that is not a power of two (a
    sample group below the pass count). By default only renders of max_depth 24 or more split."""
    spec, st, _ = scenes["gold"]
    for spp in (16, 24):
        p = _p(st, width=512, height=384, spp=spp, max_depth=6)   # 3.1 M and 4.7 M paths
        out = {}
        for pipes in (1, 2):
            for graphs in (0, 1):
                with illume_amd.Context(0, pipes=pipes, use_graphs=graphs, lone_split=2) as c:
                    sc = c.scene(spec)
                    out[(pipes, graphs)] = c.render(sc, p)
                    sc.free()
        ref = out[(1, 0)]
        big = 512 * 384 * spp >= 1 << 22
        for key, (a, px, info) in out.items():
            assert np.array_equal(a, ref[0]) and np.array_equal(px, ref[1]), (spp, key)
            assert info["segments"] == ref[2]["segments"], (spp, key)
            assert info["chunks"] == (2 if key[0] == 2 and big else 1), (spp, key, info["chunks"])
    deep = _p(st, width=512, height=384, spp=24, max_depth=24)
    res = {}
    for split in (0, 1):
        with illume_amd.Context(0, lone_split=split) as c:
            sc = c.scene(spec)
            res[split] = c.render(sc, deep)
            res[(split, "shallow")] = c.render(sc, _p(st, width=512, height=384, spp=24, max_depth=6))[2]["chunks"]
            sc.free()
    assert res[0][2]["chunks"] == 2 and res[1][2]["chunks"] == 1
    assert res[(0, "shallow")] == 1 and res[(1, "shallow")] == 1
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert res[0][2]["segments"] == res[1][2]["segments"]


def test_batch_equals_individual_renders(ctx, scenes):
    ss = [scenes[n][2] for n, _ in MIX]
    ps = [_p(scenes[n][1], **kw) for n, kw in MIX]
    batch, info = ctx.render_batch(ss, ps)
    segs = 0
    for (acc_b, px_b), s, p in zip(batch, ss, ps):
        acc, px, i = ctx.render(s, p)
        segs += i["segments"]
        assert np.array_equal(acc_b, acc)
        assert np.array_equal(px_b, px)
    assert info["segments"] == segs


def test_served_requests_equal_plain_renders(ctx, scenes):
    reqs = []
    for name, kw in MIX * 2:
        reqs.append((ctx.submit(scenes[name][2], _p(scenes[name][1], **kw)), name, kw))
    for r, name, kw in reqs:
        r.wait(60000)
        st = r.poll()
        assert st["state"] == "done" and st["end_ms"] >= st["start_ms"] >= st["submit_ms"]
        acc, px = r.result()
        ref, refpx, _ = ctx.render(scenes[name][2], _p(scenes[name][1], **kw))
        assert np.array_equal(acc, ref), (name, kw)
        assert np.array_equal(px, refpx)
        r.free()
    s = ctx.stats()
    assert s["requests_done"] >= len(reqs) and s["p99_ms"] >= s["p50_ms"] > 0


def test_fair_share_served_and_batched_equal_plain(ctx, scenes):
    """Fair-share packing with a small chunk: every chunk mixes several requests' slices (pixel
    ranges and pass groups); each result still equals its plain render."""
    with illume_amd.Context(0, sched_fair=1, paths_in_flight=20000) as c:
        local = {n: c.scene(scenes[n][0]) for n in scenes}
        ps = [(n, _p(scenes[n][1], **kw)) for n, kw in MIX]
        reqs = [(c.submit(local[n], p), n, p) for n, p in ps]
        batch, _ = c.render_batch([local[n] for n, _ in ps], [p for _, p in ps])
        for (r, n, p), (acc_b, px_b) in zip(reqs, batch):
            r.wait(120000)
            acc, px = r.result()
            ref, refpx, _ = ctx.render(scenes[n][2], p)
            assert np.array_equal(acc, ref) and np.array_equal(px, refpx), n
            assert np.array_equal(acc_b, ref) and np.array_equal(px_b, refpx), n
            r.free()
        for sc in local.values():
            sc.free()


def test_cancel_and_admission_limits(scenes):
    spec, st, _ = scenes["gold"]
    # long enough (512x512 x 4096 spp, about a billion segments) that a request is still queued
    # or running when it is cancelled
    big = _p(st, width=512, height=512, spp=4096)
    # a 512x512 request holds 512*512*56 B = 14.7 MB of HBM while admitted; the budget fits one
    with illume_amd.Context(0, max_queued=2, hbm_budget_bytes=20 * 2 ** 20) as c:
        s = c.scene(spec)
        # over the HBM budget on its own: rejected at submission
        with pytest.raises(illume_amd.IlmError) as e:
            c.submit(s, _p(st, width=2048, height=2048, spp=1))
        assert e.value.code == -2

        def wait_state(r, want):
            for _ in range(2000):
                if r.poll()["state"] == want:
                    return
                time.sleep(0.001)
            raise AssertionError("request never reached " + want)
        r1 = c.submit(s, big)          # admitted (runs)
        wait_state(r1, "running")
        r2 = c.submit(s, big)          # waits: the budget holds one
        r3 = c.submit(s, big)
        assert r2.poll()["state"] == "queued"
        with pytest.raises(illume_amd.IlmError) as e:
            c.submit(s, big)           # queue of 2 is full
        assert e.value.code == -6
        r2.cancel()
        with pytest.raises(illume_amd.IlmError) as e:
            r2.wait(60000)
        assert e.value.code == -7
        r1.cancel()
        r3.cancel()
        for r in (r1, r3):
            try:
                r.wait(60000)
            except illume_amd.IlmError as ex:
                assert ex.code == -7
        st2 = c.stats()
        assert st2["requests_rejected"] == 2 and st2["requests_cancelled"] >= 1
        for r in (r1, r2, r3):
            r.free()
        s.free()


def _kahan_f32(samples):
    """Kahan sum over axis 0 in float32, in order (the device's accumulation, no fused ops)."""
    s = np.zeros(samples.shape[1:], np.float32)
    c = np.zeros_like(s)
    for x in samples.astype(np.float32):
        y = (x - c).astype(np.float32)
        t = (s + y).astype(np.float32)
        c = ((t - s).astype(np.float32) - y).astype(np.float32)
        s = t
    return s


def test_first_sample_offsets_compose(ctx, scenes):
    """Pass k of a render seeds wang_hash(first_sample + k): one-sample renders at first_sample
    = 0..5 give each pass's radiance exactly (a Kahan sum of one value is that value), and their
    float32 Kahan sum in pass order equals the 6-sample render bit for bit; a render of passes
    2..5 equals the Kahan sum of those four."""
    spec, st, s = scenes["cornell"]
    full, _, _ = ctx.render(s, _p(st, width=32, height=32, spp=6))
    one = np.stack([ctx.render(s, _p(st, width=32, height=32, spp=1, first_sample=k))[0] for k in range(6)])
    assert np.array_equal(full, _kahan_f32(one))
    tail, _, _ = ctx.render(s, _p(st, width=32, height=32, spp=4, first_sample=2))
    assert np.array_equal(tail, _kahan_f32(one[2:]))


def test_layout_switch_by_tune_matches_a_context_opened_with_it(ctx_quant, scenes):
    """Every scene carries both node layouts: retuning a live context to the 8-bit BVH4 renders
    what a context opened with it renders, and switching back restores the FP32 image."""
    spec, st, _ = scenes["gold"]
    p = _p(st, width=64, height=48, spp=3, max_depth=6)
    with illume_amd.Context(0) as c:
        s = c.scene(spec)
        a, _, ia = c.render(s, p)
        c.tune(bvh_layout="quant8")
        assert c.settings()["bvh_layout"] == 1
        b, _, ib = c.render(s, p)
        c.tune(bvh_layout="fp32")
        a2, _, ia2 = c.render(s, p)
        s.free()
    sq = ctx_quant.scene(spec)
    q, _, iq = ctx_quant.render(sq, p)
    sq.free()
    assert np.array_equal(b, q) and ib["segments"] == iq["segments"]
    assert np.array_equal(a, a2) and ia["segments"] == ia2["segments"]


def test_quantized_and_watertight_renders_agree(ctx_quant, ctx_wt, ctx, scenes):
    spec, st, s = scenes["tiny"]
    p = _p(st, width=64, height=48, spp=4)
    ref, _, i = ctx.render(s, p)
    q, _, iq = ctx_quant.render(ctx_quant.scene(spec), p)
    assert np.array_equal(q, ref) and iq["segments"] == i["segments"]
    w, _, iw = ctx_wt.render(ctx_wt.scene(spec), p)
    # watertight differs from Moller-Trumbore only on edge/grazing hits
    rel = np.abs(w - ref) / np.maximum(np.abs(ref), 1e-3)
    assert np.mean(rel < 1e-4) > 0.95


def test_long_queues_match_across_reservations_and_oracle(scenes):
    """Queues longer than the extend grid's first reservations: with 64-ray reservations every
    extend kernel (per-step and per-lane camera rays, camera packets, leaf refill, batch) hands
    out the rays past them from the shared queue cursor. The images are the same for every
    reservation size, each ray traced once, and match the oracle."""
    spec, st, _ = scenes["gold"]
    p = _p(st, width=640, height=480, spp=8, max_depth=6)
    shapes = [dict(ext_fetch=64), dict(ext_fetch=64, camera_packets=2), dict(ext_fetch=64, extend_mode=1),
              dict(ext_fetch=64, extend_mode=2, refill_min=4), dict(ext_fetch=4096)]
    outs = []
    for kw in shapes:
        with illume_amd.Context(0, bvh_layout="quant8", **kw) as c:
            s = c.scene(spec)
            acc, px, info = c.render(s, p)
            s.free()
        outs.append((acc, px, info["segments"]))
    for (acc, px, segs), kw in zip(outs[1:], shapes[1:]):
        assert np.array_equal(acc, outs[0][0]) and np.array_equal(px, outs[0][1]), kw
        assert segs == outs[0][2], kw
    orc = OracleScene(spec)
    oacc, _, osegs, _ = orc.render(p, kahan=True)
    orc.free()
    oacc = oacc.reshape(outs[0][0].shape)
    rel = np.abs(outs[0][0] - oacc) / np.maximum(np.abs(oacc), 1e-6)
    print("640x480x8: segments gpu %d oracle %d, pixels within 1e-4 %.5f" % (outs[0][2], osegs, np.mean(rel <= 1e-4)))
    assert outs[0][2] == osegs
    assert np.mean(rel <= 1e-4) >= 0.999


def test_stats_poll_and_submit_do_not_wait_for_the_chunk_in_flight():
    """ilm_stats_get reads host-side counters and the scheduler never waits on the device while
    holding the scheduling lock, so a telemetry poll, a status poll and a submission return
    promptly while a long request occupies the device (subsurface scene, 1024x1024 x 2048 spp:
    32 chunks of 2^26 paths, about 100 ms each; the request is cancelled at the end)."""
    spec, st = S.subsurface(env_size=(256, 128))
    with illume_amd.Context(0) as c:
        s = c.scene(spec)
        long = c.submit(s, _p(st, width=1024, height=1024, spp=2048))
        for _ in range(5000):
            if long.poll()["state"] == "running":
                break
            time.sleep(0.001)
        time.sleep(0.3)   # chunks are in flight now
        worst = {}
        small = None
        for k in range(5):
            t0 = time.perf_counter()
            c.stats()
            worst["stats"] = max(worst.get("stats", 0.0), time.perf_counter() - t0)
            t0 = time.perf_counter()
            long.poll()
            worst["poll"] = max(worst.get("poll", 0.0), time.perf_counter() - t0)
            if small is None:
                t0 = time.perf_counter()
                small = c.submit(s, _p(st, width=32, height=32, spp=1))
                worst["submit"] = time.perf_counter() - t0
            time.sleep(0.05)
        assert long.poll()["state"] == "running"
        print("while a long request runs: worst stats %.2f ms, poll %.2f ms, submit %.2f ms"
              % (worst["stats"] * 1e3, worst["poll"] * 1e3, worst["submit"] * 1e3))
        assert max(worst.values()) < 0.05, worst
        long.cancel()
        for r in (long, small):
            try:
                r.wait(120000)
            except illume_amd.IlmError as ex:
                assert ex.code == -7
            r.free()
        s.free()
