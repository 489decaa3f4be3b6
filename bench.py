"""Benchmark of the path-tracer hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1] [--no-cpu] [--no-tune]

Configurations (scene generators in illume_amd/scenes.py, fixed seeds):
  0  Cornell box 256x256, 16 spp, depth 8 (8 instances, 36 triangles)
  1  displaced icosphere (100,820 triangles) in rough gold + rough dielectric sphere + checker
     plane, 1024x1024, 64 spp, depth 16                                       <- the default
  2  displaced icosphere (524,880 triangles) holding a scattering medium inside a smooth
     dielectric, 2048x1024 HDR sky + sun, 960x720, 32 spp, depth 45
  3  batched form: 64 requests over configurations 0-2 (side 128..1024, 4..256 spp) in one
     ilm_render_batch call, against the same requests rendered one at a time
  4  served form: the first 120 requests of the Poisson trace (20 req/s) over configurations 0-2
     and a 12-instance 722k-triangle mesh, submitted at their arrival times to the scheduler,
     against plain one-request-at-a-time renders of the same trace (offered vs achieved load,
     p50/p99 latency, rejections, image identity, RMSE against the FP64 oracle)

For 0-2 a step is one full render of the workload (all passes and depths) through the C ABI
(ilm_render: ray generation, BVH extend, shade + compaction per depth, Kahan accumulation,
tonemap and pixel readback), with the scene and its BVH already resident in HBM.
Prints ONE JSON line on stdout; progress goes to stderr.

value       segments/s / 1e6 over the timed steps, timed on the device (HIP events on the render
            stream from the first kernel of each render to its tonemap, each render one replay of
            captured hipGraphs: scene and path state are resident, the PCIe copy of the image to
            the caller is excluded and reported beside it as value_pcie_inclusive, the wall clock
            around the synchronous calls); a segment is one closest-hit query
            (src/renderer.cu:554 counts the same)
roofline    extend kernel, layout-neutral convention: algorithmic bytes B_seg = 28*N_box +
            36*N_tri + 128 per segment, N_box / N_tri the mean child-box / triangle tests per
            closest-hit query of per-ray ordered, t-pruned walks over the FP32 BVH2, counted by
            the kernel's own counters on the same workload (a separate, untimed render);
            achieved = B_seg * segments / summed extend-launch time, i.e. bytes per launch over
            the average launch duration (HIP events on the render stream around every extend
            launch, in a profiled pass of the same renders right after the timed steps); peak
            8 TB/s HBM3E. roofline.layout repeats it with the benchmarked node layout's own bytes
            (16 B per 8-bit BVH4 child test) and counts; traffic / hbm_measured come from the
            rocprofv3 counter summaries in profiles/
cpu_baseline  the oracle (CPU restatement of the reference per-path routine: exhaustive
            traversal of the reference's sweep-built BVH, linear instance loop) on a band of
            the same image, OpenMP over pixels on the host's cores; beside it a one-core figure
            and value_tpruned, the same band with the build's ordered, t-pruned traversal over
            the same trees (algorithmic vs hardware speedup), and the host CPU model
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "path segments/s (Mrays/s) at 1024², depth 16; p50/p99 render-request latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    pos = q * (len(xs) - 1)
    i = int(pos)
    f = pos - i
    return xs[i] * (1 - f) + xs[min(i + 1, len(xs) - 1)] * f


def workload_name(cfg, spec, st):
    tris = sum(int(np.asarray(m["indices"]).shape[0]) for m in spec["meshes"])
    names = {0: "cornell box (8 instances, 36 tris)",
             1: "displaced icosphere 100,820 tris rough gold GGX + rough dielectric sphere + checker plane",
             2: "displaced icosphere 524,880 tris subsurface medium in smooth dielectric, 2048x1024 HDR sky+sun"}
    return "config%d: %dx%d, %d spp, depth %d, %s" % (cfg, st["width"], st["height"], st["spp"], st["max_depth"],
                                                       names.get(cfg, "%d tris" % tris))


def traffic_from_profiles(cfg, layout):
    """HBM bytes per extend launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration), recorded in profiles/extend_config<cfg>_traffic.txt for the
    node layout named there; None when no profile covers this configuration and layout."""
    path = os.path.join(ROOT, "profiles", "extend_config%d_traffic.txt" % cfg)
    if not os.path.exists(path):
        return None
    value, prof_layout = None, "fp32"
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            value = float(line.split(":")[1])
        if line.startswith("bvh_layout:"):
            prof_layout = line.split(":")[1].strip()
    return value if prof_layout == layout else None


# bytes a traversal reads per child-box test: an FP32 child box (24 B) and its child index; for
# the 8-bit BVH4 a quarter of the 64-byte node
BOX_BYTES = {"fp32": 28.0, "quant8": 16.0}


def profile_value(cfg, key):
    """A `key: value` line of profiles/extend_config<cfg>_counters.txt (rocprofv3 summaries), or None."""
    path = os.path.join(ROOT, "profiles", "extend_config%d_counters.txt" % cfg)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(spec, st, target_s=15.0, target_1core_s=5.0):
    """Oracle render (FP32, reference BVH and traversal) of a band of rows of the same image at
    full spp, sized to about target_s of work on all cores; then one core on a smaller band."""
    from illume_amd._abi import render_params
    from oracle.pyoracle import OracleScene, lib as orc_lib
    threads = orc_lib().orc_threads_available()
    t0 = time.time()
    orc = OracleScene(spec, threads=threads)
    build_s = time.time() - t0
    W, H = st["width"], st["height"]
    p = render_params(**st)

    def band(rows):
        rows = max(1, min(H, rows))
        r0 = max(0, H // 2 - rows // 2)
        return r0, rows, r0 * W, (r0 + rows) * W

    # probe: two rows at full spp
    r0, rows, pb, pe = band(2)
    _, _, segs, secs = orc.render(p, pb, pe, kahan=True, threads=threads, pixels=False)
    per_row = max(secs, 1e-4) / rows
    r0, rows, pb, pe = band(int(target_s / per_row))
    _, _, segs, secs = orc.render(p, pb, pe, kahan=True, threads=threads, pixels=False)
    out = {"value": segs / secs / 1e6, "unit": "Mrays/s", "cores": int(threads), "kind": "port",
           "sample": "oracle FP32 render of rows %d..%d (%d px) x %d spp of the same workload: %d segments in %.2f s;"
                     " reference sweep-BVH build %.1f s not timed" % (r0, r0 + rows, pe - pb, st["spp"], segs, secs, build_s)}
    # one core: a run of pixels from the middle row on, at full spp
    per_px_1core = per_row * threads / W
    n1 = int(min(W * H // 2, max(64, target_1core_s / max(per_px_1core, 1e-9))))
    pb = (H // 2) * W
    _, _, segs1, secs1 = orc.render(p, pb, pb + n1, kahan=True, threads=1, pixels=False)
    out["value_1core"] = segs1 / secs1 / 1e6
    out["sample_1core"] = "%d px from row %d x %d spp: %d segments in %.2f s" % (n1, H // 2, st["spp"], segs1, secs1)
    # the same band with the build's ordered, t-pruned traversal over the same trees
    orc.set_tpruned(True)
    r0, rows, pb, pe = band(rows)
    _, _, segs_t, secs_t = orc.render(p, pb, pe, kahan=True, threads=threads, pixels=False)
    orc.set_tpruned(False)
    out["value_tpruned"] = segs_t / secs_t / 1e6
    out["sample_tpruned"] = ("same %d px x %d spp, ordered t-pruned traversal of the same sweep-SAH trees: %d segments "
                             "in %.2f s" % (pe - pb, st["spp"], segs_t, secs_t))
    model, ncpu = host_cpu()
    out["cpu_model"] = model
    out["host_logical_cpus"] = ncpu
    orc.free()
    return out


def rel_rmse(gpu_sum, ref_sum, spp):
    g = np.asarray(gpu_sum, np.float64) / spp
    o = np.asarray(ref_sum, np.float64) / spp
    return float(np.sqrt(np.mean((g - o) ** 2)) / max(np.mean(np.abs(o)), 1e-12))


class Fp64Oracle:
    """FP64 oracle renders of pixel samples of GPU results (median-split BVH: closest hits do not
    depend on the tree except for ties)."""

    def __init__(self, specs, threads):
        from oracle.pyoracle import OracleScene
        self.threads = threads
        t0 = time.time()
        self.scenes = {k: OracleScene(s, fp64=True, threads=threads, median_bvh=True) for k, s in specs.items()}
        log("fp64 oracle scenes built in %.1f s" % (time.time() - t0))

    def rmse(self, key, params, gpu_acc, npx=128):
        W, H = params.width, params.height
        n = min(npx, W)
        pb = (H // 2) * W + (W - n) // 2
        ref, _, _, _ = self.scenes[key].render(params, pb, pb + n, kahan=True, fp64=True, threads=self.threads,
                                               pixels=False)
        g = np.asarray(gpu_acc).reshape(-1, 3)
        return rel_rmse(g[pb:pb + n], ref, params.spp), "%d px of row %d" % (n, H // 2)

    def free(self):
        for s in self.scenes.values():
            s.free()


def base_line(args, metric, value, unit, ms_per_step, config, dtype="f32"):
    return {"metric": metric, "value": round(value, 3), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def ctx_opts(args):
    o = dict(bvh_layout=args.layout, tri_test=args.tri, paths_in_flight=args.paths, extend_mode=args.extend_mode,
             use_graphs=0 if args.no_graphs else 1, pipes=args.pipes)
    if args.sample_group is not None:
        o["sample_group"] = args.sample_group
    if args.tile_rows is not None:
        o["tile_rows"] = args.tile_rows
    return o


def count_tests(spec, p, shape, args, layout, packets):
    """Mean child-box and triangle tests per closest-hit query, from the extend kernel's own
    counters on a separate render of the same workload (node layout and camera-packet setting
    given; the launch shape otherwise the benchmarked one)."""
    import illume_amd
    opts = ctx_opts(args)
    opts["bvh_layout"] = layout
    with illume_amd.Context(0, count_tests=1, **opts) as cctx:
        sh = dict(shape)
        sh["bvh_layout"] = 1 if layout == "quant8" else 0
        sh["camera_packets"] = 1 if packets else 2
        cctx.tune(**sh)
        cscene = cctx.scene(spec)
        _, _, cinfo = cctx.render(cscene, p, want_accum=False, want_pixels=False)
        cscene.free()
    return cinfo["box_tests"] / max(1, cinfo["segments"]), cinfo["tri_tests"] / max(1, cinfo["segments"])


def bench_plain(args):
    import illume_amd
    from illume_amd import scenes as S
    from illume_amd._abi import render_params

    t0 = time.time()
    spec, st = S.CONFIGS[args.config]()
    log("scene generated in %.2f s" % (time.time() - t0))
    ctx = illume_amd.Context(0, **ctx_opts(args))
    t0 = time.time()
    scene = ctx.scene(spec)
    upload_s = time.time() - t0
    log("scene upload + BVH build %.2f s, %.1f MB on device" % (upload_s, scene.device_bytes / 1e6))
    p = render_params(**st)

    if args.shape:
        ctx.tune(**{k: int(v) for k, v in (kv.split("=") for kv in args.shape.split(","))})
    tuned = None
    if not args.no_tune:
        from illume_amd.autotune import autotune

        def trial(c):
            _, _, i = c.render(scene, p, want_accum=False, want_pixels=False)
            return i["segments"], i["device_ms"]
        tuned, _ = autotune(ctx, trial, repeats=1, log=log, budget_s=args.tune_budget)

    for i in range(args.warmup):
        _, _, info = ctx.render(scene, p, want_accum=False)
        log("warmup %d: %.1f ms, %d segments" % (i, info["device_ms"], info["segments"]))

    # timed steps: each render one replay of its captured graphs
    segs = 0
    lat = []
    dev_ms = 0.0
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, _, info = ctx.render(scene, p, want_accum=False)
        lat.append((time.perf_counter() - ts) * 1e3)
        segs += info["segments"]
        dev_ms += info["device_ms"]
        log("step %d: %.1f ms wall, %.2f ms device, %d segments" % (i, lat[-1], info["device_ms"], info["segments"]))
    wall = time.perf_counter() - t_start
    value = segs / (dev_ms / 1e3) / 1e6
    shape = ctx.settings()
    layout = "quant8" if shape["bvh_layout"] == 1 else "fp32"

    roof = None
    if not args.no_count:
        # profiled pass: the same renders with HIP events around every extend launch (direct
        # launches on the render stream)
        pctx = illume_amd.Context(0, profile=1, **ctx_opts(args))
        pctx.tune(**shape)
        pscene = pctx.scene(spec)
        pctx.render(pscene, p, want_accum=False, want_pixels=False)
        ext_ms, ext_launches, psegs, pdev = 0.0, 0, 0, 0.0
        for i in range(max(1, args.steps)):
            _, _, info = pctx.render(pscene, p, want_accum=False, want_pixels=False)
            ext_ms += info["extend_ms"]
            ext_launches += info["extend_launches"]
            psegs += info["segments"]
            pdev += info["device_ms"]
        pscene.free()
        pctx.close()
        n_box, n_tri = count_tests(spec, p, shape, args, "fp32", packets=False)
        l_box, l_tri = count_tests(spec, p, shape, args, layout, packets=shape["camera_packets"] != 2)
        b_seg = BOX_BYTES["fp32"] * n_box + 36.0 * n_tri + 128.0
        b_lay = BOX_BYTES[layout] * l_box + 36.0 * l_tri + 128.0
        avg_launch_s = ext_ms / 1e3 / max(1, ext_launches)
        per_launch_bytes = b_seg * psegs / max(1, ext_launches)
        achieved = per_launch_bytes / avg_launch_s / 1e9
        ach_lay = b_lay * psegs / max(1, ext_launches) / avg_launch_s / 1e9
        traffic = traffic_from_profiles(args.config, layout)
        lanes = profile_value(args.config, "active_lanes_extend")
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_extend (every extend launch of the render)",
                "convention": "B_seg = 28*N_box + 36*N_tri + 128; N_box, N_tri per closest-hit query of per-ray "
                              "ordered t-pruned walks over the FP32 BVH2, counted on the same workload",
                "b_seg": round(b_seg, 1), "n_box": round(n_box, 3), "n_tri": round(n_tri, 3),
                "avg_launch_ms": round(avg_launch_s * 1e3, 4), "launches": ext_launches,
                "algorithmic_bytes_per_launch": round(per_launch_bytes, 1),
                "segments_per_s_at_roofline_M": round(HBM_PEAK_GBS * 1e9 / b_seg / 1e6, 1),
                "layout": {"bvh_layout": layout, "b_box": BOX_BYTES[layout], "n_box": round(l_box, 3),
                           "n_tri": round(l_tri, 3), "b_seg": round(b_lay, 1), "achieved": round(ach_lay, 2),
                           "frac": round(ach_lay / HBM_PEAK_GBS, 4),
                           "camera_packets": shape["camera_packets"] != 2},
                "profiled_pass": {"renders": max(1, args.steps), "device_ms_per_render": round(pdev / max(1, args.steps), 3),
                                  "extend_ms_per_render": round(ext_ms / max(1, args.steps), 3),
                                  "segments_per_render": psegs // max(1, args.steps)}}
        if traffic is not None:
            hbm = traffic / avg_launch_s / 1e9
            roof["hbm_measured"] = round(hbm, 2)
            roof["hbm_measured_frac"] = round(hbm / HBM_PEAK_GBS, 4)
            roof["traffic_source"] = "profiles/extend_config%d_traffic.txt" % args.config
        if lanes is not None:
            roof["active_lanes"] = lanes
            roof["active_lanes_source"] = "profiles/extend_config%d_counters.txt" % args.config
        log("roofline: FP32 N_box %.2f N_tri %.2f B_seg %.1f B (layout %s: %.2f / %.2f, %.1f B), extend %.2f ms over "
            "%d launches -> %.1f GB/s (%.3f of peak), layout bytes %.1f GB/s"
            % (n_box, n_tri, b_seg, layout, l_box, l_tri, b_lay, ext_ms, ext_launches, achieved,
               achieved / HBM_PEAK_GBS, ach_lay))

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(spec, st)
            log("cpu baseline: %.2f Mrays/s on %d cores, %.3f on one, %.2f with t-pruned traversal"
                % (cpu["value"], cpu["cores"], cpu["value_1core"], cpu["value_tpruned"]))
        except Exception as e:  # the GPU number stands without it
            log("cpu baseline failed: %r" % (e,))

    config = {"workload": workload_name(args.config, spec, st), "width": st["width"], "height": st["height"],
              "spp": st["spp"], "max_depth": st["max_depth"], "bvh_layout": layout, "tri_test": args.tri,
              "segments_per_step": segs // max(1, args.steps), "parallelism": "1 GPU",
              "launch_shape": shape, "autotuned": tuned is not None, "graphs": not args.no_graphs}
    out = base_line(args, METRIC, value, "Mrays/s", dev_ms / args.steps, config)
    out["latency_ms"] = {"p50": round(percentile(lat, 0.5), 3), "p99": round(percentile(lat, 0.99), 3),
                         "form": "plain (one request per ilm_render, image copied to the host)"}
    out["value_pcie_inclusive"] = round(segs / wall / 1e6, 3)
    out["ms_per_step_wall"] = round(wall * 1e3 / args.steps, 3)
    out["roofline"] = roof
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    scene.free()
    ctx.close()


def _upload_mix(ctx, specs):
    t0 = time.time()
    out = {k: ctx.scene(s) for k, s in specs.items()}
    log("uploaded %d scenes in %.1f s" % (len(out), time.time() - t0))
    return out


def bench_batch(args):
    """Batched form: one ilm_render_batch over 64 requests vs the same requests one at a time."""
    import illume_amd
    from illume_amd import scenes as S
    from illume_amd._abi import render_params

    reqs = S.batch_requests()[:args.requests or None]
    specs, sts = {}, {}
    for k in sorted({r["scene"] for r in reqs}):
        specs[k], sts[k] = S.MIX_SCENES[k]()
    params = [render_params(**dict(sts[r["scene"]], width=r["width"], height=r["height"], spp=r["spp"])) for r in reqs]
    results = {}
    for layout in ("fp32", "quant8"):
        opts = ctx_opts(args)
        opts["bvh_layout"] = layout
        with illume_amd.Context(0, **opts) as ctx:
            sc = _upload_mix(ctx, specs)
            ss = [sc[r["scene"]] for r in reqs]
            ctx.render_batch(ss[:4], params[:4])   # warm-up (graph capture)
            t0 = time.perf_counter()
            batch, binfo = ctx.render_batch(ss, params)
            t_batch = time.perf_counter() - t0
            t0 = time.perf_counter()
            plain = []
            psegs = 0
            for s, p in zip(ss, params):
                a, px, i = ctx.render(s, p)
                plain.append((a, px))
                psegs += i["segments"]
            t_plain = time.perf_counter() - t0
            same = sum(int(np.array_equal(b[0], q[0]) and np.array_equal(b[1], q[1])) for b, q in zip(batch, plain))
            results[layout] = dict(batch_s=t_batch, plain_s=t_plain, segments=binfo["segments"], plain_segments=psegs,
                                   identical=same, accs=[b[0] for b in batch])
            log("%s: batch %.2f s (%.1f Mrays/s), plain %.2f s (%.1f Mrays/s), %d/%d identical"
                % (layout, t_batch, binfo["segments"] / t_batch / 1e6, t_plain, psegs / t_plain / 1e6, same, len(reqs)))
            for s in sc.values():
                s.free()
    rm = {}
    if not args.no_rmse:
        from oracle.pyoracle import lib as orc_lib
        orc = Fp64Oracle(specs, orc_lib().orc_threads_available())
        t0 = time.time()
        for layout in results:
            errs = [orc.rmse(r["scene"], p, acc)[0] for r, p, acc in zip(reqs, params, results[layout]["accs"])]
            rm[layout] = {"max": round(max(errs), 5), "mean": round(float(np.mean(errs)), 5)}
        orc.free()
        log("rmse vs fp64 oracle %s (%.1f s)" % (rm, time.time() - t0))
    r = results["quant8"]
    config = {"workload": "config3: %d requests (seed 3) over configs 0-2, side 128..1024, 4..256 spp, "
                          "one ilm_render_batch call, quantised 8-bit BVH child boxes" % len(reqs),
              "requests": len(reqs), "parallelism": "1 GPU", "launch_shape": "library defaults"}
    out = base_line(args, "batched path segments/s (Mrays/s)", r["segments"] / r["batch_s"] / 1e6, "Mrays/s",
                    r["batch_s"] * 1e3, config)
    out["forms"] = {k: {"batched_Mrays_s": round(v["segments"] / v["batch_s"] / 1e6, 2),
                        "plain_Mrays_s": round(v["plain_segments"] / v["plain_s"] / 1e6, 2),
                        "batched_s": round(v["batch_s"], 3), "plain_s": round(v["plain_s"], 3),
                        "speedup": round(v["plain_s"] / v["batch_s"], 3),
                        "identical_to_plain": "%d/%d" % (v["identical"], len(reqs))} for k, v in results.items()}
    out["rmse_vs_fp64"] = rm or None
    out["rmse_sample"] = "128 px of the middle row of every request, same spp, FP64 oracle with the same RNG stream"
    out["roofline"] = None
    out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def _replay(ctx_or_none, trace, start, submit_fn):
    for r in trace:
        dt = start + r["t"] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        submit_fn(r)


def bench_served(args):
    """Served form vs plain one-request-at-a-time renders on a prefix of the Poisson trace (120
    requests by default): throughput, offered vs achieved load, p50/p99 latency (completion -
    arrival), rejections, image identity, RMSE against the FP64 oracle."""
    import hashlib
    import illume_amd
    from illume_amd import scenes as S
    from illume_amd._abi import render_params

    trace = S.serve_trace()[:args.requests or 120]
    t0 = trace[0]["t"]
    trace = [dict(r, t=r["t"] - t0) for r in trace]
    specs, sts = {}, {}
    for k in sorted({r["scene"] for r in trace}):
        specs[k], sts[k] = S.MIX_SCENES[k]()
    params = [render_params(**dict(sts[r["scene"]], width=r["width"], height=r["height"], spp=r["spp"],
                                   max_depth=r["max_depth"])) for r in trace]

    def middle(p, acc):
        n = min(128, p.width)
        pb = (p.height // 2) * p.width + (p.width - n) // 2
        return np.asarray(acc, np.float32).reshape(-1, 3)[pb:pb + n].copy()

    def digest(acc, px):
        h = hashlib.blake2b(digest_size=16)
        h.update(np.ascontiguousarray(acc).tobytes())
        if px is not None:
            h.update(np.ascontiguousarray(px).tobytes())
        return h.hexdigest()

    forms, digests, slices = {}, {}, {}
    with illume_amd.Context(0, **ctx_opts(args)) as ctx:
        sc = _upload_mix(ctx, specs)
        # warm-up: capture the graphs of every depth in the trace
        for d in sorted({p.max_depth for p in params}):
            ctx.render(sc[0], render_params(**dict(sts[0], width=64, height=64, spp=1, max_depth=d)))

        def plain_form(name):
            lat, segs, dig, sl = [], 0, [], []
            start = time.perf_counter()
            for r, p in zip(trace, params):
                dt = start + r["t"] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                a, px, i = ctx.render(sc[r["scene"]], p)
                lat.append((time.perf_counter() - start - r["t"]) * 1e3)
                segs += i["segments"]
                dig.append(digest(a, px))
                sl.append(middle(p, a))
            span = time.perf_counter() - start
            forms[name] = dict(Mrays_s=segs / span / 1e6, makespan_s=span, segments=segs,
                               p50_ms=percentile(lat, 0.5), p99_ms=percentile(lat, 0.99))
            digests[name], slices[name] = dig, sl
            log("%s: %.1f Mrays/s, makespan %.2f s, p50 %.0f ms, p99 %.0f ms"
                % (name, segs / span / 1e6, span, forms[name]["p50_ms"], forms[name]["p99_ms"]))

        def served_form(name, fair):
            """submissions at arrival times; the scheduler packs whatever is admitted into chunks"""
            saved = ctx.settings()
            ctx.tune(sched_fair=fair)
            s0 = ctx.stats()
            handles = []
            start = time.perf_counter()
            _replay(None, trace, start, lambda r: handles.append(ctx.submit(sc[r["scene"]], params[len(handles)])))
            submitted = time.perf_counter() - start
            lat, dig, sl = [], [], []
            for h, p in zip(handles, params):
                h.wait()
                a, px = h.result()
                st = h.poll()
                lat.append(st["end_ms"] - st["submit_ms"])
                dig.append(digest(a, px))
                sl.append(middle(p, a))
                h.free()
            span = time.perf_counter() - start
            s1 = ctx.stats()
            segs = s1["segments"] - s0["segments"]
            ctx.tune(**saved)
            forms[name] = dict(Mrays_s=segs / span / 1e6, makespan_s=span, segments=segs, submit_span_s=submitted,
                               p50_ms=percentile(lat, 0.5), p99_ms=percentile(lat, 0.99),
                               rejected=s1["requests_rejected"] - s0["requests_rejected"])
            digests[name], slices[name] = dig, sl
            log("%s: %.1f Mrays/s, makespan %.2f s, p50 %.0f ms, p99 %.0f ms"
                % (name, segs / span / 1e6, span, forms[name]["p50_ms"], forms[name]["p99_ms"]))

        served_form("served", 0)
        if args.fair:
            served_form("served_fair", 1)
        plain_form("plain")
        for s in sc.values():
            s.free()
    same = sum(int(a == b) for a, b in zip(digests["served"], digests["plain"]))
    rm = None
    if not args.no_rmse:
        from oracle.pyoracle import lib as orc_lib
        orc = Fp64Oracle(specs, orc_lib().orc_threads_available())
        errs, errs32, within, total = [], [], 0, 0
        for r, p, sl in zip(trace, params, slices["served"]):
            n = min(128, p.width)
            pb = (p.height // 2) * p.width + (p.width - n) // 2
            o = orc.scenes[r["scene"]]
            ref, _, _, _ = o.render(p, pb, pb + n, kahan=True, fp64=True, threads=orc.threads, pixels=False)
            r32, _, _, _ = o.render(p, pb, pb + n, kahan=True, fp64=False, threads=orc.threads, pixels=False)
            errs.append(rel_rmse(sl, ref, p.spp))
            errs32.append(rel_rmse(r32, ref, p.spp))
            g = np.asarray(sl, np.float32).reshape(-1, 3)
            rel = np.abs(g - r32) / np.maximum(np.abs(r32), 1e-6)
            within += int(np.all(rel <= 1e-4, axis=1).sum())
            total += n
        orc.free()
        w = int(np.argmax(errs))
        rq = trace[w]
        rm = {"max": round(max(errs), 5), "mean": round(float(np.mean(errs)), 5),
              "worst": {"request": w, "scene": rq["scene"], "width": rq["width"], "height": rq["height"],
                        "spp": rq["spp"], "max_depth": rq["max_depth"],
                        "fp32_oracle_rmse_vs_fp64": round(errs32[w], 5)},
              "fp32_oracle_max": round(max(errs32), 5),
              "pixels_within_1e-4_of_fp32_oracle": round(within / total, 5)}
        log("rmse vs fp64 oracle: %s" % rm)
    sv, pl = forms["served"], forms["plain"]
    arrival_span = max(trace[-1]["t"], 1e-9)
    offered = pl["segments"] / arrival_span / 1e6
    config = {"workload": "config4: first %d requests of the Poisson trace (seed 4, 20 req/s) over configs 0-2 and "
                          "the 12-instance 722k-tri mesh, 256^2..1920x1080, 8..512 spp, depth 8..32" % len(trace),
              "requests": len(trace), "parallelism": "1 GPU"}
    out = base_line(args, "served path segments/s (Mrays/s); p50/p99 request latency", sv["Mrays_s"], "Mrays/s",
                    sv["makespan_s"] * 1e3, config)
    out["latency_ms"] = {"p50": round(sv["p50_ms"], 1), "p99": round(sv["p99_ms"], 1),
                         "form": "served, arrival-order packing (completion - arrival)",
                         "plain": {"p50": round(pl["p50_ms"], 1), "p99": round(pl["p99_ms"], 1)}}
    out["load"] = {"arrival_span_s": round(arrival_span, 3), "offered_Mrays_s": round(offered, 1),
                   "achieved_Mrays_s": round(sv["Mrays_s"], 1), "offered_over_achieved": round(offered / sv["Mrays_s"], 3),
                   "rejected": sv["rejected"]}
    out["forms"] = {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in forms.items()}
    out["served_vs_plain"] = round(sv["Mrays_s"] / pl["Mrays_s"], 3)
    out["identical_to_plain"] = "%d/%d" % (same, len(trace))
    out["rmse_vs_fp64"] = rm
    out["roofline"] = None
    out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_obj_load(args):
    """OBJ loader on the 12-instance scene's 722,000-triangle displaced icosphere (n=190, seed 5)
    written as an OBJ by the generator: parse and vertex-merge seconds (host only; the reference's
    reader is src/primitives/mesh.cu:159-340), and that the triangles read back bit for bit."""
    import tempfile
    import illume_amd
    from illume_amd import scenes as S
    m = S.displaced_icosphere(190, 5)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mesh.obj")
        t0 = time.time()
        S.write_obj(m, path)
        out["write_s"] = round(time.time() - t0, 3)
        out["obj_bytes"] = os.path.getsize(path)
        with open(os.path.join(d, "scene.json"), "w") as f:
            json.dump({"meshes": [{"name": "m", "file": "mesh.obj"}], "camera": {"fov": 40}}, f)
        times = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            js = illume_amd.load_json(os.path.join(d, "scene.json"))
            wall = time.perf_counter() - t0
            parse_s, merge_s = js.load_times()
            times.append((wall, parse_s, merge_s))
            dm = js.desc.meshes[0]
            pos = np.ctypeslib.as_array(dm.positions, shape=(dm.num_vertices * 3,)).reshape(-1, 3)
            idx = np.ctypeslib.as_array(dm.indices, shape=(dm.num_triangles * 3,)).reshape(-1, 3)
            exact = bool(np.array_equal(pos[idx], m["positions"][m["indices"]]))
            js.free()
    best = min(times)
    model, ncpu = host_cpu()
    line = {"metric": "OBJ load seconds (722,000 triangles)", "value": round(best[0], 3), "unit": "s",
            "higher_is_better": False, "parse_s": round(best[1], 3), "merge_s": round(best[2], 3),
            "triangles": int(len(m["indices"])), "vertices": int(len(m["positions"])),
            "triangles_bit_exact": exact, "runs": len(times), "cpu_model": model, "threads": 1}
    line.update(out)
    print(json.dumps(line), flush=True)


def calibrate_copy():
    """A plain device copy of known size (1 GiB read + 1 GiB written, three times) for scaling
    the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc to bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--layout", default="quant8", choices=["fp32", "quant8"],
                    help="node layout the run starts from (the autotuner may switch it)")
    ap.add_argument("--tri", default="mt", choices=["mt", "watertight"])
    ap.add_argument("--paths", type=int, default=0, help="paths in flight per chunk (0: library default)")
    ap.add_argument("--extend-mode", type=int, default=0,
                    help="0: library default (per-step refill), 1: one ray batch per wave, 2: refill at leaves, 3: per-step refill")
    ap.add_argument("--sample-group", type=int, default=None)
    ap.add_argument("--tile-rows", type=int, default=None)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library default launch shape)")
    ap.add_argument("--shape", default=None,
                    help="launch shape to apply before measuring, as key=value[,key=value] over the tunables "
                         "(e.g. ext_fetch=128,leaf_wait_pct=50); with --no-tune this reproduces a tuned run")
    ap.add_argument("--tune-budget", type=float, default=30.0)
    ap.add_argument("--requests", type=int, default=0, help="configs 3/4: number of requests (0: 64 / 120)")
    ap.add_argument("--fair", action="store_true", help="config 4: also run the fair-share packing")
    ap.add_argument("--no-rmse", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--pipes", type=int, default=0, choices=[0, 1, 2],
                    help="chunk pipelines (0: library default, two; 1: one stream)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel directly (for counter collection); the default replays hipGraphs")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run the 1 GiB copy used to scale HBM counters")
    ap.add_argument("--obj-load", action="store_true", help="only time the OBJ loader on the 722k-triangle mesh (host)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.obj_load:
        return bench_obj_load(args)
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by request or sample range; "
              "multi-GPU runs are not part of it). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.config in (0, 1, 2):
        bench_plain(args)
    elif args.config == 3:
        bench_batch(args)
    elif args.config == 4:
        bench_served(args)
    else:
        sys.exit("unknown config %d" % args.config)


if __name__ == "__main__":
    main()
