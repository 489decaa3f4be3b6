"""Benchmark of the path-tracer hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1] [--no-cpu]

A step is one full render of the workload (all passes and depths) through the C ABI
(ilm_render: ray generation, BVH extend, shade + compaction per depth, Kahan accumulation,
tonemap and pixel readback), with the scene and its BVH already resident in HBM.
Prints ONE JSON line on stdout; progress goes to stderr.

value       segments/s / 1e6 over the timed steps (wall clock around the synchronous calls);
            a segment is one closest-hit query (src/renderer.cu:554 counts the same)
roofline    extend kernel: algorithmic bytes B_seg = 28*N_box + 36*N_tri + 128 per segment,
            N_box / N_tri the mean child-box / triangle tests per query counted by the kernel's
            own counters on the same workload; achieved = bytes / summed extend time (HIP
            events around every extend launch on the render stream); peak 8 TB/s HBM3E
cpu_baseline  the oracle (CPU restatement of the reference per-path routine: exhaustive
            traversal of the reference's sweep-built BVH, linear instance loop) on a band of
            the same image, OpenMP over pixels on the host's cores
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


def traffic_from_profiles(cfg):
    """HBM bytes per extend launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration), recorded in profiles/extend_config<cfg>_traffic.txt."""
    path = os.path.join(ROOT, "profiles", "extend_config%d_traffic.txt" % cfg)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(spec, st, target_s=12.0):
    from illume_amd._abi import render_params
    from oracle.pyoracle import OracleScene, lib as orc_lib
    threads = orc_lib().orc_threads_available()
    t0 = time.time()
    orc = OracleScene(spec, threads=threads)
    build_s = time.time() - t0
    W, H = st["width"], st["height"]
    # a band of rows through the middle of the image (where the geometry is)
    rows = max(1, H // 8)
    r0 = H // 2 - rows // 2
    pb, pe = r0 * W, (r0 + rows) * W
    p = render_params(**dict(st, spp=1))
    _, _, segs1, secs1 = orc.render(p, pb, pe, kahan=True, threads=threads, pixels=False)
    spp = int(max(1, min(st["spp"], target_s / max(secs1, 1e-3))))
    p = render_params(**dict(st, spp=spp))
    _, _, segs, secs = orc.render(p, pb, pe, kahan=True, threads=threads, pixels=False)
    orc.free()
    return {"value": segs / secs / 1e6, "unit": "Mrays/s", "cores": int(threads), "kind": "port",
            "sample": "oracle FP32 render of rows %d..%d (%d px) x %d spp of the same workload: %d segments in %.2f s;"
                      " reference sweep-BVH build %.1f s not timed" % (r0, r0 + rows, pe - pb, spp, segs, secs, build_s)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--layout", default="fp32", choices=["fp32", "quant8"])
    ap.add_argument("--tri", default="mt", choices=["mt", "watertight"])
    ap.add_argument("--paths", type=int, default=0, help="paths in flight per chunk (0: library default)")
    ap.add_argument("--extend-mode", type=int, default=0, help="0: one ray batch per wave, 1: lane refill")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards by request or sample range; "
              "multi-GPU runs are not part of it). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import illume_amd
    from illume_amd import scenes as S
    from illume_amd._abi import render_params

    t0 = time.time()
    spec, st = S.CONFIGS[args.config]()
    log("scene generated in %.2f s" % (time.time() - t0))
    ctx = illume_amd.Context(0, bvh_layout=args.layout, tri_test=args.tri, profile=1, paths_in_flight=args.paths,
                             extend_mode=args.extend_mode)
    t0 = time.time()
    scene = ctx.scene(spec)
    upload_s = time.time() - t0
    log("scene upload + BVH build %.2f s, %.1f MB on device" % (upload_s, scene.device_bytes / 1e6))
    p = render_params(**st)

    for i in range(args.warmup):
        _, _, info = ctx.render(scene, p, want_accum=False)
        log("warmup %d: %.1f ms, %d segments" % (i, info["device_ms"], info["segments"]))

    segs = 0
    ext_ms = 0.0
    ext_launches = 0
    lat = []
    dev_ms = 0.0
    t_start = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        _, _, info = ctx.render(scene, p, want_accum=False)
        lat.append((time.perf_counter() - ts) * 1e3)
        segs += info["segments"]
        ext_ms += info["extend_ms"]
        ext_launches += info["extend_launches"]
        dev_ms += info["device_ms"]
        log("step %d: %.1f ms wall, %.1f ms device, %d segments, extend %.1f ms" % (i, lat[-1], info["device_ms"],
                                                                                    info["segments"], info["extend_ms"]))
    wall = time.perf_counter() - t_start
    value = segs / wall / 1e6

    # traversal test counts on the same workload (a separate, untimed render)
    roof = None
    if not args.no_count:
        cctx = illume_amd.Context(0, bvh_layout=args.layout, tri_test=args.tri, count_tests=1,
                                  paths_in_flight=args.paths, extend_mode=args.extend_mode)
        cscene = cctx.scene(spec)
        _, _, cinfo = cctx.render(cscene, p, want_accum=False, want_pixels=False)
        n_box = cinfo["box_tests"] / max(1, cinfo["segments"])
        n_tri = cinfo["tri_tests"] / max(1, cinfo["segments"])
        cscene.free()
        cctx.close()
        b_seg = 28.0 * n_box + 36.0 * n_tri + 128.0
        per_launch_bytes = b_seg * segs / max(1, ext_launches)
        avg_launch_s = ext_ms / 1e3 / max(1, ext_launches)
        achieved = per_launch_bytes / avg_launch_s / 1e9
        traffic = traffic_from_profiles(args.config)
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_extend", "b_seg": round(b_seg, 1), "n_box": round(n_box, 2), "n_tri": round(n_tri, 2),
                "avg_launch_ms": round(avg_launch_s * 1e3, 4), "launches": ext_launches,
                "algorithmic_bytes_per_launch": round(per_launch_bytes, 1)}
        log("roofline: N_box %.2f N_tri %.2f B_seg %.1f B, extend %.1f ms over %d launches -> %.1f GB/s"
            % (n_box, n_tri, b_seg, ext_ms, ext_launches, achieved))

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(spec, st)
            log("cpu baseline: %.2f Mrays/s on %d cores" % (cpu["value"], cpu["cores"]))
        except Exception as e:  # the GPU number stands without it
            log("cpu baseline failed: %r" % (e,))

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload_name(args.config, spec, st), "width": st["width"], "height": st["height"],
                   "spp": st["spp"], "max_depth": st["max_depth"], "bvh_layout": args.layout, "tri_test": args.tri,
                   "segments_per_step": segs // max(1, args.steps), "parallelism": "1 GPU"},
        "latency_ms": {"p50": round(percentile(lat, 0.5), 3), "p99": round(percentile(lat, 0.99), 3),
                       "form": "plain (one request per ilm_render)"},
        "value_device": round(segs / (dev_ms / 1e3) / 1e6, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    scene.free()
    ctx.close()


if __name__ == "__main__":
    main()
