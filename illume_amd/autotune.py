"""Autotuner for the launch shape of the wavefront pipeline.

The tunables change how paths are laid out, queued and walked, never what is computed: every
choice gives bit-identical images (samples are summed per pixel in pass order whatever the path
order or chunking, and a ray's closest hit does not depend on the order its walk meets
triangles: exact distance ties go to the smaller instance and primitive). The tuner only
measures time.

    bvh_layout       0: FP32 BVH2 nodes, 1: 8-bit BVH4 nodes (both are resident for every scene)
    extend_mode      1: each wave traces one batch of 64 rays to completion
                     2: idle lanes refill from a wave-local queue at BLAS leaves
                     3 (and 0, the default): idle lanes refill after every traversal step
    refill_min       idle lanes that trigger a refill (modes 1 and 2)
    ext_fetch        rays an extend wave reserves from the queue at a time (work-queue granularity)
    leaf_wait_pct    per-step refill: lanes waiting at BLAS leaves (as a percentage of the other
                     busy lanes) before a wave spends a trip on triangle tests
    shade_sort       1: each 512-path shading block past the first depth runs its paths grouped by
                     material stand-in (staged through LDS), 2: in queue order
    camera_packets   1: the first depth traces each wave's 64 camera rays as one packet, 2: per lane
    sample_group     consecutive path ids per pixel (ray coherence of the first bounces)
    tile_rows        pixel walk in bands of this many rows, column by column (wave footprint)
    paths_in_flight  paths per chunk (queue size): larger chunks amortise launch tails,
                     smaller ones bound the memory of the path state
    chunk_passes     passes of one request per chunk (0: as many as fit)
    lone_split       1: a lone request renders as one chunk, 2: as two chunks of half its pixels
                     on the two pipelines (0, the default: two when its max_depth is 24 or more)

The search is coordinate descent from the context's current settings: each knob in turn is set
to every candidate value with the others held, and the fastest is kept if it beats the current
choice by more than min_gain (run-to-run noise does not move the settings). The objective is device
time per segment of a caller-supplied trial (normally the workload itself or a reduced-spp copy
of it), the best of `repeats` runs.
"""
import time

# A knob is a setting name, or a tuple of names searched jointly (refill_min only matters for
# the refill modes, so it is searched together with extend_mode).
DEFAULT_SPACE = (
    ("bvh_layout", (0, 1)),
    ("sample_group", (1, 4, 16, 64)),
    ("tile_rows", (1, 4, 8)),
    (("extend_mode", "refill_min"), ((1, 0), (2, 16), (2, 32), (3, 8), (3, 16), (3, 32), (3, 48))),
    ("leaf_wait_pct", (40, 50, 60, 80, 100, 150)),
    ("shade_sort", (1, 2)),
    ("camera_packets", (1, 2)),
    ("ext_fetch", (128, 256, 512, 1024, 2048)),
    ("paths_in_flight", (1 << 24, 1 << 25, 1 << 26)),
    ("lone_split", (1, 2)),
)


def _get(cur, knob):
    return tuple(cur[k] for k in knob) if isinstance(knob, tuple) else cur[knob]


def _set(cur, knob, v):
    out = dict(cur)
    if isinstance(knob, tuple):
        out.update(zip(knob, v))
    else:
        out[knob] = v
    return out


def _measure(ctx, run, repeats):
    best = None
    for _ in range(repeats):
        segs, ms = run(ctx)
        rate = segs / max(ms, 1e-6)
        best = rate if best is None else max(best, rate)
    return best


def autotune(ctx, run, space=DEFAULT_SPACE, repeats=2, log=None, budget_s=60.0, min_gain=0.02):
    """Tune ctx in place. run(ctx) renders the trial workload and returns (segments, device_ms).
    Returns (best settings dict, list of trial records)."""
    t0 = time.time()
    cur = ctx.settings()
    run(ctx)   # warm-up: graph capture, first-touch of buffers
    best_rate = _measure(ctx, run, repeats)
    trials = [dict(cur, rate=best_rate)]
    if log:
        log("autotune start %s: %.1f Mseg/s" % (cur, best_rate / 1e3))
    for knob, values in space:
        for v in values:
            if time.time() - t0 > budget_s:
                break
            if v == _get(cur, knob):
                continue
            cand = _set(cur, knob, v)
            ctx.tune(**cand)
            run(ctx)
            rate = _measure(ctx, run, repeats)
            trials.append(dict(cand, rate=rate))
            if log:
                log("autotune %s=%s: %.1f Mseg/s" % (knob, v, rate / 1e3))
            if rate > best_rate * (1.0 + min_gain):
                best_rate, cur = rate, cand
        ctx.tune(**cur)
    if log:
        log("autotune pick %s: %.1f Mseg/s (%d trials, %.1f s)" % (cur, best_rate / 1e3, len(trials), time.time() - t0))
    return cur, trials
