"""The autotuner's search logic, with a stand-in context whose speed is a known function of
the settings (no GPU needed)."""
from illume_amd.autotune import autotune


class FakeCtx:
    def __init__(self):
        self.s = dict(bvh_layout=0, camera_packets=0, extend_mode=0, refill_min=0, ext_fetch=0, leaf_wait_pct=0, shade_sort=0, sched_fair=0, sample_group=0, tile_rows=0, chunk_passes=0, paths_in_flight=0,
                      lone_split=0)
        self.tunes = 0

    def settings(self):
        return dict(self.s)

    def tune(self, **kw):
        self.s.update(kw)
        self.tunes += 1


def speed(s):
    # best at sample_group 16, tile_rows 8, extend_mode 0, paths 2^25; separable
    v = 100.0
    v -= abs((s["sample_group"] or 1).bit_length() - 5) * 7
    v -= abs((s["tile_rows"] or 1) - 8) * 2
    v -= 15 * s["extend_mode"]
    v -= abs(((s["paths_in_flight"] or (1 << 26)).bit_length() - 1) - 25) * 3
    return v


def test_coordinate_descent_finds_separable_optimum():
    ctx = FakeCtx()

    def run(c):
        return speed(c.settings()) * 1000.0, 1.0

    best, trials = autotune(ctx, run, repeats=1)
    assert best["sample_group"] == 16 and best["tile_rows"] == 8
    assert best["extend_mode"] == 0 and best["paths_in_flight"] == 1 << 25
    assert ctx.settings() == best               # the context is left tuned
    rates = [t["rate"] for t in trials]
    assert max(rates) == speed(best) * 1000.0


def test_budget_stops_search():
    ctx = FakeCtx()
    best, trials = autotune(ctx, lambda c: (1.0, 1.0), repeats=1, budget_s=0.0)
    assert len(trials) == 1 and best == FakeCtx().settings()


def test_gains_within_noise_do_not_move_settings():
    """A candidate must beat the current choice by more than min_gain: a 0.3 % better layout is
    left alone (run-to-run noise), a 5 % better one is taken."""
    for gain, want in [(0.003, 0), (0.05, 1)]:
        ctx = FakeCtx()

        def run(c):
            return 1000.0 * (1.0 + gain * c.settings()["bvh_layout"]), 1.0

        best, _ = autotune(ctx, run, repeats=1)
        assert best["bvh_layout"] == want
