"""ilm_render_noalloc against ilm_render through the C ABI: the same accumulated sums, pixels and
segment count, bit for bit, on two calls in a row with the second smaller than the first (the
static scratch array keeps the first call's contents), on a scene with spheres and instances
and on a scene whose BVH needs the spilling traversal stack; an image beyond the scratch array
is refused before any device work."""
import numpy as np
import pytest

import illume_amd
import illume_amd.scenes as S
from illume_amd._abi import render_params

pytestmark = pytest.mark.gpu


def _p(st, **kw):
    d = dict(st)
    d.update(kw)
    return render_params(**d)


def _same(ctx, scene, p):
    ref, ref_px, ri = ctx.render(scene, p)
    got, px, gi = ctx.render_noalloc(scene, p)
    print("noalloc %dx%dx%d: %d segments in %d chunks (plain %d chunk(s))"
          % (p.width, p.height, p.spp, gi["segments"], gi["chunks"], ri["chunks"]))
    assert gi["segments"] == ri["segments"] > 0
    assert np.array_equal(got, ref)
    assert np.array_equal(px, ref_px)
    return gi


def test_noalloc_equals_plain_render_twice_second_smaller():
    spec, st = S.gold_spheres(n=20)
    with illume_amd.Context(0) as c:
        s = c.scene(spec)
        # 1024x768x16: 12.6 M paths, several chunks of what the array holds
        big = _same(c, s, _p(st, width=1024, height=768, spp=16, max_depth=8))
        assert big["chunks"] > 1
        small = _same(c, s, _p(st, width=96, height=64, spp=3, max_depth=6, first_sample=5, tonemap=1))
        assert small["chunks"] == 1
        s.free()


def test_noalloc_equals_plain_render_with_spilling_stack():
    """The 524,880-triangle subsurface mesh under FP32 nodes is deeper than the LDS stack, so
    the extend kernels spill to the area carved from the scratch array (with a smaller grid)."""
    spec, st = S.subsurface(env_size=(256, 128))
    with illume_amd.Context(0) as c:
        s = c.scene(spec)
        _same(c, s, _p(st, width=160, height=120, spp=4))
        _same(c, s, _p(st, width=40, height=30, spp=2, max_depth=12))
        s.free()


def test_noalloc_rejects_images_beyond_the_scratch_array():
    spec, st = S.cornell()
    with illume_amd.Context(0) as c:
        s = c.scene(spec)
        with pytest.raises(illume_amd.IlmError) as e:
            c.render_noalloc(s, _p(st, width=4096, height=4096, spp=1))
        assert e.value.code == -9
        # the context still renders afterwards
        _same(c, s, _p(st, width=64, height=64, spp=2))
        s.free()
