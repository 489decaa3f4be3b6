"""Image error of the FP32 GPU renders against the FP64 oracle with the same RNG stream, per
configuration (reduced sizes so the FP64 oracle finishes in seconds). The difference is pure
precision: where FP32 and FP64 take different branches a path diverges, which is what the
relative RMSE measures. Stated bar: relative RMSE <= 1% of the mean pixel value."""
import numpy as np
import pytest

import illume_amd.scenes as S
from illume_amd._abi import render_params
from oracle.pyoracle import OracleScene

pytestmark = pytest.mark.gpu

CASES = [
    ("config0 cornell 256x256x16 d8", S.cornell, dict()),
    ("config1 gold 128x128x16 d16", S.gold_spheres, dict(width=128, height=128, spp=16)),
    ("config2 subsurface 120x90x8 d45", S.subsurface, dict(width=120, height=90, spp=8)),
    ("instanced 160x90x8 d16", S.instanced, dict(width=160, height=90, spp=8)),
]


def rel_rmse(a, b):
    a = np.asarray(a, np.float64).reshape(-1, 3)
    b = np.asarray(b, np.float64).reshape(-1, 3)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.mean(np.abs(b)), 1e-12))


@pytest.fixture(scope="module")
def oracles():
    return {}


@pytest.mark.parametrize("name,maker,kw", CASES, ids=[c[0].split()[0] for c in CASES])
@pytest.mark.parametrize("variant", ["fp32", "quant8", "watertight"])
def test_rmse_against_fp64_oracle(name, maker, kw, variant, ctx, ctx_quant, ctx_wt, oracles):
    spec, st = maker()
    d = dict(st)
    d.update(kw)
    p = render_params(**d)
    c = {"fp32": ctx, "quant8": ctx_quant, "watertight": ctx_wt}[variant]
    sc = c.scene(spec)
    acc, _, info = c.render(sc, p)
    sc.free()
    if name not in oracles:
        orc = OracleScene(spec, fp64=True, median_bvh=True)
        ref, _, segs, secs = orc.render(p, kahan=True, fp64=True, pixels=False)
        orc.free()
        oracles[name] = ref
    ref = oracles[name]
    e = rel_rmse(acc / p.spp, ref / p.spp)
    print("%s [%s]: relative RMSE vs FP64 oracle %.5f (%d segments)" % (name, variant, e, info["segments"]))
    assert e <= 0.01
