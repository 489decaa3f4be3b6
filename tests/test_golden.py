"""Oracle against the committed fixtures in tests/golden (see make_golden.py there)."""
import json
import os

import numpy as np

from oracle.pyoracle import lib, OracleScene
import illume_amd.scenes as S
from illume_amd._abi import render_params

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_pcg32_published_outputs_fixture():
    import ctypes as C
    d = json.load(open(os.path.join(G, "pcg32_known_answers.json")))
    out = np.empty(len(d["outputs"]), np.uint32)
    lib().orc_pcg32(C.c_uint64(d["seed"]), C.c_uint64(d["seq"]), len(out), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert [int(x, 16) for x in d["outputs"]] == [int(v) for v in out]


def test_oracle_fp64_render_regression():
    f = np.load(os.path.join(G, "oracle_cornell16.npz"))
    spec, st = S.cornell()
    p = render_params(**dict(st, width=16, height=16, spp=4))
    orc = OracleScene(spec, fp64=True)
    acc, px, segs, _ = orc.render(p, kahan=True, fp64=True)
    orc.free()
    assert segs == int(f["segments"][0])
    assert np.array_equal(acc, f["accum"])
    assert np.array_equal(px, f["pixels"])
