"""Writes the fixtures in this folder.

pcg32_known_answers.json  the first outputs of the published PCG32 demonstration program
                          (pcg-random.org, pcg32-demo with seed 42 and sequence 54); typed in
                          from the published listing, not computed here.
oracle_cornell16.npz      an FP64 oracle render of the Cornell scene at 16x16, 4 spp, depth 8:
                          a regression fixture for the oracle itself (it is the checker of the
                          GPU path, so an accidental change to it must show up).

Run from the repository root after building the oracle:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PCG32_DEMO_42_54 = ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b", "0xcbed606e"]


def main():
    with open(os.path.join(HERE, "pcg32_known_answers.json"), "w") as f:
        json.dump({"seed": 42, "seq": 54, "outputs": PCG32_DEMO_42_54,
                   "source": "pcg32-demo output listing, pcg-random.org"}, f, indent=1)
    from oracle.pyoracle import OracleScene
    from illume_amd import scenes as S
    from illume_amd._abi import render_params
    spec, st = S.cornell()
    p = render_params(**dict(st, width=16, height=16, spp=4))
    orc = OracleScene(spec, fp64=True)
    acc, px, segs, _ = orc.render(p, kahan=True, fp64=True)
    orc.free()
    np.savez_compressed(os.path.join(HERE, "oracle_cornell16.npz"), accum=acc, pixels=px,
                        segments=np.array([segs], np.int64))
    print("wrote fixtures; segments", segs)


if __name__ == "__main__":
    main()
