"""Writes tests/golden/real_pairs.npz from image files that the reference
ships in res/ (tsukuba_L/R.png, Benchmark/Teddy/q_left/right/true.png): the
decoded pixels only (uint8 RGB pairs; Teddy rows 160-415 of the 512x512
images and of its ground truth), loaded with numpy.load(allow_pickle=False)
by the tests. Run from the repository root:

    python tests/golden/make_fixtures.py /root/reference/res
"""
import os
import sys

import numpy as np
from PIL import Image


def load(path, mode):
    return np.ascontiguousarray(np.asarray(Image.open(path).convert(mode)), dtype=np.uint8)


def main(res):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "real_pairs.npz")
    teddy = os.path.join(res, "Benchmark", "Teddy")
    np.savez_compressed(
        out,
        tsukuba_left=load(os.path.join(res, "tsukuba_L.png"), "RGB"),
        tsukuba_right=load(os.path.join(res, "tsukuba_R.png"), "RGB"),
        teddy_left=load(os.path.join(teddy, "q_left.png"), "RGB")[160:416].copy(),
        teddy_right=load(os.path.join(teddy, "q_right.png"), "RGB")[160:416].copy(),
        teddy_true=load(os.path.join(teddy, "q_true.png"), "L")[160:416].copy())
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/res")
