"""Seeded synthetic stereo pairs (libtdvsynth.so, csrc/synth.c).

The workload generator for tests and the benchmark: blurred uniform noise,
rectangle disparity layout, forward warp with occlusion fill, sigma-2 noise.
"""
import ctypes as ct
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_PKG, "libtdvsynth.so")
        if not os.path.exists(path):
            raise RuntimeError("%s is missing: run `python -m thundervision_amd.build`" % path)
        h = ct.CDLL(path)
        h.tdv_synth_pair.restype = ct.c_int
        h.tdv_synth_pair.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_uint64,
                                     ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        h.tdv_synth_uniform_int.restype = ct.c_int
        h.tdv_synth_uniform_int.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_int, ct.c_int]
        h.tdv_synth_exponential.restype = ct.c_double
        h.tdv_synth_exponential.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_double]
        h.tdv_synth_hash.restype = ct.c_uint64
        h.tdv_synth_hash.argtypes = [ct.c_void_p, ct.c_size_t]
        _LIB = h
    return _LIB


def pair(width, height, dmax, nrect, seed, frame=0):
    """Returns (left, right, truth): uint8, uint8, int16 arrays of shape (H, W)."""
    L = np.empty((height, width), np.uint8)
    R = np.empty((height, width), np.uint8)
    gt = np.empty((height, width), np.int16)
    rc = _lib().tdv_synth_pair(width, height, dmax, nrect, seed, frame, L.ctypes.data,
                               R.ctypes.data, gt.ctypes.data)
    if rc != 0:
        raise ValueError("tdv_synth_pair failed (%d)" % rc)
    return L, R, gt


def uniform_int(seed, stream, lo, hi):
    return _lib().tdv_synth_uniform_int(seed, stream, lo, hi)


def exponential(seed, stream, rate):
    return _lib().tdv_synth_exponential(seed, stream, rate)


def hash_bytes(a):
    a = np.ascontiguousarray(a)
    return _lib().tdv_synth_hash(a.ctypes.data, a.nbytes)


# The workloads (image size, disparity range of the truth, rectangles, depth,
# seed) used by the tests and the benchmark.
PLAIN_384 = dict(width=384, height=288, dmax=15, nrect=8, depth=16, seed=1)
PLAIN_720P = dict(width=1280, height=720, dmax=127, nrect=12, depth=128, seed=2)
WINDOW_1080P = dict(width=1920, height=1080, dmax=255, nrect=12, depth=256, seed=3)


def workload_pair(spec, frame=0):
    return pair(spec["width"], spec["height"], spec["dmax"], spec["nrect"], spec["seed"], frame)


def batch_specs(n=64, seed_base=1000):
    """Mixed batch: resolution uniform over 4 sizes, D uniform over
    {32, 64, 128, 256} clamped to <= W/4, seed seed_base + i."""
    sizes = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]
    depths = [32, 64, 128, 256]
    out = []
    for i in range(n):
        s = seed_base + i
        W, H = sizes[uniform_int(s, 1, 0, 3)]
        D = min(depths[uniform_int(s, 2, 0, 3)], W // 4)
        out.append(dict(width=W, height=H, dmax=D - 1, nrect=12, depth=D, seed=s))
    return out
