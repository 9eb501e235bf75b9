"""thundervision_amd: the stereo-matching hot path of ThunderVision (matching
cost -> SGM / scanline DP -> winner-take-all) as hand-written HIP kernels for
MI355X (gfx950) behind a C ABI (include/tdvh.h).

Python access goes through `thundervision_amd.tdvh` (ctypes over libtdvh.so);
the C++ stage layer mirroring the reference's MatchingCost / Optimizer API is
in thundervision_amd/csrc/stages.
"""
from . import synth  # noqa: F401

__all__ = ["synth", "tdvh"]


def __getattr__(name):
    # tdvh imports torch; load it on first use
    if name == "tdvh":
        import importlib
        return importlib.import_module(__name__ + ".tdvh")
    raise AttributeError(name)
