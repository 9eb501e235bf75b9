"""ctypes binding of the C ABI in include/tdvh.h, with torch tensors as the
device-memory carrier.

Every function here calls the HIP library libtdvh.so; there is no CPU or
PyTorch fallback. If the library is missing, `lib()` raises.
"""
import ctypes as ct
import os

import torch

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U16, U32, F32, I32 = 1, 2, 3, 4
LAYOUT_STRIDED, LAYOUT_SPLIT16 = 0, 1
SGM_REF_GEOMETRY, SGM_START_ASSIGN, SGM_MIN_SKIP_LAST = 1, 2, 4
SGM_REF_ALL = SGM_REF_GEOMETRY | SGM_START_ASSIGN | SGM_MIN_SKIP_LAST
WTA_LOWEST, WTA_TREE = 0, 1
DP_START_ZERO, DP_REF_BACKTRACK, DP_COST_BOUNDED = 1, 2, 4
FUSE_STAGED, FUSE_COST = 0, 1
MAX_DEPTH = 256      # SGM (the reference's limit, semiglobal.cu:44)
DP_MAX_DEPTH = 512   # DP and the matching costs


class TdvhError(RuntimeError):
    """A tdvh call returned a negative status."""

    def __init__(self, status, message):
        super().__init__("tdvh error %d: %s" % (status, message))
        self.status = status


class DSI(ct.Structure):
    _fields_ = [("ptr", ct.c_void_p), ("width", ct.c_int32), ("height", ct.c_int32),
                ("depth", ct.c_int32), ("dtype", ct.c_int32), ("layout", ct.c_int32),
                ("stride_x", ct.c_int64), ("stride_y", ct.c_int64)]


class SgmParams(ct.Structure):
    _fields_ = [("p1", ct.c_int32), ("p2", ct.c_int32), ("dir_mask", ct.c_int32),
                ("flags", ct.c_int32), ("fp1", ct.c_float), ("fp2", ct.c_float)]


class Pair(ct.Structure):
    _fields_ = [("left", ct.c_void_p), ("right", ct.c_void_p), ("width", ct.c_int32),
                ("height", ct.c_int32), ("pitch", ct.c_int32), ("depth", ct.c_int32),
                ("disp", ct.c_void_p)]


class SgPath(ct.Structure):
    _fields_ = [("sx", ct.c_uint16), ("sy", ct.c_uint16), ("ex", ct.c_uint16),
                ("ey", ct.c_uint16), ("dx", ct.c_int16), ("dy", ct.c_int16),
                ("size", ct.c_uint16), ("pad", ct.c_uint16)]


# name -> (restype, argtypes) for every symbol of include/tdvh.h
_VP, _I, _U32, _SZ, _I64 = ct.c_void_p, ct.c_int, ct.c_uint32, ct.c_size_t, ct.c_int64
SIGNATURES = {
    "tdvh_last_error": (ct.c_char_p, []),
    "tdvh_ctx_create": (_I, [_I, ct.POINTER(_VP)]),
    "tdvh_ctx_destroy": (_I, [_VP]),
    "tdvh_ctx_trim": (_I, [_VP]),
    "tdvh_ctx_scratch_bytes": (_SZ, [_VP]),
    "tdvh_profile_enable": (_I, [_VP, _I]),
    "tdvh_profile_reset": (_I, [_VP]),
    "tdvh_profile_read": (_I, [_VP, _I, ct.c_char_p, _SZ, ct.POINTER(ct.c_double),
                               ct.POINTER(_I64)]),
    "tdvh_dsi_bytes": (_SZ, [_I, _I, _I, _I, _I]),
    "tdvh_dsi_init": (_I, [ct.POINTER(DSI), _VP, _I, _I, _I, _I, _I]),
    "tdvh_dsi_convert": (_I, [_VP, ct.POINTER(DSI), ct.POINTER(DSI), _VP]),
    "tdvh_cost_bt": (_I, [_VP, _VP, _VP, _I, _I, _I, ct.POINTER(DSI), _VP]),
    "tdvh_cost_ssd": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, ct.POINTER(DSI), _VP]),
    "tdvh_cost_ncc": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, ct.POINTER(DSI), _VP]),
    "tdvh_sgm": (_I, [_VP, ct.POINTER(DSI), ct.POINTER(SgmParams), ct.POINTER(DSI), _VP, _I,
                      _VP]),
    "tdvh_wta": (_I, [_VP, ct.POINTER(DSI), _I, _VP, _I, _VP]),
    "tdvh_dp": (_I, [_VP, ct.POINTER(DSI), _U32, _I, _VP, _I, _VP]),
    "tdvh_dp_f32": (_I, [_VP, ct.POINTER(DSI), ct.c_float, _I, _VP, _I, _VP]),
    "tdvh_fast_wta_ssd": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_fast_wta_ncc": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_stereo_ssd_dp": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _U32, _I, _VP, _I, _VP]),
    "tdvh_fast_wta_ncc_noalloc": (_I, [_VP, _VP, _I, _I, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_stereo_bt_sgm": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, ct.POINTER(SgmParams), _I, _VP,
                                _I, _VP]),
    "tdvh_disp_to_unit_float": (_I, [_VP, _VP, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_unit_float_to_u8": (_I, [_VP, _VP, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_sgm_paths": (_SZ, [_I, _I, _VP]),
    "tdvh_prefilter_gauss5": (_I, [_VP, _VP, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_postfilter_open": (_I, [_VP, _VP, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_rgb_to_gray": (_I, [_VP, _VP, _I, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_remap_u8": (_I, [_VP, _VP, _I, _I, _I, _VP, _VP, _I, _I, _I, _I, _VP, _I, _VP]),
    "tdvh_reproject": (_I, [_VP, _VP, _I, _I, _I, _VP, ct.c_float, ct.c_float, _VP, _I, _VP]),
    "tdvh_host_alloc": (_I, [_SZ, ct.POINTER(_VP)]),
    "tdvh_host_free": (_I, [_VP]),
    "tdvh_copy_2d": (_I, [_VP, _VP, _SZ, _VP, _SZ, _SZ, _I, _VP]),
    "tdvh_dsi_download": (_I, [_VP, ct.POINTER(DSI), _VP, _VP]),
    "tdvh_pipeline_create": (_I, [_VP, _I, _I, _I, ct.POINTER(SgmParams), _I, ct.POINTER(_VP)]),
    "tdvh_pipeline_run": (_I, [_VP, _VP, _VP, _I, _VP, _I, _VP]),
    "tdvh_pipeline_destroy": (_I, [_VP]),
    "tdvh_batch_create": (_I, [_VP, ct.POINTER(Pair), _I, ct.POINTER(SgmParams),
                               ct.POINTER(_VP)]),
    "tdvh_batch_destroy": (_I, [_VP]),
    "tdvh_batch_config": (_I, [_VP, _I, _I, _I, _I, _I]),
    "tdvh_batch_run": (_I, [_VP, _I, _VP]),
    "tdvh_batch_wait": (_I, [_VP, _VP]),
    "tdvh_batch_cells": (_I64, [_VP]),
    "tdvh_batch_scratch_bytes": (_SZ, [_VP]),
}


def lib_path():
    return os.path.join(_PKG, "libtdvh.so")


def lib():
    """Loads libtdvh.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError("%s is missing: run `python -m thundervision_amd.build`" % path)
        h = ct.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def _check(rc):
    if rc != 0:
        raise TdvhError(rc, lib().tdvh_last_error().decode())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ct.c_void_p(s.cuda_stream)


def _ptr(t):
    return ct.c_void_p(t.data_ptr())


class Context:
    """Per-device tdvh context (owns scratch memory)."""

    def __init__(self, device=0):
        h = ct.c_void_p()
        _check(lib().tdvh_ctx_create(device, ct.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            _check(lib().tdvh_ctx_destroy(self.handle))
            self.handle = None

    def trim(self):
        _check(lib().tdvh_ctx_trim(self.handle))

    def scratch_bytes(self):
        return lib().tdvh_ctx_scratch_bytes(self.handle)

    def profile(self, on=True):
        """Per-kernel hipEvent timing of this context's launches."""
        _check(lib().tdvh_profile_enable(self.handle, int(on)))

    def profile_reset(self):
        _check(lib().tdvh_profile_reset(self.handle))

    def profile_table(self):
        """{kernel name: (total ms, launches)}"""
        out = {}
        i = 0
        name = ct.create_string_buffer(128)
        ms = ct.c_double()
        n = ct.c_int64()
        while lib().tdvh_profile_read(self.handle, i, name, 128, ct.byref(ms), ct.byref(n)) == 0:
            out[name.value.decode()] = (ms.value, n.value)
            i += 1
        return out

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT = {}


def default_context(device=0):
    if device not in _DEFAULT:
        _DEFAULT[device] = Context(device)
    return _DEFAULT[device]


def _ctx(ctx):
    return (ctx or default_context(torch.cuda.current_device())).handle


def split_words(depth):
    return 16 * ((depth + 31) // 32)


def dsi_for(t, depth, dtype, layout=LAYOUT_STRIDED, stride_x=None, stride_y=None):
    """DSI descriptor over tensor `t`: strided tensors are (H, W, D) with d
    contiguous unless explicit element strides are given; SPLIT16 tensors are
    (H, W, words) int32."""
    d = DSI()
    H, W = t.shape[0], t.shape[1]
    _check(lib().tdvh_dsi_init(ct.byref(d), _ptr(t), W, H, depth, dtype, layout))
    if layout == LAYOUT_STRIDED:
        d.stride_x = t.stride(1) if stride_x is None else stride_x
        d.stride_y = t.stride(0) if stride_y is None else stride_y
    return d


def _img_args(left, right):
    assert left.dtype == torch.uint8 and right.dtype == torch.uint8
    assert left.is_cuda and right.is_cuda and left.shape == right.shape and left.dim() == 2
    assert left.stride(1) == 1 and right.stride(1) == 1 and left.stride(0) == right.stride(0)
    H, W = left.shape
    return W, H, left.stride(0)


def new_split_volume(H, W, depth, device):
    return torch.empty((H, W, split_words(depth)), dtype=torch.int32, device=device)


def cost_bt(left, right, depth, layout=LAYOUT_STRIDED, ctx=None, stream=None):
    """Birchfield-Tomasi (2x units) -> (H, W, D) uint16 bits in an int16
    tensor, or a SPLIT16 (H, W, words) int32 tensor."""
    W, H, pitch = _img_args(left, right)
    if layout == LAYOUT_SPLIT16:
        out = new_split_volume(H, W, depth, left.device)
    else:
        out = torch.empty((H, W, depth), dtype=torch.int16, device=left.device)
    d = dsi_for(out, depth, U16, layout)
    _check(lib().tdvh_cost_bt(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, ct.byref(d),
                              _stream(stream)))
    return out


def cost_ssd(left, right, depth, radius=3, ctx=None, stream=None):
    W, H, pitch = _img_args(left, right)
    out = torch.empty((H, W, depth), dtype=torch.int32, device=left.device)
    d = dsi_for(out, depth, U32)
    _check(lib().tdvh_cost_ssd(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, radius,
                               ct.byref(d), _stream(stream)))
    return out


def cost_ncc(left, right, depth, radius=4, zero_mean=True, ctx=None, stream=None):
    W, H, pitch = _img_args(left, right)
    out = torch.empty((H, W, depth), dtype=torch.float32, device=left.device)
    d = dsi_for(out, depth, F32)
    _check(lib().tdvh_cost_ncc(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, radius,
                               int(zero_mean), ct.byref(d), _stream(stream)))
    return out


def sgm_params(p1=8, p2=32, dir_mask=0xFF, flags=0, fp1=30.0 / 255, fp2=150.0 / 255):
    return SgmParams(p1, p2, dir_mask, flags, fp1, fp2)


def sgm(C, depth=None, p1=8, p2=32, dir_mask=0xFF, flags=0, return_S=False, layout=None,
        fp1=30.0 / 255, fp2=150.0 / 255, ctx=None, stream=None):
    """SGM + WTA over a cost volume: (H, W, D) strided int16 (uint16 BT bits),
    int32 (uint32 SSD bits) or float32 (NCC), or a SPLIT16 uint16 volume when
    layout == LAYOUT_SPLIT16. Float volumes use the penalties fp1/fp2.
    Returns disp (and S, in the cost's dtype; int32 in reference geometry)."""
    H, W = C.shape[0], C.shape[1]
    if layout is None:
        layout = LAYOUT_STRIDED
    if depth is None:
        assert layout == LAYOUT_STRIDED
        depth = C.shape[2]
    dt = U16 if layout == LAYOUT_SPLIT16 else {torch.int16: U16, torch.int32: U32,
                                                torch.float32: F32}[C.dtype]
    cd = dsi_for(C, depth, dt, layout)
    params = sgm_params(p1, p2, dir_mask, flags, fp1, fp2)
    disp = torch.empty((H, W), dtype=torch.int16, device=C.device)
    S = None
    sp = None
    if return_S:
        if flags & SGM_REF_GEOMETRY:
            S = torch.empty((H, W, depth), dtype=torch.int32, device=C.device)
            sd = dsi_for(S, depth, I32)
        else:
            tdt = {U16: torch.int16, U32: torch.int32, F32: torch.float32}[dt]
            S = torch.empty((H, W, depth), dtype=tdt, device=C.device)
            sd = dsi_for(S, depth, dt)
        sp = ct.byref(sd)
    _check(lib().tdvh_sgm(_ctx(ctx), ct.byref(cd), ct.byref(params), sp, _ptr(disp), W,
                          _stream(stream)))
    return (disp, S) if return_S else disp


def wta(S, mode=WTA_LOWEST, ctx=None, stream=None):
    """WTA over an (H, W, D) volume: int16 tensors are read as uint16, int32
    as uint32 unless `signed` dtype I32 is wanted (use wta_dsi)."""
    dt = {torch.int16: U16, torch.int32: U32, torch.float32: F32}[S.dtype]
    return wta_dsi(S, dsi_for(S, S.shape[2], dt), mode, ctx, stream)


def wta_dsi(S, dsi, mode=WTA_LOWEST, ctx=None, stream=None):
    H, W = dsi.height, dsi.width
    disp = torch.empty((H, W), dtype=torch.int16, device=S.device)
    _check(lib().tdvh_wta(_ctx(ctx), ct.byref(dsi), mode, _ptr(disp), W, _stream(stream)))
    return disp


def dp(C, p_occ=20, flags=0, ctx=None, stream=None):
    """Scanline DP over an (H, W, D) volume: int16 (uint16 BT bits), int32
    (uint32 SSD bits) or float32 (NCC, float penalty p_occ)."""
    H, W, D = C.shape
    cd = dsi_for(C, D, {torch.int16: U16, torch.int32: U32, torch.float32: F32}[C.dtype])
    disp = torch.empty((H, W), dtype=torch.int16, device=C.device)
    if C.dtype == torch.float32:
        _check(lib().tdvh_dp_f32(_ctx(ctx), ct.byref(cd), float(p_occ), flags, _ptr(disp), W,
                                 _stream(stream)))
    else:
        _check(lib().tdvh_dp(_ctx(ctx), ct.byref(cd), p_occ, flags, _ptr(disp), W,
                             _stream(stream)))
    return disp


def fast_wta_ssd(left, right, depth, radius=3, ctx=None, stream=None):
    W, H, pitch = _img_args(left, right)
    disp = torch.empty((H, W), dtype=torch.int16, device=left.device)
    _check(lib().tdvh_fast_wta_ssd(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, depth,
                                   radius, _ptr(disp), W, _stream(stream)))
    return disp


def stereo_ssd_dp(left, right, depth, radius=3, p_occ=20, flags=0, out=None, ctx=None,
                  stream=None):
    """SSD window + scanline DP without a volume: dp(cost_ssd(...), p_occ, flags)."""
    W, H, pitch = _img_args(left, right)
    disp = out if out is not None else torch.empty((H, W), dtype=torch.int16,
                                                   device=left.device)
    _check(lib().tdvh_stereo_ssd_dp(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, depth,
                                    radius, p_occ, flags, _ptr(disp), W, _stream(stream)))
    return disp


def fast_wta_ncc(left, right, depth, radius=4, zero_mean=True, ctx=None, stream=None):
    """NCC + lowest-index WTA without a volume: wta(cost_ncc(...))."""
    W, H, pitch = _img_args(left, right)
    disp = torch.empty((H, W), dtype=torch.int16, device=left.device)
    _check(lib().tdvh_fast_wta_ncc(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, depth,
                                   radius, int(zero_mean), _ptr(disp), W, _stream(stream)))
    return disp


def stereo_bt_sgm(left, right, depth, p1=8, p2=32, dir_mask=0xFF, flags=0, fusion=FUSE_COST,
                  out=None, ctx=None, stream=None):
    """BT2 + SGM + WTA from images to int16 disparities."""
    W, H, pitch = _img_args(left, right)
    disp = out if out is not None else torch.empty((H, W), dtype=torch.int16,
                                                   device=left.device)
    params = sgm_params(p1, p2, dir_mask, flags)
    _check(lib().tdvh_stereo_bt_sgm(_ctx(ctx), _ptr(left), _ptr(right), W, H, pitch, depth,
                                    ct.byref(params), fusion, _ptr(disp), disp.stride(0),
                                    _stream(stream)))
    return disp


def disp_to_unit_float(disp, depth, ctx=None, stream=None):
    H, W = disp.shape
    out = torch.empty((H, W), dtype=torch.float32, device=disp.device)
    _check(lib().tdvh_disp_to_unit_float(_ctx(ctx), _ptr(disp), disp.stride(0), W, H, depth,
                                         _ptr(out), W, _stream(stream)))
    return out


def unit_float_to_u8(img, ctx=None, stream=None):
    H, W = img.shape
    out = torch.empty((H, W), dtype=torch.uint8, device=img.device)
    _check(lib().tdvh_unit_float_to_u8(_ctx(ctx), _ptr(img), img.stride(0), W, H, _ptr(out), W,
                                       _stream(stream)))
    return out


RGB, BGR = 0, 1


def prefilter_gauss5(img, out=None, ctx=None, stream=None):
    """5x5 binomial Gaussian of a uint8 (H, W) device image."""
    H, W = img.shape
    if out is None:
        out = torch.empty((H, W), dtype=torch.uint8, device=img.device)
    _check(lib().tdvh_prefilter_gauss5(_ctx(ctx), _ptr(img), img.stride(0), W, H, _ptr(out),
                                       out.stride(0), _stream(stream)))
    return out


def postfilter_open(disp, iterations=3, out=None, ctx=None, stream=None):
    """`iterations` 3x3 erosions then dilations of int16 (H, W) disparities."""
    H, W = disp.shape
    if out is None:
        out = torch.empty((H, W), dtype=torch.int16, device=disp.device)
    _check(lib().tdvh_postfilter_open(_ctx(ctx), _ptr(disp), disp.stride(0), W, H, iterations,
                                      _ptr(out), out.stride(0), _stream(stream)))
    return out


def rgb_to_gray(img, order=RGB, ctx=None, stream=None):
    """uint8 (H, W, 3 or 4) colour image on the device -> uint8 (H, W) gray."""
    H, W, C = img.shape
    assert img.stride(2) == 1 and img.stride(1) == C
    out = torch.empty((H, W), dtype=torch.uint8, device=img.device)
    _check(lib().tdvh_rgb_to_gray(_ctx(ctx), _ptr(img), img.stride(0), W, H, C, order, _ptr(out),
                                  W, _stream(stream)))
    return out


def remap(img, map_x, map_y, border=0, ctx=None, stream=None):
    """Bilinear remap of a uint8 (H, W) image by float (Ho, Wo) maps."""
    H, W = img.shape
    Ho, Wo = map_x.shape
    assert map_y.shape == map_x.shape and map_x.stride(0) == map_y.stride(0)
    out = torch.empty((Ho, Wo), dtype=torch.uint8, device=img.device)
    _check(lib().tdvh_remap_u8(_ctx(ctx), _ptr(img), img.stride(0), W, H, _ptr(map_x),
                               _ptr(map_y), map_x.stride(0), Wo, Ho, border, _ptr(out), Wo,
                               _stream(stream)))
    return out


def reproject(disp, Q, disp_scale=1.0, z_scale=24.0, ctx=None, stream=None):
    """int16 (H, W) disparities -> float32 (H, W, 3) points via the 4x4 Q."""
    H, W = disp.shape
    q = (ct.c_float * 16)(*[float(v) for v in list(Q.reshape(16))])
    out = torch.empty((H, W, 3), dtype=torch.float32, device=disp.device)
    _check(lib().tdvh_reproject(_ctx(ctx), _ptr(disp), disp.stride(0), W, H,
                                ct.cast(q, ct.c_void_p), disp_scale, z_scale, _ptr(out), 3 * W,
                                _stream(stream)))
    return out


class PinnedBuffer:
    """Page-locked host bytes (tdvh_host_alloc) viewed as a numpy array."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * self.dtype.itemsize
        h = ct.c_void_p()
        _check(lib().tdvh_host_alloc(n, ct.byref(h)))
        self.ptr = h
        raw = (ct.c_uint8 * n).from_address(h.value)
        self.array = np.frombuffer(raw, dtype=self.dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            _check(lib().tdvh_host_free(self.ptr))
            self.ptr = None


def copy_2d(dst_ptr, dst_pitch, src_ptr, src_pitch, width_bytes, rows, ctx=None, stream=None):
    """One asynchronous 2D copy between any two buffers (pitches in bytes)."""
    _check(lib().tdvh_copy_2d(_ctx(ctx), ct.c_void_p(dst_ptr), dst_pitch, ct.c_void_p(src_ptr),
                              src_pitch, width_bytes, rows, _stream(stream)))


def dsi_download(dsi, host_ptr, ctx=None, stream=None):
    """Copies a dense DSI (a DSI struct) to host memory in one transfer."""
    _check(lib().tdvh_dsi_download(_ctx(ctx), ct.byref(dsi), ct.c_void_p(host_ptr),
                                   _stream(stream)))


PIPE_PREFILTER, PIPE_POSTFILTER = 1, 2


class Pipeline:
    """Fixed-shape stereo pipeline replayed as a hipGraph (tdvh_pipeline_*)."""

    def __init__(self, width, height, depth, p1=8, p2=32, flags=PIPE_PREFILTER, ctx=None):
        self.width, self.height, self.depth = width, height, depth
        params = sgm_params(p1, p2)
        h = ct.c_void_p()
        _check(lib().tdvh_pipeline_create(_ctx(ctx), width, height, depth, ct.byref(params),
                                          flags, ct.byref(h)))
        self.handle = h

    def run(self, left, right, out=None, stream=None):
        if out is None:
            out = torch.empty((self.height, self.width), dtype=torch.int16, device=left.device)
        assert left.stride(0) == right.stride(0)
        _check(lib().tdvh_pipeline_run(self.handle, _ptr(left), _ptr(right), left.stride(0),
                                       _ptr(out), out.stride(0), _stream(stream)))
        return out

    def close(self):
        if self.handle:
            _check(lib().tdvh_pipeline_destroy(self.handle))
            self.handle = None


def sgm_paths(width, height):
    """Reference path table (host call); returns a list of SgPath."""
    n = lib().tdvh_sgm_paths(width, height, None)
    arr = (SgPath * n)()
    lib().tdvh_sgm_paths(width, height, ct.cast(arr, ct.c_void_p))
    return arr


class Batch:
    """Many pairs as one graph of persistent (or per-phase) launches
    (tdvh_batch_*)."""

    def __init__(self, pairs, p1=8, p2=32, ctx=None):
        self._keep = pairs  # (left, right, depth, disp) tuples keep tensors alive
        arr = (Pair * len(pairs))()
        for i, (left, right, depth, disp) in enumerate(pairs):
            W, H, pitch = _img_args(left, right)
            arr[i] = Pair(left.data_ptr(), right.data_ptr(), W, H, pitch, depth,
                          disp.data_ptr())
        params = sgm_params(p1, p2)
        h = ct.c_void_p()
        _check(lib().tdvh_batch_create(_ctx(ctx), arr, len(pairs), ct.byref(params),
                                       ct.byref(h)))
        self.handle = h

    def config(self, lines_per_wave=0, waves_per_group=0, prefetch=0, fusion_depth=0,
               sum_lag=0):
        """Knobs of tdvh_batch_config (0 keeps the current value)."""
        _check(lib().tdvh_batch_config(self.handle, lines_per_wave, waves_per_group, prefetch,
                                       fusion_depth, sum_lag))

    def run(self, use_graph=True, stream=None):
        _check(lib().tdvh_batch_run(self.handle, int(use_graph), _stream(stream)))

    def wait(self, stream=None):
        """Synchronises and raises if the work queue stalled."""
        _check(lib().tdvh_batch_wait(self.handle, _stream(stream)))

    def cells(self):
        return lib().tdvh_batch_cells(self.handle)

    def scratch_bytes(self):
        return lib().tdvh_batch_scratch_bytes(self.handle)

    def close(self):
        if self.handle:
            _check(lib().tdvh_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
