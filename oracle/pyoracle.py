"""numpy wrappers over liboracle.so (tdvref.c). Test infrastructure only."""
import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

INV16 = 0xFFFF
INV32 = 0xFFFFFFFF
INF_I32 = 0x3FFFFFFF
SGM_START_ASSIGN = 1
SGM_MIN_SKIP_LAST = 2
DP_START_ZERO = 1
DP_REF_BACKTRACK = 2
DIR_DX = (1, -1, 0, 0, 1, -1, -1, 1)
DIR_DY = (0, 0, 1, -1, 1, -1, 1, -1)


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ct.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _img(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    assert a.ndim == 2
    return a


def cost_bt2(L, R, D):
    L, R = _img(L), _img(R)
    H, W = L.shape
    C = np.empty((H, W, D), np.uint16)
    lib().tdvref_cost_bt2(_p(L), _p(R), W, H, D, _p(C))
    return C


def cost_ssd(L, R, D, radius=3, box=True):
    L, R = _img(L), _img(R)
    H, W = L.shape
    C = np.empty((H, W, D), np.uint32)
    fn = lib().tdvref_cost_ssd_box if box else lib().tdvref_cost_ssd
    fn(_p(L), _p(R), W, H, D, radius, _p(C))
    return C


def cost_ncc(L, R, D, radius=4, zero_mean=True):
    """Returns (cost = 1 - rho as float64, rho) with +inf / NaN where invalid."""
    L, R = _img(L), _img(R)
    H, W = L.shape
    C = np.empty((H, W, D), np.float64)
    rho = np.empty((H, W, D), np.float64)
    lib().tdvref_cost_ncc(_p(L), _p(R), W, H, D, radius, int(zero_mean), _p(C), _p(rho))
    return C, rho


def sgm_lr(C, P1, P2, direction):
    C = np.ascontiguousarray(C, np.uint16)
    H, W, D = C.shape
    Lr = np.empty_like(C)
    lib().tdvref_sgm_lr(_p(C), W, H, D, P1, P2, direction, _p(Lr))
    return Lr


def sgm_std(C, P1, P2, dir_mask=0xFF):
    C = np.ascontiguousarray(C, np.uint16)
    H, W, D = C.shape
    S = np.empty_like(C)
    lib().tdvref_sgm_std(_p(C), W, H, D, P1, P2, dir_mask, _p(S))
    return S


def sgm_u32(C, P1, P2, dir_mask=0xFF):
    """SGM over uint32 costs (saturating at INV32), directions summed in order."""
    C = np.ascontiguousarray(C, np.uint32)
    H, W, D = C.shape
    S = np.empty_like(C)
    lib().tdvref_sgm_u32(_p(C), W, H, D, ct.c_uint32(P1), ct.c_uint32(P2), dir_mask, _p(S))
    return S


def sgm_f32(C, P1, P2, dir_mask=0xFF):
    """SGM over float32 costs (IEEE single precision, fixed operation order)."""
    C = np.ascontiguousarray(C, np.float32)
    H, W, D = C.shape
    S = np.empty_like(C)
    lib().tdvref_sgm_f32(_p(C), W, H, D, ct.c_float(P1), ct.c_float(P2), dir_mask, _p(S))
    return S


def sgm_ref(C, P1, P2, flags=SGM_START_ASSIGN | SGM_MIN_SKIP_LAST):
    C = np.ascontiguousarray(C, np.uint16)
    H, W, D = C.shape
    S = np.empty((H, W, D), np.int32)
    lib().tdvref_sgm_ref(_p(C), W, H, D, P1, P2, flags, _p(S))
    return S


def sgm_paths(W, H):
    n = lib().tdvref_sgm_path_count(W, H)
    lib().tdvref_sgm_path_count.restype = ct.c_size_t
    n = lib().tdvref_sgm_path_count(W, H)
    a = np.zeros((n, 8), np.uint16)
    lib().tdvref_sgm_paths(W, H, _p(a))
    out = a.astype(np.int32)
    out[:, 4:6] = a[:, 4:6].view(np.int16)
    return out  # columns sx sy ex ey dx dy size pad


def sgm_ref_visits(W, H):
    v = np.empty((8, H, W), np.int32)
    lib().tdvref_sgm_ref_visits(W, H, _p(v))
    return v


def wta(S):
    S = np.ascontiguousarray(S)
    H, W, D = S.shape
    disp = np.empty((H, W), np.int16)
    fn = {np.dtype(np.uint16): lib().tdvref_wta_u16, np.dtype(np.uint32): lib().tdvref_wta_u32,
          np.dtype(np.int32): lib().tdvref_wta_i32, np.dtype(np.float64): lib().tdvref_wta_f64,
          np.dtype(np.float32): lib().tdvref_wta_f32}[S.dtype]
    fn(_p(S), W, H, D, _p(disp))
    return disp


def wta_tree(S):
    S = np.ascontiguousarray(S, np.uint16)
    H, W, D = S.shape
    disp = np.empty((H, W), np.int16)
    lib().tdvref_wta_tree_u16(_p(S), W, H, D, _p(disp))
    return disp


def dp(C, p_occ=20, flags=0, return_acc=False):
    C = np.ascontiguousarray(C, np.uint32)
    H, W, D = C.shape
    disp = np.empty((H, W), np.int16)
    acc = np.empty((H, D), np.uint64) if return_acc else None
    lib().tdvref_dp(_p(C), W, H, D, ct.c_uint32(p_occ), flags, _p(disp),
                    _p(acc) if acc is not None else None)
    return (disp, acc) if return_acc else disp


def dp_u16(C, p_occ=20, flags=0):
    """DP over uint16 (BT) costs: INV16 read as INV32, then tdvref_dp."""
    C = np.asarray(C, np.uint16)
    C32 = np.where(C == INV16, np.uint32(INV32), C.astype(np.uint32))
    return dp(C32, p_occ, flags)


def dp_f32(C, p_occ=0.0, flags=0):
    C = np.ascontiguousarray(C, np.float32)
    H, W, D = C.shape
    disp = np.empty((H, W), np.int16)
    lib().tdvref_dp_f32(_p(C), W, H, D, ct.c_float(p_occ), flags, _p(disp))
    return disp


def fast_wta_ssd(L, R, D, radius=3):
    L, R = _img(L), _img(R)
    H, W = L.shape
    disp = np.empty((H, W), np.int16)
    lib().tdvref_fast_wta_ssd(_p(L), _p(R), W, H, D, radius, _p(disp))
    return disp


def stereo_bt_sgm(L, R, D, P1=8, P2=32):
    """BT2 -> 8-direction standard SGM -> lowest-index WTA."""
    return wta(sgm_std(cost_bt2(L, R, D), P1, P2))


# ---- around the path (tdvref_io.c) ------------------------------------------

def gauss5(img):
    img = _img(img)
    H, W = img.shape
    out = np.empty_like(img)
    lib().tdvref_gauss5_u8(_p(img), W, H, _p(out))
    return out


def open_disp(disp, iterations=3):
    d = np.ascontiguousarray(disp, dtype=np.int16)
    H, W = d.shape
    out = np.empty_like(d)
    lib().tdvref_open_i16(_p(d), W, H, iterations, _p(out))
    return out


def rgb_to_gray(img, order=0):
    a = np.ascontiguousarray(img, dtype=np.uint8)
    H, W, ch = a.shape
    out = np.empty((H, W), np.uint8)
    lib().tdvref_rgb_to_gray_u8(_p(a), W, H, ch, order, _p(out))
    return out


def remap(img, mapx, mapy, border=0):
    img = _img(img)
    H, W = img.shape
    mx = np.ascontiguousarray(mapx, dtype=np.float32)
    my = np.ascontiguousarray(mapy, dtype=np.float32)
    Ho, Wo = mx.shape
    out = np.empty((Ho, Wo), np.uint8)
    lib().tdvref_remap_u8(_p(img), W, H, _p(mx), _p(my), Wo, Ho, border, _p(out))
    return out


def reproject(disp, Q, disp_scale=1.0, z_scale=24.0):
    d = np.ascontiguousarray(disp, dtype=np.int16)
    H, W = d.shape
    q = np.ascontiguousarray(Q, dtype=np.float32).reshape(16)
    out = np.empty((H, W, 3), np.float32)
    lib().tdvref_reproject(_p(d), W, H, _p(q), ct.c_float(disp_scale), ct.c_float(z_scale),
                           _p(out))
    return out
