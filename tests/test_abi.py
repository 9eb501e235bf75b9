"""The C-ABI library loads and exports every function declared in
include/tdvh.h (no compute calls: runs without a GPU)."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(tdvh_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("tdvh_cost_bt", "tdvh_sgm", "tdvh_wta", "tdvh_dp", "tdvh_cost_ssd",
                 "tdvh_cost_ncc", "tdvh_fast_wta_ssd", "tdvh_fast_wta_ncc", "tdvh_stereo_bt_sgm",
                 "tdvh_batch_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from thundervision_amd import tdvh
    lib = ct.CDLL(tdvh.lib_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from thundervision_amd import tdvh
    assert set(declared_functions()) == set(tdvh.SIGNATURES)


def test_host_only_calls():
    from thundervision_amd import tdvh
    lib = tdvh.lib()
    assert lib.tdvh_dsi_bytes(1280, 720, 128, tdvh.U16, tdvh.LAYOUT_SPLIT16) == 1280 * 720 * 256
    assert lib.tdvh_dsi_bytes(10, 10, 80, tdvh.U16, tdvh.LAYOUT_SPLIT16) == 100 * 96 * 2
    assert lib.tdvh_sgm_paths(384, 288, None) == 2014
    d = tdvh.DSI()
    assert lib.tdvh_dsi_init(ct.byref(d), None, 0, 5, 5, tdvh.U16, 0) == -1
    assert b"bad shape" in lib.tdvh_last_error()


def test_path_table_matches_oracle():
    import numpy as np
    import oracle
    from thundervision_amd import tdvh
    for W, H in ((64, 48), (48, 64), (31, 31), (384, 288)):
        a = tdvh.sgm_paths(W, H)
        got = np.array([[p.sx, p.sy, p.ex, p.ey, p.dx, p.dy, p.size] for p in a])
        assert (got == oracle.sgm_paths(W, H)[:, :7]).all()
