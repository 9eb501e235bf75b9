"""Runs the C++ stage-layer tests (tests/cpp/test_stages.cpp, built by
__graft_entry__.build into thundervision_amd/_bin/test_stages)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from thundervision_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "thundervision_amd", "_bin", "test_stages")


def run(*args):
    assert os.path.exists(BIN), "build first: python -m thundervision_amd.build"
    r = subprocess.run([BIN] + list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_runtime_pipes_and_runner():
    assert "cpu: ok" in run("cpu")


@pytest.mark.gpu
def test_stage_chains_on_gpu(tmp_path):
    L, R, _ = synth.workload_pair(synth.PLAIN_384)
    D = 16
    H, W = L.shape
    (tmp_path / "shape.txt").write_text("%d %d %d\n" % (W, H, D))
    L.tofile(tmp_path / "left.u8")
    R.tofile(tmp_path / "right.u8")
    C = oracle.cost_bt2(L, R, D)
    oracle.wta(oracle.sgm_std(C, 8, 32)).tofile(tmp_path / "bt_sgm.i16")
    np.ascontiguousarray(C.transpose(1, 0, 2)).tofile(tmp_path / "bt_xyd.u16")
    oracle.wta(oracle.cost_ssd(L, R, D, 3)).tofile(tmp_path / "ssd_wta.i16")
    # DevStereoMatcher chains see the Gaussian-prefiltered pair
    # (devstereomatcher.cpp:22-29); the fused matchers see the raw pair
    Lf, Rf = oracle.gauss5(L), oracle.gauss5(R)
    Cf = oracle.cost_bt2(Lf, Rf, D)
    sgm_f = oracle.wta(oracle.sgm_std(Cf, 8, 32))
    sgm_f.tofile(tmp_path / "bt_sgm_pf.i16")
    oracle.open_disp(sgm_f, 3).tofile(tmp_path / "bt_sgm_pf_open.i16")
    oracle.wta(Cf).tofile(tmp_path / "bt_wta_pf.i16")
    Sf = oracle.cost_ssd(Lf, Rf, D, 3)
    oracle.dp(Sf, 20).tofile(tmp_path / "ssd_dp_pf.i16")
    oracle.wta(Sf).tofile(tmp_path / "ssd_wta_pf.i16")
    assert "gpu: ok" in run("gpu", str(tmp_path))
