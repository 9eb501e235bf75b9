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
    # all nine factory pairings on the prefiltered pair; the NCC volume
    # (CrossCorrelationDev's 3x3 plain NCC) comes from the device, whose
    # float costs are pinned against the FP64 oracle elsewhere, and the
    # optimisers' float arithmetic is checked bit-exact against the oracle's
    import torch
    import thundervision_amd.tdvh as tv
    dev = torch.device("cuda:0")
    Nf = tv.cost_ncc(torch.from_numpy(Lf).to(dev), torch.from_numpy(Rf).to(dev), D, 1,
                     False).cpu().numpy()
    fp1, fp2 = np.float32(30.0 / 255), np.float32(150.0 / 255)
    vols = {"ssd": (Sf, lambda: oracle.dp(Sf, 20), lambda: oracle.sgm_u32(Sf, 8, 32)),
            "ncc": (Nf, lambda: oracle.dp_f32(Nf, 0.0), lambda: oracle.sgm_f32(Nf, fp1, fp2)),
            "bt": (Cf, lambda: oracle.dp_u16(Cf, 20), lambda: oracle.sgm_std(Cf, 8, 32))}
    for name, (vol, dpf, sgmf) in vols.items():
        oracle.wta(vol).tofile(tmp_path / ("pair_%s_wta.i16" % name))
        dpf().tofile(tmp_path / ("pair_%s_dp.i16" % name))
        oracle.wta(sgmf()).tofile(tmp_path / ("pair_%s_sgm.i16" % name))
    # TestSemiGlobalDev.WithSSD / TestSemiGlobalCPU.WithSSD: SSD -> SGM at 64
    # and 256 disparities
    for Dd in (64, 256):
        oracle.wta(oracle.sgm_u32(oracle.cost_ssd(Lf, Rf, Dd, 3), 8, 32)).tofile(
            tmp_path / ("ssd_sgm_pf_%d.i16" % Dd))
    assert "gpu: ok" in run("gpu", str(tmp_path))
