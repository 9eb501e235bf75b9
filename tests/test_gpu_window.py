"""GPU parity of the window costs (SSD, NCC), fused SSD+WTA and scanline DP
against the CPU oracle. SSD/DP/WTA bit-exact; NCC within 1e-5 relative of
the FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
from thundervision_amd import synth

pytestmark = pytest.mark.gpu

# NCC tolerance: |cost_gpu - cost_fp64| <= 1e-5 * max(|rho|, 1e-2)
NCC_REL, NCC_FLOOR = 1e-5, 1e-2
NEAR_TIE = 1e-6


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def pair(W, H, dmax, seed):
    L, R, _ = synth.pair(W, H, dmax, 8, seed)
    return L, R


@pytest.mark.parametrize("W,H,D,r", [(33, 21, 16, 3), (70, 45, 64, 3), (100, 37, 130, 3),
                                     (64, 64, 256, 3), (9, 40, 5, 1), (50, 33, 70, 5),
                                     (3, 2, 4, 2)])
def test_ssd_volume(tv, dev, W, H, D, r):
    L, R = pair(W, H, D - 1, W + H + r)
    want = oracle.cost_ssd(L, R, D, r)
    got = tv.cost_ssd(to_dev(L, dev), to_dev(R, dev), D, r)
    assert (got.cpu().numpy().view(np.uint32) == want).all()


@pytest.mark.parametrize("W,H,D,r", [(61, 29, 32, 3), (140, 50, 200, 3), (17, 80, 64, 2),
                                     (300, 70, 256, 3), (45, 33, 100, 1)])
def test_fast_wta_ssd(tv, dev, W, H, D, r):
    L, R = pair(W, H, D - 1, 7 * W + H)
    want = oracle.fast_wta_ssd(L, R, D, r)
    got = tv.fast_wta_ssd(to_dev(L, dev), to_dev(R, dev), D, r)
    assert (got.cpu().numpy() == want).all()


def _check_ncc(got, C64, rho):
    fin = np.isfinite(C64)
    assert (np.isfinite(got) == fin).all()
    err = np.abs(got[fin].astype(np.float64) - C64[fin])
    bound = NCC_REL * np.maximum(np.abs(rho[fin]), NCC_FLOOR)
    assert (err <= bound).all(), float((err / bound).max())


@pytest.mark.parametrize("W,H,D,r,zm", [(45, 30, 16, 4, True), (80, 41, 64, 4, True),
                                        (64, 40, 100, 1, False), (33, 33, 256, 4, True),
                                        (50, 20, 32, 2, False)])
def test_ncc_volume(tv, dev, W, H, D, r, zm):
    L, R = pair(W, H, D - 1, 3 * W + r)
    C64, rho = oracle.cost_ncc(L, R, D, r, zm)
    got = tv.cost_ncc(to_dev(L, dev), to_dev(R, dev), D, r, zm).cpu().numpy()
    _check_ncc(got, C64, rho)


def test_ncc_flat_windows_are_invalid(tv, dev):
    L = np.full((20, 30), 100, np.uint8)
    R = L.copy()
    got = tv.cost_ncc(to_dev(L, dev), to_dev(R, dev), 8, 4, True).cpu().numpy()
    assert np.isinf(got).all()


def test_ncc_window_1080p_crop_with_near_ties(tv, dev):
    # the 1920x1080 D=256 generator, cropped so the FP64 oracle stays small
    L, R, _ = synth.workload_pair(synth.WINDOW_1080P)
    L, R = np.ascontiguousarray(L[400:490, 1000:1320]), np.ascontiguousarray(R[400:490, 1000:1320])
    D = 256
    C64, rho = oracle.cost_ncc(L, R, D, 4, True)
    got = tv.cost_ncc(to_dev(L, dev), to_dev(R, dev), D, 4, True)
    _check_ncc(got.cpu().numpy(), C64, rho)
    # argmax rho (argmin cost): identical except at near-ties of the FP64 costs
    d_gpu = tv.wta(got).cpu().numpy()
    d_ref = oracle.wta(C64)
    srt = np.sort(np.where(np.isfinite(C64), C64, np.inf), axis=2)
    near = (srt[:, :, 1] - srt[:, :, 0]) < NEAR_TIE
    assert ((d_gpu == d_ref) | near).all()


@pytest.mark.parametrize("W,H,D,r,zm", [(45, 30, 16, 4, True), (80, 41, 64, 4, True),
                                        (64, 40, 100, 1, False), (33, 33, 256, 4, True),
                                        (50, 20, 32, 2, False), (130, 70, 200, 5, True),
                                        (300, 40, 256, 4, True), (5, 9, 3, 3, True)])
def test_fast_wta_ncc(tv, dev, W, H, D, r, zm):
    """Fused NCC + lowest-index WTA: bit-exact to the volume path (cost_ncc,
    then wta), including depths that pad the last chunk and the x < d border."""
    L, R = pair(W, H, D - 1, 5 * W + r)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    want = tv.wta(tv.cost_ncc(Ld, Rd, D, r, zm)).cpu().numpy()
    got = tv.fast_wta_ncc(Ld, Rd, D, r, zm).cpu().numpy()
    assert (got == want).all()


def test_fast_wta_ncc_ties_and_flat_regions(tv, dev):
    """Exact ties go to the lowest d, and all-invalid pixels give d = 0, as
    in the volume path's WTA. The ties come from a texture periodic along x:
    identical windows give identical costs at d = 3, 11, 19, ... Zero-variance
    windows in a flat block make every cost of a pixel invalid."""
    rng = np.random.default_rng(9)
    H, W, D = 40, 200, 64
    tile = rng.integers(0, 256, (H, 8), dtype=np.uint8)
    L = np.ascontiguousarray(np.tile(tile, (1, W // 8)))
    R = np.ascontiguousarray(np.roll(L, -3, axis=1))
    L[:, 150:] = 77
    R[:, 140:] = 77
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    for zm in (True, False):
        want = tv.wta(tv.cost_ncc(Ld, Rd, D, 4, zm)).cpu().numpy()
        got = tv.fast_wta_ncc(Ld, Rd, D, 4, zm).cpu().numpy()
        assert (got == want).all()
        assert (want[:, 20:140] % 8 == 3).mean() > 0.9  # the ties were there to break


def test_fast_wta_ncc_1080p(tv, dev):
    """1920x1080, D = 256, 9x9 zero-mean: fused equals the volume path."""
    L, R, _ = synth.workload_pair(synth.WINDOW_1080P)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    C = tv.cost_ncc(Ld, Rd, 256, 4, True)
    want = tv.wta(C).cpu().numpy()
    del C
    got = tv.fast_wta_ncc(Ld, Rd, 256, 4, True).cpu().numpy()
    assert (got == want).all()


@pytest.mark.parametrize("bounded", [0, 4])
@pytest.mark.parametrize("W,H,D,flags", [(40, 12, 16, 0), (97, 20, 64, 0), (64, 9, 200, 0),
                                         (70, 15, 33, 1), (70, 15, 33, 2), (70, 15, 33, 3),
                                         (1, 5, 8, 0), (130, 7, 256, 0), (300, 6, 100, 0),
                                         (700, 4, 256, 3)])
def test_dp(tv, dev, W, H, D, flags, bounded):
    """SSD volumes (costs <= 49 * 255^2, invalid exactly where d > x), in
    64-bit sums and in the bounded-cost 32-bit form."""
    L, R = pair(W, H, D - 1, 11 * W + D)
    C = oracle.cost_ssd(L, R, D, 3)
    want = oracle.dp(C, 20, flags)
    got = tv.dp(to_dev(C.view(np.int32), dev), 20, flags | bounded)
    assert (got.cpu().numpy() == want).all()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dp_bounded_extreme_costs(tv, dev, seed):
    """The bounded-cost contract at its edges: costs up to 2^22 - 1 (far above
    SSD's), long rows, p_occ up to 2^20 - 1, ties from a coarse cost alphabet,
    against the oracle's 64-bit sums; both volume layouts."""
    rng = np.random.default_rng(seed)
    H, W, D = 5, 1500, 256
    hi = [1 << 22, 4, 1 << 22][seed - 1]
    p_occ = [20, 1, (1 << 20) - 1][seed - 1]
    C = rng.integers(0, hi, (H, W, D), dtype=np.uint64).astype(np.uint32)
    x = np.arange(W)[None, :, None]
    d = np.arange(D)[None, None, :]
    C[np.broadcast_to(d > x, C.shape)] = 0xFFFFFFFF
    want = oracle.dp(C, p_occ)
    got = tv.dp(to_dev(C.view(np.int32), dev), p_occ, tv.DP_COST_BOUNDED).cpu().numpy()
    assert (got == want).all()
    Cx = to_dev(np.ascontiguousarray(C.transpose(1, 0, 2)).view(np.int32), dev).permute(1, 0, 2)
    assert (tv.dp(Cx, p_occ, tv.DP_COST_BOUNDED).cpu().numpy() == want).all()


@pytest.mark.parametrize("D", [64, 256, 512])
def test_dp_bounded_renormalisation(tv, dev, D):
    """Every valid cost just below 2^22: path sums pass 2^32 after about
    1000 columns, so the 32-bit form is exact only if its periodic wave-
    minimum subtraction works (a wrong or stale minimum saturates or wraps
    the sums and changes the decisions)."""
    rng = np.random.default_rng(D)
    H, W = 4, 1700
    C = ((1 << 22) - 1 - rng.integers(0, 3000, (H, W, D))).astype(np.uint32)
    x = np.arange(W)[None, :, None]
    d = np.arange(D)[None, None, :]
    C[np.broadcast_to(d > x, C.shape)] = 0xFFFFFFFF
    want = oracle.dp(C, 700)
    got = tv.dp(to_dev(C.view(np.int32), dev), 700, tv.DP_COST_BOUNDED).cpu().numpy()
    assert (got == want).all()


def test_dp_bounded_flag_with_large_penalty_runs_64_bit(tv, dev):
    # the 32-bit form needs p_occ < 2^20; above it the flag is ignored
    rng = np.random.default_rng(12)
    C = rng.integers(0, 1 << 21, (3, 40, 9)).astype(np.uint32)
    for p in ((1 << 20), (1 << 31) + 5):
        got = tv.dp(to_dev(C.view(np.int32), dev), p, tv.DP_COST_BOUNDED).cpu().numpy()
        assert (got == oracle.dp(C, p)).all()


@pytest.mark.parametrize("hi,inv", [(1 << 20, 0.0), (1 << 31, 0.05), (0xFFFFFFF0, 0.3)])
def test_dp_arbitrary_costs(tv, dev, hi, inv):
    """Costs far above SSD's range and scattered invalid cells (64-bit sums),
    both layouts of the volume's strides."""
    rng = np.random.default_rng(hi % 1000 + int(inv * 100))
    H, W, D = 7, 300, 256
    C = rng.integers(0, hi, (H, W, D), dtype=np.uint64).astype(np.uint32)
    C[rng.random(C.shape) < inv] = 0xFFFFFFFF
    want = oracle.dp(C, 20)
    assert (tv.dp(to_dev(C.view(np.int32), dev), 20).cpu().numpy() == want).all()
    # x-major storage ([x][y][d], the reference's DSI order) seen as (H, W, D)
    Cx = to_dev(np.ascontiguousarray(C.transpose(1, 0, 2)).view(np.int32), dev).permute(1, 0, 2)
    assert (tv.dp(Cx, 20).cpu().numpy() == want).all()


def test_ssd_dp_window_1080p(tv, dev):
    """1920x1080, D = 256: 7x7 SSD (uint32) -> DP (P_occ = 20), bit-exact."""
    L, R, gt = synth.workload_pair(synth.WINDOW_1080P)
    C = oracle.cost_ssd(L, R, 256, 3)
    Ld, Rd = to_dev(L, dev), to_dev(R, dev)
    Cg = tv.cost_ssd(Ld, Rd, 256, 3)
    assert (Cg.cpu().numpy().view(np.uint32) == C).all()
    want = oracle.dp(C, 20)
    got = tv.dp(Cg, 20).cpu().numpy()
    assert (got == want).all()
    assert (tv.dp(Cg, 20, tv.DP_COST_BOUNDED).cpu().numpy() == want).all()
    assert (np.abs(got.astype(int) - gt) <= 1).mean() > 0.5
    # fused SSD + WTA (no volume) at the full size: the lowest-index argmin of
    # the same volume
    fused = tv.fast_wta_ssd(Ld, Rd, 256, 3).cpu().numpy()
    assert (fused == oracle.wta(C)).all()
