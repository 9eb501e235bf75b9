"""Pins the CPU oracle (oracle/tdvref.c) against closed forms, symmetries,
brute force and the reference's own path-table arithmetic. CPU only."""
import itertools

import numpy as np
import pytest

import oracle
from thundervision_amd import synth

rng = np.random.default_rng(1234)


def rand_img(H, W):
    return rng.integers(0, 256, (H, W), dtype=np.uint8)


# ---- Birchfield-Tomasi -------------------------------------------------------

def test_bt_constant_images_are_zero_or_invalid():
    L = np.full((5, 9), 77, np.uint8)
    C = oracle.cost_bt2(L, L, 6)
    for x in range(9):
        for d in range(6):
            want = 0 if x - d >= 0 else oracle.INV16
            assert (C[:, x, d] == want).all()


def test_bt_hand_computed_step_edge():
    # left row 10 10 50 50 50 ; right row = left shifted by one (R(x) = L(x+1))
    L = np.array([[10, 10, 50, 50, 50]], np.uint8)
    R = np.array([[10, 50, 50, 50, 50]], np.uint8)
    C = oracle.cost_bt2(L, R, 2)
    # x=1, d=0: left pixel 10 (neighbours 10, 50 -> 2x terms 20, 20, 60) range
    # [20, 60]; right pixel R(1)=50 (neighbours 10, 50 -> 100, 60, 100) range
    # [60, 100]. l2=20 vs [60,100] -> 40; r2=100 vs [20,60] -> 40 -> 40.
    assert C[0, 1, 0] == 40
    # x=2, d=1: left 50 (neighbours 10, 50 -> 100, 60, 100), right R(1)=50
    # (100, 60, 100): identical -> 0
    assert C[0, 2, 1] == 0
    # x=1, d=1: left 10 -> l2=20 range [20,60]; right R(0)=10 (clamped
    # neighbour 10 and 50 -> 20, 20, 60) -> 0
    assert C[0, 1, 1] == 0
    assert C[0, 0, 1] == oracle.INV16


def test_bt_mirror_symmetry():
    # C_{L,R}(x, d) == C_{flip R, flip L}(W-1-x+d, d): BT is symmetric in its
    # two images and the clamp is mirror-symmetric.
    H, W, D = 6, 23, 7
    L, R = rand_img(H, W), rand_img(H, W)
    C = oracle.cost_bt2(L, R, D)
    Cm = oracle.cost_bt2(R[:, ::-1], L[:, ::-1], D)
    for x in range(W):
        for d in range(D):
            if x - d < 0:
                continue
            assert (C[:, x, d] == Cm[:, W - 1 - x + d, d]).all()


def test_bt_range_and_invalid_mask():
    H, W, D = 7, 31, 12
    C = oracle.cost_bt2(rand_img(H, W), rand_img(H, W), D)
    x = np.arange(W)[None, :, None]
    d = np.arange(D)[None, None, :]
    inv = np.broadcast_to(x - d < 0, C.shape)
    assert (C[inv] == oracle.INV16).all()
    assert (C[~inv] <= 510).all()


# ---- SSD / NCC -----------------------------------------------------------------

def test_ssd_box_equals_brute_force():
    H, W, D = 9, 17, 6
    L, R = rand_img(H, W), rand_img(H, W)
    for r in (1, 3):
        assert (oracle.cost_ssd(L, R, D, r, box=True) == oracle.cost_ssd(L, R, D, r, box=False)).all()


def test_ssd_constant_offset_closed_form():
    L = rand_img(12, 20) // 2
    R = L + 7
    C = oracle.cost_ssd(L, R, 1, 3)
    assert (C[:, :, 0] == 49 * 49).all()


def _ncc_brute(L, R, D, r, zero_mean):
    H, W = L.shape
    out = np.full((H, W, D), np.inf)
    for y in range(H):
        for x in range(W):
            ys = np.clip(np.arange(y - r, y + r + 1), 0, H - 1)
            xr = np.arange(x - r, x + r + 1)
            a = L[np.ix_(ys, np.clip(xr, 0, W - 1))].astype(np.float64)
            for d in range(D):
                if x - d < 0:
                    continue
                b = R[np.ix_(ys, np.clip(xr - d, 0, W - 1))].astype(np.float64)
                if zero_mean:
                    a0, b0 = a - a.mean(), b - b.mean()
                else:
                    a0, b0 = a, b
                den = np.sqrt((a0 * a0).sum() * (b0 * b0).sum())
                if den > 0:
                    out[y, x, d] = 1.0 - (a0 * b0).sum() / den
    return out


@pytest.mark.parametrize("zero_mean,r", [(True, 4), (False, 1), (True, 1)])
def test_ncc_matches_brute_force(zero_mean, r):
    H, W, D = 11, 14, 5
    L, R = rand_img(H, W), rand_img(H, W)
    C, _ = oracle.cost_ncc(L, R, D, r, zero_mean)
    B = _ncc_brute(L, R, D, r, zero_mean)
    fin = np.isfinite(B)
    assert (np.isfinite(C) == fin).all()
    assert np.allclose(C[fin], B[fin], rtol=0, atol=1e-12)


def test_ncc_affine_gives_rho_one_and_flat_is_invalid():
    L = rand_img(10, 12) // 3 + 1
    R = (2 * L.astype(np.int32) + 5).astype(np.uint8)
    C, rho = oracle.cost_ncc(L, R, 1, 2, True)
    assert np.allclose(rho[:, :, 0], 1.0, atol=1e-12)
    flat = np.full((6, 6), 9, np.uint8)
    C, rho = oracle.cost_ncc(flat, flat, 2, 1, True)
    assert np.isinf(C).all()


# ---- SGM -----------------------------------------------------------------------

def test_sgm_zero_penalties_give_eight_times_cost():
    C = oracle.cost_bt2(rand_img(9, 13), rand_img(9, 13), 5)
    S = oracle.sgm_std(C, 0, 0)
    v = C != oracle.INV16
    assert (S[v] == 8 * C[v]).all() and (S[~v] == oracle.INV16).all()


def test_sgm_zero_volume():
    C = np.zeros((6, 7, 4), np.uint16)
    assert (oracle.sgm_std(C, 8, 32) == 0).all()


def _lr_bruteforce(c, P1, P2):
    """Path costs along one scan line by exhaustive enumeration of disparity
    sequences: U(k, d) = min over sequences ending in d of sum C + penalties
    (0 same, P1 for +-1, P2 otherwise); L(k, d) = U(k, d) - min U(k-1, .)."""
    n, D = c.shape
    INF = float("inf")
    U = np.full((n, D), INF)
    for k in range(n):
        for seq in itertools.product(range(D), repeat=k + 1):
            tot = 0.0
            for i, d in enumerate(seq):
                v = c[i, d]
                if v == oracle.INV16:
                    tot = INF
                    break
                tot += v
                if i:
                    jump = abs(d - seq[i - 1])
                    tot += 0 if jump == 0 else (P1 if jump == 1 else P2)
            U[k, seq[-1]] = min(U[k, seq[-1]], tot)
    L = np.empty_like(U)
    L[0] = U[0]
    for k in range(1, n):
        L[k] = U[k] - U[k - 1][np.isfinite(U[k - 1])].min()
    return L


def test_sgm_single_direction_matches_path_enumeration():
    H, W, D = 3, 5, 4
    C = oracle.cost_bt2(rand_img(H, W), rand_img(H, W), D)
    C[C != oracle.INV16] = rng.integers(0, 60, (C != oracle.INV16).sum())
    Lr = oracle.sgm_lr(C, 5, 17, 0)  # direction (1,0)
    for y in range(H):
        want = _lr_bruteforce(C[y], 5, 17)
        got = Lr[y].astype(np.float64)
        got[Lr[y] == oracle.INV16] = np.inf
        assert np.array_equal(got, want)


def test_sgm_direction_symmetries():
    H, W, D = 7, 9, 6
    C = oracle.cost_bt2(rand_img(H, W), rand_img(H, W), D)
    Cs = rng.integers(0, 300, C.shape).astype(np.uint16)  # no invalid cells
    flip = lambda v: v[:, ::-1]
    tr = lambda v: v.transpose(1, 0, 2)
    # mirror in x maps (1,0)->(-1,0), (1,1)->(-1,1), (1,-1)->(-1,-1)
    for a, b in ((0, 1), (4, 6), (7, 5), (2, 2)):
        assert (oracle.sgm_lr(np.ascontiguousarray(flip(Cs)), 8, 32, a) == flip(oracle.sgm_lr(Cs, 8, 32, b))).all()
    # transpose maps (1,0)->(0,1), (-1,0)->(0,-1), (1,1)->(1,1), (-1,1)->(1,-1)
    for a, b in ((2, 0), (3, 1), (4, 4), (7, 6)):
        assert (oracle.sgm_lr(np.ascontiguousarray(tr(Cs)), 8, 32, a) == tr(oracle.sgm_lr(Cs, 8, 32, b))).all()


def test_sgm_bounds():
    C = oracle.cost_bt2(rand_img(20, 30), rand_img(20, 30), 9)
    for d in range(8):
        Lr = oracle.sgm_lr(C, 8, 32, d)
        v = C != oracle.INV16
        assert (Lr[v] >= C[v]).all() and (Lr[v] <= C[v].astype(int) + 32).all()
        assert (Lr[~v] == oracle.INV16).all()


def test_path_table_counts():
    for (W, H), n in (((384, 288), 2014), ((1280, 720), 5998), ((1920, 1080), 8998)):
        assert len(oracle.sgm_paths(W, H)) == n


@pytest.mark.parametrize("W,H,dirs,skipped,revisited,maxv", [
    (384, 288, (5, 7), 27648, 288, 97),
    (64, 48, (5, 7), 768, 48, 17),
    (48, 64, (5,), 720, 48, 16),
])
def test_reference_path_coverage(W, H, dirs, skipped, revisited, maxv):
    v = oracle.sgm_ref_visits(W, H)
    for di in range(8):
        if di in dirs:
            assert (v[di] == 0).sum() == skipped
            assert (v[di] > 1).sum() == revisited
            assert v[di].max() == maxv
        else:
            assert (v[di] == 1).all(), di


def test_reference_path_coverage_square_is_exact():
    assert (oracle.sgm_ref_visits(32, 32) == 1).all()


def test_ref_geometry_on_square_equals_standard_without_quirks():
    C = oracle.cost_bt2(rand_img(12, 12), rand_img(12, 12), 6)
    S = oracle.sgm_ref(C, 8, 32, flags=0)
    Sstd = oracle.sgm_std(C, 8, 32)
    v = C != oracle.INV16
    assert (S[v] == Sstd[v].astype(np.int32)).all() and (S[~v] == oracle.INF_I32).all()


def test_ref_min_skip_last_is_inert_when_last_disparity_invalid():
    C = oracle.cost_bt2(rand_img(10, 14), rand_img(10, 14), 5)
    C[:, :, 4] = oracle.INV16
    a = oracle.sgm_ref(C, 8, 32, flags=oracle.SGM_START_ASSIGN)
    b = oracle.sgm_ref(C, 8, 32, flags=oracle.SGM_START_ASSIGN | oracle.SGM_MIN_SKIP_LAST)
    assert (a == b).all()


# ---- WTA -----------------------------------------------------------------------

def test_wta_ties_go_to_lowest_index_and_invalid_never_wins():
    S = np.array([[[5, 3, 3, 9], [7, 7, 7, 7], [0xFFFF, 4, 2, 2], [0xFFFF] * 4]], np.uint16)
    assert oracle.wta(S).tolist() == [[1, 0, 2, 0]]


def test_wta_tree_ties_go_right():
    # pairs (0,2),(1,3) then (w02, w13); equal values take the right operand
    S = np.array([[[4, 4, 4, 4], [1, 5, 1, 5], [3, 2, 9, 2]]], np.uint16)
    # all equal: (0,2)->2, (1,3)->3, (2,3)->3
    # [1,5,1,5]: (0,2)->2 (tie), (1,3)->3, (1 vs 5) -> left=2
    # [3,2,9,2]: (0,2)->0, (1,3)->3 (tie), (3 vs 2) -> right=3
    assert oracle.wta_tree(S).tolist() == [[3, 2, 3]]


# ---- DP ------------------------------------------------------------------------

def test_dp_optimal_cost_matches_enumeration():
    W, D, P = 6, 4, 20
    for trial in range(5):
        C = rng.integers(0, 100, (1, W, D)).astype(np.uint32)
        C[0, 0, 1:] = oracle.INV32
        disp, acc = oracle.dp(C, P, return_acc=True)
        best = None
        for seq in itertools.product(range(D), repeat=W):
            if any(abs(seq[i] - seq[i - 1]) > 1 for i in range(1, W)):
                continue
            if any(C[0, i, d] == oracle.INV32 for i, d in enumerate(seq)):
                continue
            tot = sum(int(C[0, i, d]) for i, d in enumerate(seq))
            tot += P * sum(seq[i] != seq[i - 1] for i in range(1, W))
            best = tot if best is None else min(best, tot)
        assert int(acc[0].min()) == best
        d = disp[0]
        assert all(abs(int(d[i]) - int(d[i - 1])) <= 1 for i in range(1, W))
        got = sum(int(C[0, i, d[i]]) for i in range(W)) + P * sum(d[i] != d[i - 1] for i in range(1, W))
        assert got == best


def test_dp_reference_backtrack_is_shifted_by_one():
    C = rng.integers(0, 1000, (3, 9, 5)).astype(np.uint32)
    a = oracle.dp(C, 20)
    b = oracle.dp(C, 20, flags=oracle.DP_REF_BACKTRACK)
    # the reference writes d(x-1) at x (and d(0) at x = 0)
    assert (b[:, 1:] == a[:, :-1]).all() and (b[:, 0] == a[:, 0]).all()


# ---- generator -----------------------------------------------------------------

def test_generator_known_answers():
    L, R, gt = synth.workload_pair(synth.PLAIN_384)
    assert synth.hash_bytes(L) == 0x67DA62008B67E205
    assert synth.hash_bytes(R) == 0xF098B1E556AD000E
    assert gt.min() >= 0 and gt.max() <= 15
    L2, R2, gt2 = synth.workload_pair(synth.PLAIN_384)
    assert (L == L2).all() and (R == R2).all() and (gt == gt2).all()


def test_fast_wta_equals_ssd_then_wta():
    L, R = rand_img(10, 16), rand_img(10, 16)
    assert (oracle.fast_wta_ssd(L, R, 5) == oracle.wta(oracle.cost_ssd(L, R, 5))).all()


# ---- reference-geometry start-assign: closed form -----------------------------

@pytest.mark.parametrize("W,H", [(12, 1), (1, 11), (17, 1)])
def test_ref_start_assign_one_row_or_column_closed_form(W, H):
    # On a single row (or column) every column and diagonal path of the
    # reference table (semiglobal.cu:116-230) is one pixel long and comes
    # after the row (column) path in table order; with start-assign
    # (semiglobalcpu.cpp:81) the last traversal starting at each pixel sets
    # S = C there and nothing is added afterwards, so S = C exactly.
    C = oracle.cost_bt2(rand_img(H, W), rand_img(H, W), 5)
    S = oracle.sgm_ref(C, 8, 32, flags=oracle.SGM_START_ASSIGN)
    v = C != oracle.INV16
    assert (S[v] == C[v].astype(np.int32)).all() and (S[~v] == oracle.INF_I32).all()
    # without it each pixel sums the two traversals of the long path and
    # one C per one-pixel traversal over it (the table's W != H end points
    # pile some of them onto the end pixels; the visit counts are pinned by
    # test_reference_path_coverage's closed forms)
    S0 = oracle.sgm_ref(C, 8, 32, flags=0)
    long_dir = (0, 1) if H == 1 else (2, 3)
    Lsum = sum(oracle.sgm_lr(C, 8, 32, d).astype(np.int64) for d in long_dir)
    visits = oracle.sgm_ref_visits(W, H)
    short = sum(visits[d] for d in range(8) if d not in long_dir)[:, :, None]
    assert (S0[v] == (Lsum + short * C.astype(np.int64))[v]).all()


# ---- SGM restart after a predecessor with no valid cell ------------------------

def test_sgm_restarts_after_an_all_invalid_pixel():
    # one row, direction (1,0): pixel k has no valid cell, so the path
    # restarts at k + 1 exactly as if the row began there (L = C, then the
    # textbook recurrence), in the uint16, uint32 and float forms alike
    W, D, k = 9, 4, 4
    C = rng.integers(0, 80, (1, W, D)).astype(np.uint16)
    C[0, k, :] = oracle.INV16
    Lr = oracle.sgm_lr(C, 5, 17, 0)
    tail = oracle.sgm_lr(np.ascontiguousarray(C[:, k + 1:]), 5, 17, 0)
    assert (Lr[0, k] == oracle.INV16).all()
    assert (Lr[:, k + 1:] == tail).all()
    C32 = np.where(C == oracle.INV16, np.uint32(oracle.INV32), C.astype(np.uint32))
    S32 = oracle.sgm_u32(C32, 5, 17, dir_mask=1)
    assert (S32[:, k + 1:] == tail.astype(np.uint32)).all()
    Cf = np.where(C == oracle.INV16, np.float32(np.inf), C.astype(np.float32))
    Sf = oracle.sgm_f32(Cf, 5.0, 17.0, dir_mask=1)
    assert (Sf[:, k + 1:] == tail.astype(np.float32)).all() and np.isinf(Sf[0, k]).all()


# ---- relative costs: the path-cost increments depend on C only through -------
# ---- min(C - Cmin, 2 P2) per pixel -------------------------------------------------

def _relative_costs(C, P2):
    """The device's relative cost form of a uint16 volume: per pixel
    min(C - Cmin, 2 P2) over its valid cells, 2 P2 on invalid cells, and
    0xFFFF on every cell of a pixel without a valid cell."""
    valid = C != oracle.INV16
    c = C.astype(np.int64)
    cmin = np.where(valid, c, 1 << 20).min(axis=2, keepdims=True)
    rel = np.where(valid, np.minimum(c - cmin, 2 * P2), 2 * P2)
    rel = np.where(valid.any(axis=2, keepdims=True), rel, oracle.INV16)
    return rel.astype(np.uint16), valid


@pytest.mark.parametrize("P1,P2", [(8, 32), (0, 17), (40, 5), (100, 63), (3, 0), (8, 1)])
def test_sgm_relative_costs_give_the_same_increments(P1, P2):
    # every direction's E = L - C on the valid cells is the same from the
    # true costs and from the relative costs (edge cells d > x invalid, a
    # few scattered invalid cells, pixels without a valid cell, BT-range
    # and far larger costs)
    H, W, D = 9, 23, 12
    C = rng.integers(0, 511, (H, W, D)).astype(np.uint16)
    C[rng.random((H, W, D)) < 0.1] = rng.integers(600, 3000)
    for x in range(W):
        C[:, x, x + 1:] = oracle.INV16
    C[rng.random((H, W, D)) < 0.05] = oracle.INV16
    C[4, 7, :] = oracle.INV16
    C[0, 3, :] = oracle.INV16
    C[:, 0, 0] = np.where(np.arange(H) % 3 == 0, oracle.INV16, C[:, 0, 0])
    rel, valid = _relative_costs(C, P2)
    for direction in range(8):
        e_true = oracle.sgm_lr(C, P1, P2, direction).astype(np.int64) - C
        e_rel = oracle.sgm_lr(rel, P1, P2, direction).astype(np.int64) - rel
        assert (e_true[valid] == e_rel[valid]).all(), direction
        assert (e_true[valid] >= 0).all() and (e_true[valid] <= P2).all()


@pytest.mark.parametrize("P1,P2,mask", [(7, 32, 0xFF), (0, 127, 0xFF), (30, 5, 0x5A), (3, 0, 0x81)])
def test_sgm_u32_relative_costs_give_the_same_sums(P1, P2, mask):
    # the uint32 form (SSD costs up to 2^21): the oracle's S over the true
    # costs equals its S over the relative costs (invalid cells 2 P2,
    # all-invalid pixels invalid) plus n (C - CR) on every valid cell, i.e.
    # every direction adds the same increments; all-invalid pixels stay
    # invalid in both
    H, W, D = 11, 26, 20
    C = rng.integers(0, 1 << 21, (H, W, D)).astype(np.uint32)
    C[rng.random((H, W, D)) < 0.5] = rng.integers(0, 200)  # near-minimum cells, small gaps
    for x in range(W):
        C[:, x, x + 1:] = oracle.INV32
    C[rng.random((H, W, D)) < 0.05] = oracle.INV32
    C[5, 9, :] = oracle.INV32
    C[2, 20, :] = oracle.INV32
    valid = C != oracle.INV32
    c = C.astype(np.int64)
    cmin = np.where(valid, c, 1 << 40).min(axis=2, keepdims=True)
    rel = np.where(valid, np.minimum(c - cmin, 2 * P2), 2 * P2)
    rel = np.where(valid.any(axis=2, keepdims=True), rel, oracle.INV32).astype(np.uint32)
    n = bin(mask).count("1")
    s_true = oracle.sgm_u32(C, P1, P2, mask).astype(np.int64)
    s_rel = oracle.sgm_u32(rel, P1, P2, mask).astype(np.int64)
    assert (s_true[valid] == (s_rel + n * (c - rel.astype(np.int64)))[valid]).all()
    none = ~valid.any(axis=2)
    assert (s_true[none] == oracle.INV32).all() and (s_rel[none] == oracle.INV32).all()


def test_sgm_relative_costs_fail_with_a_tighter_clip():
    # the clip at 2 P2 is needed: clipping one below it changes some
    # increments on random volumes (so the test above can fail)
    H, W, D, P1, P2 = 12, 40, 16, 8, 32
    C = rng.integers(0, 511, (H, W, D)).astype(np.uint16)
    valid = np.ones_like(C, bool)
    c = C.astype(np.int64)
    cmin = c.min(axis=2, keepdims=True)
    changed = False
    for clip in (2 * P2 - 1, P2):
        tight = np.minimum(c - cmin, clip).astype(np.uint16)
        for direction in range(8):
            e_true = oracle.sgm_lr(C, P1, P2, direction).astype(np.int64) - C
            e_t = oracle.sgm_lr(tight, P1, P2, direction).astype(np.int64) - tight
            changed |= bool((e_true[valid] != e_t[valid]).any())
    assert changed


# ---- SGM over uint32 and float volumes -------------------------------------------

def _u16_to_u32(C):
    return np.where(C == oracle.INV16, np.uint32(oracle.INV32), C.astype(np.uint32))


@pytest.mark.parametrize("mask", [0xFF, 0x01, 0x5A])
def test_sgm_u32_equals_uint16_sgm_on_its_domain(mask):
    # the uint32 form (saturating, never reached here) and the pinned
    # uint16 form are separate code; on uint16 volumes they must agree
    H, W, D = 9, 14, 7
    C = oracle.cost_bt2(rand_img(H, W), rand_img(H, W), D)
    C[2, 5, :] = oracle.INV16  # a pixel with no valid cell
    S16 = oracle.sgm_std(C, 8, 32, mask)
    S32 = oracle.sgm_u32(_u16_to_u32(C), 8, 32, mask)
    assert (S32 == _u16_to_u32(S16)).all()


def test_sgm_u32_zero_penalties_and_saturation():
    C = rng.integers(0, 1 << 28, (5, 8, 6)).astype(np.uint32)
    S = oracle.sgm_u32(C, 0, 0)
    assert (S.astype(np.uint64) == 8 * C.astype(np.uint64)).all()
    # costs near 2^32: every sum saturates at the invalid value
    big = np.full((3, 4, 3), 0xF0000000, np.uint32)
    assert (oracle.sgm_u32(big, 0, 0) == oracle.INV32).all()


def test_sgm_f32_dyadic_costs_equal_integer_sgm():
    # costs and penalties k/8 with small k: every float operation is exact,
    # so the float form is the integer form divided by 8
    H, W, D = 8, 11, 6
    C = oracle.cost_bt2(rand_img(H, W), rand_img(H, W), D)
    Cf = np.where(C == oracle.INV16, np.float32(np.inf), C.astype(np.float32) / 8)
    for mask in (0xFF, 0x81):
        S16 = oracle.sgm_std(C, 8, 32, mask)
        Sf = oracle.sgm_f32(Cf, 1.0, 4.0, mask)
        v = C != oracle.INV16
        assert (Sf[v] == S16[v].astype(np.float32) / 8).all() and np.isinf(Sf[~v]).all()


def test_sgm_f32_nan_and_negative_infinity_read_as_invalid():
    C = rng.random((4, 6, 5)).astype(np.float32)
    Cn = C.copy()
    Cn[1, 2, 3] = np.nan
    Cn[2, 4, 0] = -np.inf
    Ci = C.copy()
    Ci[1, 2, 3] = np.inf
    Ci[2, 4, 0] = np.inf
    a = oracle.sgm_f32(Cn, 0.1, 0.5)
    b = oracle.sgm_f32(Ci, 0.1, 0.5)
    assert np.array_equal(a, b)


# ---- DP over uint16 and float volumes ------------------------------------------

def test_dp_f32_dyadic_costs_equal_integer_dp():
    C = rng.integers(0, 400, (4, 23, 6)).astype(np.uint32)
    C[:, 0, 1:] = oracle.INV32
    Cf = np.where(C == oracle.INV32, np.float32(np.inf), C.astype(np.float32) / 8)
    for flags in (0, oracle.DP_START_ZERO, oracle.DP_REF_BACKTRACK):
        assert (oracle.dp_f32(Cf, 20 / 8, flags) == oracle.dp(C, 20, flags)).all()


def test_dp_u16_reads_inv16_as_invalid():
    C = oracle.cost_bt2(rand_img(5, 17), rand_img(5, 17), 6)
    assert (oracle.dp_u16(C, 20) == oracle.dp(_u16_to_u32(C), 20)).all()
