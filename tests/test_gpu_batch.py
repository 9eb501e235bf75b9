"""Batched launch (per-phase launches or persistent work queues): every
pair's disparities equal the plain per-pair call's (bit-exact), across graph
replays, eager runs and every tuning knob."""
import numpy as np
import pytest
import torch

import oracle
from thundervision_amd import synth

pytestmark = pytest.mark.gpu


def make_pairs(specs, dev):
    out = []
    for s in specs:
        L, R, _ = synth.pair(s["width"], s["height"], s["dmax"], s["nrect"], s["seed"])
        Ld = torch.from_numpy(L).to(dev)
        Rd = torch.from_numpy(R).to(dev)
        disp = torch.full((s["height"], s["width"]), -1, dtype=torch.int16, device=dev)
        out.append((Ld, Rd, s["depth"], disp))
    return out


MIXED = [dict(width=W, height=H, dmax=D - 1, nrect=6, depth=D, seed=100 + i)
         for i, (W, H, D) in enumerate([(64, 48, 16), (97, 61, 80), (160, 90, 32), (33, 120, 33),
                                        (200, 40, 160), (48, 48, 256), (1, 9, 4), (130, 1, 8)])]


def test_batch_matches_plain_and_oracle(tv, dev):
    pairs = make_pairs(MIXED, dev)
    b = tv.Batch(pairs)
    b.run(use_graph=False)
    b.wait()
    for (Ld, Rd, D, disp), s in zip(pairs, MIXED):
        plain = tv.stereo_bt_sgm(Ld, Rd, D).cpu().numpy()
        assert (disp.cpu().numpy() == plain).all()
        want = oracle.stereo_bt_sgm(Ld.cpu().numpy(), Rd.cpu().numpy(), D)
        assert (plain == want).all()
    assert b.cells() == sum(s["width"] * s["height"] * s["depth"] for s in MIXED)


def test_batch_graph_replays(tv, dev):
    pairs = make_pairs(MIXED[:5], dev)
    want = [tv.stereo_bt_sgm(Ld, Rd, D).cpu().numpy() for Ld, Rd, D, _ in pairs]
    b = tv.Batch(pairs)
    for rep in range(4):
        for _, _, _, disp in pairs:
            disp.fill_(-1)  # eager work between replays
        b.run(use_graph=True)
        b.wait()
        for (_, _, _, disp), w in zip(pairs, want):
            assert (disp.cpu().numpy() == w).all(), rep


@pytest.mark.parametrize("depth", [1, 2, 3])
@pytest.mark.parametrize("lpw,wpg,pf", [(4, 1, 8), (8, 2, 4), (4, 4, 4), (8, 1, 8)])
def test_batch_knobs_do_not_change_results(tv, dev, depth, lpw, wpg, pf):
    pairs = make_pairs(MIXED, dev)
    want = [tv.stereo_bt_sgm(Ld, Rd, D).cpu().numpy() for Ld, Rd, D, _ in pairs]
    b = tv.Batch(pairs)
    b.config(lpw, wpg, pf, depth)
    for use_graph in (False, True, True):
        for _, _, _, disp in pairs:
            disp.fill_(-1)
        b.run(use_graph=use_graph)
        b.wait()
        for (_, _, _, disp), w in zip(pairs, want):
            assert (disp.cpu().numpy() == w).all(), (depth, lpw, wpg, pf, use_graph)


def test_batch_reconfigured_between_runs(tv, dev):
    pairs = make_pairs(MIXED[:6], dev)
    want = [tv.stereo_bt_sgm(Ld, Rd, D).cpu().numpy() for Ld, Rd, D, _ in pairs]
    b = tv.Batch(pairs)
    for knobs in ((4, 1, 8, 3, 1), (8, 2, 4, 1, 0), (4, 4, 8, 2, 2), (8, 1, 8, 3, -1),
                  (8, 1, 8, 2, 1)):
        b.config(*knobs)
        for _, _, _, disp in pairs:
            disp.fill_(-1)
        b.run()
        b.wait()
        for (_, _, _, disp), w in zip(pairs, want):
            assert (disp.cpu().numpy() == w).all(), knobs


def test_autotuner_picks_a_candidate_and_caches(tv, dev, tmp_path):
    from thundervision_amd import autotune
    pairs = make_pairs(MIXED[:3], dev)
    shapes = [(s["width"], s["height"], s["depth"]) for s in MIXED[:3]]
    want = [tv.stereo_bt_sgm(Ld, Rd, D).cpu().numpy() for Ld, Rd, D, _ in pairs]
    b = tv.Batch(pairs)
    path = str(tmp_path / "tune.json")
    r1 = autotune.tune(b, shapes, reps=1, path=path)
    assert not r1["cached"] and 6 <= len(r1["trials"]) <= 13
    assert r1["waves_per_group"] in autotune.WAVES_PER_GROUP
    assert {t["fusion_depth"] for t in r1["trials"]} == {1, 2, 3}
    assert {t["lines_per_wave"] for t in r1["trials"]} == {4, 8}
    r2 = autotune.tune(b, shapes, reps=1, path=path)
    assert r2["cached"] and r2["waves_per_group"] == r1["waves_per_group"]
    b.run()
    b.wait()
    for (_, _, _, disp), w in zip(pairs, want):
        assert (disp.cpu().numpy() == w).all()


@pytest.mark.parametrize("depth,lag", [(1, 0), (2, 1), (3, -1), (3, 4)])
def test_batch_mixed_64_pairs_bit_exact(tv, dev, depth, lag):
    """The 64-pair mix (sizes 320x240..1920x1080, D 32..256 clamped to W/4),
    per-phase launches and persistent work queues, graph replays."""
    specs = synth.batch_specs(64)
    pairs = make_pairs(specs, dev)
    plain = [tv.stereo_bt_sgm(Ld, Rd, D) for Ld, Rd, D, _ in pairs]
    b = tv.Batch(pairs)
    b.config(8 if depth == 3 else 4, 1, 8, depth, lag)
    for rep in range(2):
        for _, _, _, disp in pairs:
            disp.fill_(-1)
        b.run(use_graph=True)
        b.wait()
        for (_, _, _, disp), p, s in zip(pairs, plain, specs):
            assert torch.equal(p, disp), (depth, lag, rep, s)


def test_batch_rejects_bad_input(tv, dev):
    L = torch.zeros((8, 8), dtype=torch.uint8, device=dev)
    d = torch.zeros((8, 8), dtype=torch.int16, device=dev)
    with pytest.raises(tv.TdvhError, match="depth"):
        tv.Batch([(L, L, 300, d)])
    with pytest.raises(tv.TdvhError, match="P2"):
        tv.Batch([(L, L, 16, d)], p2=300)
