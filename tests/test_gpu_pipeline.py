"""The fixed-shape graph pipeline (tdvh_pipeline_*): every replay equals the
oracle's prefilter -> BT + SGM -> WTA -> postfilter on the same pair, with
eager library work and synchronisations between replays."""
import numpy as np
import pytest
import torch

import oracle
from thundervision_amd import synth

pytestmark = pytest.mark.gpu


def want(L, R, D, flags):
    if flags & 1:
        L, R = oracle.gauss5(L), oracle.gauss5(R)
    d = oracle.stereo_bt_sgm(L, R, D)
    if flags & 2:
        d = oracle.open_disp(d, 3)
    return d


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_pipeline_replays(tv, dev, flags):
    W, H, D = 200, 120, 64
    p = tv.Pipeline(W, H, D, flags=flags)
    try:
        for k in range(4):
            L, R, _ = synth.pair(W, H, D - 1, 6, 40 + k)
            Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
            got = p.run(Ld, Rd)
            # eager work of the library on another context between replays
            tv.stereo_bt_sgm(Rd, Ld, 32)
            torch.cuda.synchronize()
            assert (got.cpu().numpy() == want(L, R, D, flags)).all(), k
    finally:
        p.close()


def test_pipeline_720p_pitched(tv, dev):
    L, R, _ = synth.workload_pair(synth.PLAIN_720P)
    H, W = L.shape
    p = tv.Pipeline(W, H, 128, flags=3)
    try:
        big = torch.zeros((2, H, W + 64), dtype=torch.uint8, device=dev)
        big[0, :, :W] = torch.from_numpy(L).to(dev)
        big[1, :, :W] = torch.from_numpy(R).to(dev)
        w = want(L, R, 128, 3)
        for _ in range(3):
            out = torch.full((H, W + 8), -7, dtype=torch.int16, device=dev)
            p.run(big[0, :, :W], big[1, :, :W], out=out[:, :W])
            torch.cuda.synchronize()
            assert (out[:, :W].cpu().numpy() == w).all()
            assert (out[:, W:] == -7).all()
    finally:
        p.close()


def test_pipeline_host_and_device_buffers_interleaved(tv, dev):
    """Pinned host buffers (copied in and out around the pipeline's own
    graph) and device buffers (a graph over the caller's pointers, replayed
    while the pointers stay the same) in alternation; new pixels in the same
    device buffers between replays."""
    W, H, D = 160, 96, 48
    p = tv.Pipeline(W, H, D, flags=3)
    Ld = torch.empty((H, W), dtype=torch.uint8, device=dev)
    Rd = torch.empty_like(Ld)
    outd = torch.empty((H, W), dtype=torch.int16, device=dev)
    try:
        for k in range(6):
            L, R, _ = synth.pair(W, H, D - 1, 5, 70 + k)
            if k % 2 == 0:
                Lh = torch.from_numpy(L).pin_memory()
                Rh = torch.from_numpy(R).pin_memory()
                outh = torch.full((H, W), -1, dtype=torch.int16).pin_memory()
                p.run(Lh, Rh, out=outh)
                torch.cuda.synchronize()
                got = outh.numpy()
            else:
                Ld.copy_(torch.from_numpy(L))
                Rd.copy_(torch.from_numpy(R))
                p.run(Ld, Rd, out=outd)
                torch.cuda.synchronize()
                got = outd.cpu().numpy()
            assert (got == want(L, R, D, 3)).all(), k
    finally:
        p.close()


def test_two_pipelines_on_two_streams(tv, dev):
    """Two pipelines, each with its own context, stream, device buffers and
    pinned output, take alternate frames with no synchronisation between
    frames (so one stream's copies and kernels overlap the other's): every
    frame equals the oracle chain."""
    W, H, D = 160, 96, 48
    frames = [synth.pair(W, H, D - 1, 5, 90 + k)[:2] for k in range(6)]
    lanes = []
    for _ in range(2):
        ctx = tv.Context(0)
        lanes.append(dict(ctx=ctx, st=torch.cuda.Stream(), p=tv.Pipeline(W, H, D, flags=3, ctx=ctx),
                          L=torch.empty((H, W), dtype=torch.uint8, device=dev),
                          R=torch.empty((H, W), dtype=torch.uint8, device=dev),
                          out=torch.empty((H, W), dtype=torch.int16, device=dev)))
    hosts = [(torch.from_numpy(L).pin_memory(), torch.from_numpy(R).pin_memory()) for L, R in frames]
    outs = [torch.full((H, W), -1, dtype=torch.int16).pin_memory() for _ in frames]
    try:
        for k, ((Lh, Rh), oh) in enumerate(zip(hosts, outs)):
            ln = lanes[k % 2]
            with torch.cuda.stream(ln["st"]):
                ln["L"].copy_(Lh, non_blocking=True)
                ln["R"].copy_(Rh, non_blocking=True)
                ln["p"].run(ln["L"], ln["R"], out=ln["out"], stream=ln["st"])
                oh.copy_(ln["out"], non_blocking=True)
        torch.cuda.synchronize()
        for k, ((L, R), oh) in enumerate(zip(frames, outs)):
            assert (oh.numpy() == want(L, R, D, 3)).all(), k
    finally:
        for ln in lanes:
            ln["p"].close()
            ln["ctx"].close()
