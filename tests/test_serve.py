"""Serving loop: admission/budget logic on the CPU with a fake executor, and
on the GPU: every frame's disparities are the same alone and in the mix."""
import numpy as np
import pytest

from thundervision_amd import serve


class FakeExecutor:
    """Deterministic clock; a frame takes cells / rate seconds."""

    def __init__(self, rate=1e9):
        self.t = 0.0
        self.rate = rate
        self.running = {}
        self.launched = []

    def now(self):
        self.t += 1e-4
        return self.t

    def launch(self, f, slot):
        self.running[slot] = self.t + f.cells / self.rate
        self.launched.append((f.stream, f.index, slot))

    def poll(self, slot):
        if slot in self.running and self.running[slot] <= self.t:
            return self.running.pop(slot)
        return None


def frames_from(schedule, shapes):
    return [serve.Frame(t, s, k, *shapes[s]) for t, s, k in schedule]


def test_schedule_is_poisson_like_and_seeded():
    a = serve.poisson_schedule(16, 600.0, 2.0)
    b = serve.poisson_schedule(16, 600.0, 2.0)
    assert a == b
    assert 1000 < len(a) < 1400  # ~1200 expected
    assert all(x[0] <= y[0] for x, y in zip(a, a[1:]))


def test_budget_and_slots_are_respected():
    sched = serve.poisson_schedule(16, 600.0, 0.5)
    frames = frames_from(sched, serve.SERVED_STREAMS)
    big = serve.frame_footprint(1920, 1080, 256)
    ex = FakeExecutor(rate=2e11)
    sch = serve.Scheduler(ex, slots=4, budget_bytes=int(2.5 * big))
    done = sch.run(frames)
    assert len(done) == len(frames)
    assert sch.peak_in_use <= int(2.5 * big)
    assert sch.max_concurrent <= 4
    assert all(f.t_done >= f.t_arrival for f in done)
    # frames of one stream complete in arrival order under FIFO admission
    for s in range(16):
        idx = [f.index for f in sorted(done, key=lambda f: f.t_start) if f.stream == s]
        assert idx == sorted(idx)


def test_cell_budget_keeps_large_frames_apart():
    sched = serve.poisson_schedule(16, 600.0, 0.5)
    frames = frames_from(sched, serve.SERVED_STREAMS)
    big = 1920 * 1080 * 256
    ex = FakeExecutor(rate=2e11)
    sch = serve.Scheduler(ex, slots=4, budget_bytes=64 << 30, cell_budget=1.3 * big)
    seen = []
    launch = ex.launch

    def watch(f, slot):
        seen.append(sch.in_flight_cells)
        launch(f, slot)

    ex.launch = watch
    done = sch.run(frames)
    assert len(done) == len(frames)
    # with more than one frame in flight the cells stay within the budget;
    # a frame larger than the budget would still start on an idle device
    assert max(seen) <= 1.3 * big
    assert sch.max_concurrent >= 2


def test_backfill_passes_a_blocked_head_only_for_a_while():
    # a large frame in flight; a second large frame (blocked by the cell
    # budget) heads the queue; small frames behind it may start while the
    # head has waited less than backfill_s, not after
    big = (1920, 1080, 256)
    small = (320, 240, 32)
    frames = [serve.Frame(0.0, 0, 0, *big), serve.Frame(0.00005, 1, 0, *big),
              serve.Frame(0.0001, 2, 0, *small), serve.Frame(0.02, 3, 0, *small)]
    ex = FakeExecutor(rate=2e10)  # a large frame takes ~26 ms
    sch = serve.Scheduler(ex, slots=4, budget_bytes=64 << 30,
                          cell_budget=1.3 * 1920 * 1080 * 256, backfill_s=0.005)
    done = {(f.stream, f.index): f for f in sch.run(frames)}
    # the early small frame ran next to the first large one
    assert done[(2, 0)].t_start < done[(0, 0)].t_done
    # the late small frame arrived after the backfill window: it waits for
    # the blocked large frame to start first
    assert done[(3, 0)].t_start >= done[(1, 0)].t_start


def test_join_cells_runs_large_frames_one_at_a_time_in_arrival_order():
    # with join_cells only small frames start next to frames in flight:
    # medium and large frames run alone, in arrival order, as on one slot
    sched = serve.poisson_schedule(16, 600.0, 0.5)
    frames = frames_from(sched, serve.SERVED_STREAMS)
    big = 1920 * 1080 * 256
    join = 0.1 * big
    ex = FakeExecutor(rate=2e11)
    sch = serve.Scheduler(ex, slots=4, budget_bytes=64 << 30, cell_budget=1.3 * big,
                          backfill_s=0.003, join_cells=join)
    starts = []
    launch = ex.launch

    def watch(f, slot):
        starts.append((f, sch.in_flight_cells - f.cells))  # the others in flight
        launch(f, slot)

    ex.launch = watch
    done = sch.run(frames)
    assert len(done) == len(frames)
    # every frame above join_cells started on an idle device
    assert all(inflight == 0 for f, inflight in starts if f.cells > join)
    # small frames did join frames in flight
    assert any(inflight > 0 for f, inflight in starts if f.cells <= join)
    # frames above join_cells started in arrival order
    large = [f for f, _ in starts if f.cells > join]
    assert all(a.t_arrival <= b.t_arrival for a, b in zip(large, large[1:]))
    # and never overlapped one another
    spans = sorted((f.t_start, f.t_done) for f in done if f.cells > join)
    assert all(a[1] <= b[0] + 1e-12 for a, b in zip(spans, spans[1:]))


def test_smallest_first_with_an_age_bound():
    # one slot; while a large frame runs, a large, a medium and a small frame
    # queue up in that order: smallest first starts them small, medium,
    # large; with an age bound they have all passed, arrival order decides
    big, mid, small = (1920, 1080, 256), (1280, 720, 128), (320, 240, 32)
    def order(**kw):
        frames = [serve.Frame(0.0, 0, 0, *big), serve.Frame(0.001, 1, 0, *big),
                  serve.Frame(0.0015, 2, 0, *mid), serve.Frame(0.002, 3, 0, *small)]
        ex = FakeExecutor(rate=2e10)  # the first frame takes ~26 ms
        done = serve.Scheduler(ex, slots=1, budget_bytes=64 << 30, **kw).run(frames)
        return [f.stream for f in sorted(done, key=lambda f: f.t_start)]
    assert order() == [0, 1, 2, 3]
    assert order(smallest_first=True) == [0, 3, 2, 1]
    assert order(smallest_first=True, max_wait_s=0.05) == [0, 3, 2, 1]
    assert order(smallest_first=True, max_wait_s=0.01) == [0, 1, 2, 3]
    assert order(largest_first=True) == [0, 1, 2, 3]
    with pytest.raises(ValueError):
        serve.Scheduler(FakeExecutor(), 1, 1 << 30, largest_first=True, smallest_first=True)


def test_gc_is_paused_while_serving_and_restored():
    import gc
    ex = FakeExecutor()
    states = []
    launch = ex.launch
    ex.launch = lambda f, slot: (states.append(gc.isenabled()), launch(f, slot))
    sch = serve.Scheduler(ex, slots=2, budget_bytes=64 << 30)
    assert gc.isenabled()
    sch.run([serve.Frame(0.0, 0, 0, 64, 48, 16), serve.Frame(0.001, 1, 0, 64, 48, 16)])
    assert states == [False, False] and gc.isenabled()


def test_oversized_frame_is_rejected():
    ex = FakeExecutor()
    sch = serve.Scheduler(ex, slots=2, budget_bytes=1000)
    with pytest.raises(ValueError, match="budget"):
        sch.run([serve.Frame(0.0, 0, 0, 640, 480, 64)])


def test_latency_stats():
    fs = []
    for k in range(100):
        f = serve.Frame(k * 0.01, 0, k, 10, 10, 10)
        f.t_done = f.t_arrival + (k + 1) * 1e-3
        fs.append(f)
    st = serve.latency_stats(fs)
    assert abs(st["p50_ms"] - 51) < 1.01 and abs(st["p99_ms"] - 99) < 1.01


@pytest.mark.gpu
def test_frames_identical_alone_and_in_the_mix(tv, dev):
    import torch
    from thundervision_amd import synth
    shapes = [(160, 120, 32), (320, 240, 64), (200, 150, 80), (640, 480, 128)]
    streams = [shapes[i % 4] for i in range(8)]
    sched = serve.poisson_schedule(8, 2000.0, 0.05, seed_base=3000)
    frames = []
    for t, s, k in sched:
        W, H, D = streams[s]
        L, R, _ = synth.pair(W, H, D - 1, 8, 3000 + s, frame=k)
        Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
        disp = torch.full((H, W), -1, dtype=torch.int16, device=dev)
        frames.append(serve.Frame(t, s, k, W, H, D, (Ld, Rd, disp)))
    ex = serve.GpuExecutor(slots=4)
    sch = serve.Scheduler(ex, slots=4, budget_bytes=8 << 30)
    ex.start()
    done = sch.run(frames)
    torch.cuda.synchronize()
    assert len(done) == len(frames) and sch.max_concurrent >= 2
    for f in frames:
        Ld, Rd, disp = f.payload
        alone = tv.stereo_bt_sgm(Ld, Rd, f.depth)
        assert torch.equal(alone, disp), (f.stream, f.index)
    st = serve.latency_stats(done)
    assert st["p99_ms"] >= st["p50_ms"] > 0


@pytest.mark.gpu
def test_cfg5_served_mix_frames_identical_alone(tv, dev):
    """The served workload: 16 streams (4 each at 640x480/D64, 1280x720/D128,
    1280x720/D192, 1920x1080/D256), Poisson arrivals at 600 pairs/s in total,
    a fresh frame per arrival (rectangles drifting with the frame index),
    64 GB budget, 4 slots with the benchmark's cell budget, backfill and join
    size.
    Every served disparity map equals the same frame computed alone."""
    import torch
    from thundervision_amd import synth
    shapes = serve.SERVED_STREAMS
    sched = serve.poisson_schedule(len(shapes), 600.0, 0.15)
    frames = []
    for t, s, k in sched:
        W, H, D = shapes[s]
        L, R, _ = synth.pair(W, H, D - 1, 12, 2000 + s, frame=k)
        Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
        disp = torch.full((H, W), -1, dtype=torch.int16, device=dev)
        frames.append(serve.Frame(t, s, k, W, H, D, (Ld, Rd, disp)))
    big = max(W * H * D for W, H, D in shapes)
    ex = serve.GpuExecutor(slots=4)
    sch = serve.Scheduler(ex, 4, 64 << 30, cell_budget=1.3 * big, backfill_s=0.003,
                          join_cells=0.1 * big)
    ex.warm(frames)
    for f in frames:
        f.payload[2].fill_(-1)
    ex.start()
    done = sch.run(frames)
    torch.cuda.synchronize()
    assert len(done) == len(frames) >= 40
    assert {(f.width, f.height, f.depth) for f in frames} == set(shapes)
    assert sch.max_concurrent >= 2 and sch.peak_in_use <= 64 << 30
    for f in frames:
        Ld, Rd, disp = f.payload
        alone = tv.stereo_bt_sgm(Ld, Rd, f.depth)
        assert torch.equal(alone, disp), (f.stream, f.index)
