"""Serving loop: continuous batching of stereo frames from many streams.

Frames arrive on a schedule (Poisson per stream in the benchmark). The loop
admits a queued frame when a slot is free and the device-memory footprint of
the frames in flight plus the new one stays within the budget; admitted
frames run concurrently, one per slot, each slot owning a HIP stream and a
tdvh context (so slots never share scratch). Inputs and outputs stay device-
resident per stream. Latency is measured on the device clock: from the
frame's arrival time to the completion event of its disparity map.

The scheduling logic is independent of the executor, so it is tested on the
CPU with a fake executor; GpuExecutor runs tdvh_stereo_bt_sgm.
"""
import collections
import gc
import time

BT_TERM_WORDS = 3


def frame_footprint(width, height, depth):
    """Device bytes a frame needs while in flight: byte path-cost planes of
    8 directions, BT term planes, both images and the disparity map."""
    m = (depth + 31) // 32
    planes = 8 * width * height * 32 * m
    terms = 4 * BT_TERM_WORDS * (width * height + (width + 16 * m) * height)
    return planes + terms + 2 * width * height + 2 * width * height


class Frame:
    __slots__ = ("t_arrival", "stream", "index", "width", "height", "depth", "payload",
                 "t_start", "t_done", "slot")

    def __init__(self, t_arrival, stream, index, width, height, depth, payload=None):
        self.t_arrival = t_arrival
        self.stream = stream
        self.index = index
        self.width, self.height, self.depth = width, height, depth
        self.payload = payload
        self.t_start = self.t_done = None
        self.slot = None

    @property
    def cells(self):
        return self.width * self.height * self.depth

    @property
    def footprint(self):
        return frame_footprint(self.width, self.height, self.depth)


class Scheduler:
    """Admission and completion bookkeeping (no device calls).

    executor.launch(frame, slot) starts a frame; executor.poll(slot) returns
    the completion time (seconds on the schedule clock) or None;
    executor.now() is the schedule clock.

    Admission is in arrival order (largest_first: the largest queued frame
    first, then arrival order; smallest_first: the smallest first, except
    that frames queued for max_wait_s seconds or longer go before all others,
    in arrival order), needs a free slot and memory within
    budget_bytes, and, with cell_budget, keeps the cells (W*H*D) of the
    frames in flight within that budget, so that large frames do not share
    the device with each other (a frame always starts on an idle device).
    With backfill_s, frames behind a head frame that does not fit yet may
    start first while the head has waited less than backfill_s seconds.
    With join_cells, only frames of at most that many cells start while
    other frames are in flight: larger frames then run one after another in
    arrival order, as on one slot, and small frames join them instead of
    waiting behind them.
    """

    def __init__(self, executor, slots, budget_bytes, cell_budget=None, backfill_s=None,
                 largest_first=False, join_cells=None, smallest_first=False, max_wait_s=None):
        if largest_first and smallest_first:
            raise ValueError("largest_first and smallest_first exclude each other")
        self.ex = executor
        self.nslots = slots
        self.budget = budget_bytes
        self.cell_budget = cell_budget
        self.backfill_s = backfill_s
        self.largest_first = largest_first
        self.join_cells = join_cells
        self.smallest_first = smallest_first
        self.max_wait_s = max_wait_s
        self.in_use = 0
        self.in_flight_cells = 0
        self.peak_in_use = 0
        self.max_concurrent = 0

    def _fits(self, f, busy):
        if self.in_use + f.footprint > self.budget:
            return False
        if busy and self.join_cells is not None and f.cells > self.join_cells:
            return False
        if self.cell_budget is not None and busy:
            return self.in_flight_cells + f.cells <= self.cell_budget
        return True

    def run(self, frames, spin_sleep=0.0):
        """Serves `frames`; Python's cyclic garbage collector is paused while
        the loop runs (a full collection takes tens of milliseconds, which
        would land on some frame's latency)."""
        was_enabled = gc.isenabled()
        gc.disable()
        try:
            return self._run(frames, spin_sleep)
        finally:
            if was_enabled:
                gc.enable()

    def _run(self, frames, spin_sleep):
        pending = sorted(frames, key=lambda f: (f.t_arrival, f.stream, f.index))
        queue = collections.deque()
        free = list(range(self.nslots))
        busy = {}
        done = []
        i = 0
        while i < len(pending) or queue or busy:
            now = self.ex.now()
            while i < len(pending) and pending[i].t_arrival <= now:
                f = pending[i]
                if f.footprint > self.budget:
                    raise ValueError("frame %dx%dx%d needs more than the memory budget"
                                     % (f.width, f.height, f.depth))
                queue.append(f)
                i += 1
            # admit in arrival order (or largest frame first) while a slot is
            # free and the budgets allow (optionally backfilling past a
            # blocked head frame)
            if self.largest_first and len(queue) > 1:
                queue = collections.deque(sorted(queue, key=lambda f: (-f.cells, f.t_arrival)))
            elif self.smallest_first and len(queue) > 1:
                aged = self.max_wait_s if self.max_wait_s is not None else float("inf")
                queue = collections.deque(sorted(
                    queue, key=lambda f: (0, f.t_arrival, 0) if now - f.t_arrival >= aged
                    else (1, f.cells, f.t_arrival)))
            while queue and free:
                pick = None
                for j, f in enumerate(queue):
                    if self._fits(f, busy):
                        pick = j
                        break
                    if self.backfill_s is None or now - queue[0].t_arrival >= self.backfill_s:
                        break
                if pick is None:
                    break
                f = queue[pick]
                del queue[pick]
                slot = free.pop()
                f.slot = slot
                f.t_start = now
                self.in_use += f.footprint
                self.in_flight_cells += f.cells
                self.peak_in_use = max(self.peak_in_use, self.in_use)
                self.ex.launch(f, slot)
                busy[slot] = f
                self.max_concurrent = max(self.max_concurrent, len(busy))
            progressed = False
            for slot in list(busy):
                t = self.ex.poll(slot)
                if t is not None:
                    f = busy.pop(slot)
                    f.t_done = t
                    self.in_use -= f.footprint
                    self.in_flight_cells -= f.cells
                    free.append(slot)
                    done.append(f)
                    progressed = True
            if not progressed and spin_sleep:
                time.sleep(spin_sleep)
        return done


def latency_by_shape(frames):
    """latency_stats per (width, height, depth)."""
    shapes = sorted({(f.width, f.height, f.depth) for f in frames})
    out = {}
    for sh in shapes:
        st = latency_stats([f for f in frames if (f.width, f.height, f.depth) == sh])
        out["%dx%d/D%d" % sh] = {k: st[k] for k in ("frames", "p50_ms", "p99_ms", "max_ms")}
    return out


def latency_stats(frames):
    lat = sorted(f.t_done - f.t_arrival for f in frames)
    if not lat:
        return {}

    def pct(q):
        k = min(len(lat) - 1, max(0, int(round(q * (len(lat) - 1)))))
        return lat[k]

    t0 = min(f.t_arrival for f in frames)
    t1 = max(f.t_done for f in frames)
    cells = sum(f.cells for f in frames)
    return {"frames": len(lat), "p50_ms": pct(0.50) * 1e3, "p99_ms": pct(0.99) * 1e3,
            "max_ms": lat[-1] * 1e3, "mean_ms": sum(lat) / len(lat) * 1e3,
            "span_s": t1 - t0, "mdes": cells / max(t1 - t0, 1e-9) / 1e6}


class GpuExecutor:
    """Runs frames with tdvh_stereo_bt_sgm, one HIP stream and one tdvh
    context per slot; times are device event times relative to start()."""

    def __init__(self, slots, device=0):
        import torch

        import thundervision_amd.tdvh as tv
        self.torch = torch
        self.tv = tv
        self.streams = [torch.cuda.Stream(device=device) for _ in range(slots)]
        self.ctxs = [tv.Context(device) for _ in range(slots)]
        self.events = [None] * slots
        self.t0_host = None
        self.t0_event = None

    def warm(self, frames):
        """Runs one frame of every shape in `frames` on every slot and polls
        them to completion, so that scratch, per-shape tables and the event
        and timing paths are initialised before start()."""
        torch = self.torch
        firsts = {}
        for f in frames:
            firsts.setdefault((f.width, f.height, f.depth), f)
        shapes = sorted(firsts.values(), key=lambda f: -f.cells)
        self.start()
        for f in shapes:
            for slot in range(len(self.streams)):
                self.launch(f, slot)
            for slot in range(len(self.streams)):
                while self.poll(slot) is None:
                    time.sleep(0)
        torch.cuda.synchronize()
        gc.collect()

    def start(self):
        torch = self.torch
        torch.cuda.synchronize()
        self.t0_event = torch.cuda.Event(enable_timing=True)
        self.t0_event.record(self.streams[0])
        self.t0_host = time.perf_counter()

    def now(self):
        return time.perf_counter() - self.t0_host

    def launch(self, frame, slot):
        torch = self.torch
        s = self.streams[slot]
        L, R, disp = frame.payload
        with torch.cuda.stream(s):
            self.tv.stereo_bt_sgm(L, R, frame.depth, out=disp, ctx=self.ctxs[slot], stream=s)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(s)
        self.events[slot] = ev

    def poll(self, slot):
        ev = self.events[slot]
        if ev is None or not ev.query():
            return None
        self.events[slot] = None
        return self.t0_event.elapsed_time(ev) / 1e3


# The served workload: 16 streams, 4 each of these (width, height, depth).
SERVED_STREAMS = [(640, 480, 64)] * 4 + [(1280, 720, 128)] * 4 + [(1280, 720, 192)] * 4 + \
                 [(1920, 1080, 256)] * 4


def poisson_schedule(nstreams, rate_total, duration, seed_base=2000):
    """Arrival times per stream: Poisson at rate_total / nstreams each."""
    from thundervision_amd import synth
    out = []
    rate = rate_total / nstreams
    for s in range(nstreams):
        t, k = 0.0, 0
        while True:
            t += synth.exponential(seed_base + s, k, rate)
            if t > duration:
                break
            out.append((t, s, k))
            k += 1
    out.sort()
    return out
