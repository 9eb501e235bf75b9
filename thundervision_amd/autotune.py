"""Autotuner of the batched launch (tdvh_batch_config).

Sweeps the batch knobs on the batch itself, one knob at a time from the
current best (coordinate descent): fusion depth (one launch per phase, or a
persistent work-queue launch per depth class for line scans and sums, or
for all three phases), tile height (lines per wave: 4, or 8 with 8-lane
lines), waves per workgroup, the prefetch depth of the SGM line scans and,
for the persistent queues, how far a pair's sums trail its scans.
The fastest setting is cached per batch signature (the multiset of pair
shapes) in a JSON file, so a service tunes a given mix of resolutions once.
"""
import hashlib
import json
import os

import torch

# knob name -> candidate values, in sweep order
KNOBS = (("fusion_depth", (1, 2, 3)), ("lines_per_wave", (4, 8)),
         ("waves_per_group", (1, 2, 4)), ("prefetch", (4, 8)), ("sum_lag", (-1, 1, 4)))
DEFAULT = {"fusion_depth": 1, "lines_per_wave": 4, "waves_per_group": 1, "prefetch": 8,
           "sum_lag": -1}
WAVES_PER_GROUP = dict(KNOBS)["waves_per_group"]


def cache_path():
    return os.environ.get("TDVH_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "thundervision_amd",
                                       "batch_tune.json"))


def signature(shapes):
    """shapes: iterable of (width, height, depth)."""
    s = ",".join("%dx%dx%d" % t for t in sorted(shapes))
    return hashlib.sha1(s.encode()).hexdigest()[:16]


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_batch(batch, reps=3, use_graph=False):
    """Mean milliseconds of one batch run (hipEvents on the current stream)."""
    batch.run(use_graph=use_graph)
    batch.wait()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        batch.run(use_graph=use_graph)
    end.record()
    batch.wait()
    return start.elapsed_time(end) / reps


def _apply(batch, knobs):
    batch.config(knobs["lines_per_wave"], knobs["waves_per_group"], knobs["prefetch"],
                 knobs["fusion_depth"], knobs.get("sum_lag", -1))


def tune(batch, shapes, reps=3, path=None, force=False, use_graph=True):
    """Configures `batch` with the cached or freshly measured best knobs.
    Returns the knobs plus {"ms", "cached", "trials"}."""
    path = path or cache_path()
    key = signature(shapes)
    data = _load(path)
    if key in data and not force:
        best = data[key]
        _apply(batch, best)
        return dict(best, cached=True, trials=[])
    cur = dict(DEFAULT)
    seen = {}
    trials = []

    def measure(knobs):
        k = tuple(sorted(knobs.items()))
        if k not in seen:
            _apply(batch, knobs)
            seen[k] = time_batch(batch, reps, use_graph=use_graph)
            trials.append(dict(knobs, ms=seen[k]))
        return seen[k]

    best_ms = measure(cur)
    for name, values in KNOBS:
        if name == "sum_lag" and cur["fusion_depth"] == 1:
            continue  # only the persistent queues order sums
        for v in values:
            cand = dict(cur, **{name: v})
            ms = measure(cand)
            if ms < best_ms:
                best_ms, cur = ms, cand
    _apply(batch, cur)
    best = dict(cur, ms=best_ms)
    data[key] = best
    _store(path, data)
    return dict(best, cached=False, trials=trials)
