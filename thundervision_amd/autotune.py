"""Autotuner of the persistent batch launch.

Sweeps the batch knobs (waves per workgroup, prefetch depth of the SGM line
scans) on the batch itself, keeps the fastest and caches it per batch
signature (the multiset of pair shapes) in a JSON file, so a service tunes a
given mix of resolutions once.
"""
import hashlib
import json
import os

import torch

WAVES_PER_GROUP = (1, 2, 4)
PREFETCH = (4, 8)


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


def tune(batch, shapes, reps=3, path=None, force=False):
    """Configures `batch` with the cached or freshly measured best knobs.
    Returns {"waves_per_group", "prefetch", "ms", "cached", "trials"}."""
    path = path or cache_path()
    key = signature(shapes)
    data = _load(path)
    if key in data and not force:
        best = data[key]
        batch.config(0, best["waves_per_group"], best["prefetch"])
        return dict(best, cached=True, trials=[])
    trials = []
    for wpg in WAVES_PER_GROUP:
        for pf in PREFETCH:
            batch.config(0, wpg, pf)
            ms = time_batch(batch, reps)
            trials.append({"waves_per_group": wpg, "prefetch": pf, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    batch.config(0, best["waves_per_group"], best["prefetch"])
    data[key] = best
    _store(path, data)
    return dict(best, cached=False, trials=trials)
