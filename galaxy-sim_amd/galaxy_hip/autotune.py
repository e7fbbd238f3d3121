"""Autotuner for pooled stepping: walk workgroup size, leaf bucket and sort
digit width, chosen by timing full steps of the actual workload.

The search is exhaustive over a small space.  For each (walk_wg, radix_bits)
one pool is built and loaded with the workload; every bucket is then timed on
the same pool (gs_slot_set_tree changes the leaf rule in place, and every
trial starts from a re-upload of the initial particles, so all trials step the
same states).  A trial is `warmup` untimed steps then `steps` timed ones
(sort every step, graph replay), wall-clock between stream synchronisations.

The bucket changes the tree and therefore the force approximation (a larger
leaf turns more monopoles into exact pairs); the tuner only ranks speed, and
the chosen bucket is reported beside the results it produced.
"""
import hashlib
import json
import os
import time

import numpy as np

from . import GalaxyError, GalaxyPool, lib

DEFAULT_SPACE = {
    "walk_wg": (64, 128, 256, 512),
    "bucket": (8, 16, 24, 32, 48),
    "radix_bits": (8, 10),
}


def _supported_radix():
    """Digit widths the loaded library accepts (probed once per process)."""
    if not hasattr(_supported_radix, "cache"):
        ok = []
        for r in DEFAULT_SPACE["radix_bits"]:
            try:
                GalaxyPool([1024], radix_bits=r).destroy()
                ok.append(r)
            except GalaxyError:
                pass
        _supported_radix.cache = tuple(ok)
    return _supported_radix.cache


def workload_key(sizes, params, space):
    h = hashlib.sha256()
    h.update(np.asarray(sorted(int(n) for n in sizes), np.int64).tobytes())
    h.update(json.dumps(params, sort_keys=True).encode())
    h.update(json.dumps({k: list(v) for k, v in space.items()}, sort_keys=True).encode())
    h.update(lib().gs_version())
    return h.hexdigest()[:24]


def _load_pool(pool, galaxies, params, bucket):
    for k, g in enumerate(galaxies):
        pool.upload(k, g, **params)
        pool.set_tree(k, bucket)


def time_pool(pool, galaxies, params, bucket, steps, warmup):
    _load_pool(pool, galaxies, params, bucket)
    pool.step(warmup, sort_every=1)
    pool.sync()
    t0 = time.perf_counter()
    pool.step(steps, sort_every=1)
    pool.sync()
    return (time.perf_counter() - t0) / steps * 1e3


def tune(galaxies, params, space=None, steps=3, warmup=1, cache=None, log=None):
    """Pick (walk_wg, bucket, radix_bits) for stepping `galaxies` (a list of
    n x 8 record arrays) together with updateParams `params`.

    Returns {"walk_wg", "bucket", "radix_bits", "ms_per_step", "trials",
    "key"}; with `cache` (a JSON file path) a previous result for the same
    workload, space and library is reused."""
    space = dict(DEFAULT_SPACE if space is None else space)
    space["radix_bits"] = tuple(r for r in space["radix_bits"] if r in _supported_radix())
    sizes = [g.shape[0] for g in galaxies]
    key = workload_key(sizes, params, space)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            return stored[key]
    trials = []
    for radix in space["radix_bits"]:
        for wg in space["walk_wg"]:
            pool = GalaxyPool(sizes, walk_wg=wg, radix_bits=radix)
            try:
                for bucket in space["bucket"]:
                    ms = time_pool(pool, galaxies, params, bucket, steps, warmup)
                    trials.append({"walk_wg": wg, "bucket": bucket, "radix_bits": radix,
                                   "ms_per_step": ms})
                    if log:
                        log(f"autotune wg={wg} bucket={bucket} radix={radix}: {ms:.3f} ms/step")
            finally:
                pool.destroy()
    best = min(trials, key=lambda t: t["ms_per_step"])
    out = dict(best, trials=trials, key=key)
    if cache:
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[key] = out
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(stored, f, indent=1)
    return out
