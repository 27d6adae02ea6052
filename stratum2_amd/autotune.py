"""Launch-shape autotuner for batched rendering (st_batch_*).

Coordinate descent over the four fields of st_batch_config: each field in
turn takes the candidate value with the shortest batch time, the others
held, and the search repeats until a pass changes nothing (or `passes` runs
out).  Every candidate is timed on a fresh Batch whose configuration is set
before its first render (which captures the graph); the mean of `repeats`
further renders is its time.  Images do not depend on the configuration, so
the tuner never looks at them.

The reference has no counterpart: it renders one view per frame from a
host-driven dispatch loop (TestRenderer.cpp:613-621).
"""
import time

from .renderer import Batch

CANDIDATES = {
    "wave_paths": [1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25],
    "streams": [1, 2, 4],
    "trace_blocks_per_cu": [8, 12, 16],
    "shade_blocks_per_cu": [4, 8, 16],
}


def time_config(ctx, requests, config, repeats=2):
    """Mean wall time in ms of `repeats` renders of a fresh batch with `config`."""
    b = Batch(ctx, requests)
    try:
        b.set_config(**config)
        b.render()  # capture + first launch
        a = time.perf_counter()
        for _ in range(repeats):
            b.render()  # returns when the images are complete
        return (time.perf_counter() - a) / repeats * 1e3
    finally:
        b.close()


def autotune_batch(ctx, requests, repeats=2, candidates=None, start=None, passes=2, log=None):
    """Returns (best config dict, its ms, [(config, ms) for every candidate timed])."""
    cands = dict(CANDIDATES if candidates is None else candidates)
    if start is None:
        b = Batch(ctx, requests)
        start = {k: v for k, v in b.config().items() if k != "ms"}
        b.close()
    best = dict(start)
    seen = {}

    def measure(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in seen:
            seen[key] = time_config(ctx, requests, cfg, repeats)
            if log:
                log(f"autotune {cfg}: {seen[key]:.1f} ms")
        return seen[key]

    best_ms = measure(best)
    for _ in range(passes):
        changed = False
        for field, values in cands.items():
            for v in values:
                if v == best[field]:
                    continue
                cfg = dict(best, **{field: v})
                ms = measure(cfg)
                if ms < best_ms:
                    best, best_ms, changed = cfg, ms, True
        if not changed:
            break
    history = [(dict(k), ms) for k, ms in seen.items()]
    return best, best_ms, history
