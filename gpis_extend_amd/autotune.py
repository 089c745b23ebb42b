"""Launch-shape autotuner for the persistent flight kernels.

Knobs (all exposed by the C ABI): waves per CU of the persistent grid
(occupancy of the launch), threads per block (waves sharing one block's LDS
counters), and the finish/refill batch (how many lanes of a wave wait for each
other before the divergent finish, bounce and refill code runs once for all of
them: 1 serves each lane at once, 64 keeps 64-ray tiles together).  The tuner times every combination on
the caller's own batch with HIP events on the current stream and keeps the
fastest; results are cached per (entry point, medium signature, batch-size
bucket) for the life of the process.
"""
from __future__ import annotations

import itertools
import math
from typing import Callable, Dict, Tuple

_CACHE: Dict[Tuple, Tuple[int, int, int]] = {}

WAVES_PER_CU = (8, 12, 16, 20, 24)
BLOCK_SIZES = (64, 128, 256)
REFILL = (1, 8, 16, 32, 64)


def _key(dev, mode: str, n: int):
    m = dev.medium
    return (mode, m.cov_kind, bool(m.ns_multires), bool(m.single_realization), m.correlation_context,
            round(m.length_scale, 6), int(m.impulse_density), int(math.log2(max(n, 1))))


def tune(dev, mode: str, n: int, run: Callable[[], None], repeats: int = 1, verbose: bool = False):
    """Pick (refill, waves_per_cu, block_size) for `run` (which launches once) and
    apply it to `dev`.  Returns the chosen triple."""
    import torch
    key = _key(dev, mode, n)
    if key in _CACHE:
        r, w, b = _CACHE[key]
        dev.set_refill(r)
        dev.set_launch(w, b)
        return _CACHE[key]
    best, best_ms = None, float("inf")
    stream = torch.cuda.current_stream()
    for r, w, b in itertools.product(REFILL, WAVES_PER_CU, BLOCK_SIZES):
        dev.set_refill(r)
        dev.set_launch(w, b)
        run()  # settle
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(repeats):
            run()
        e.record(stream)
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / repeats
        if verbose:
            print(f"autotune {mode}: refill={r} waves/CU={w} block={b}: {ms:.3f} ms", flush=True)
        if ms < best_ms:
            best, best_ms = (r, w, b), ms
    _CACHE[key] = best
    dev.set_refill(best[0])
    dev.set_launch(best[1], best[2])
    return best
