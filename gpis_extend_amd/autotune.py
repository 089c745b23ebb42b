"""Launch-shape autotuner for the persistent flight kernels.

Knobs (all exposed by the C ABI):
  * sub-cells per cell edge of the support lists (gpis_set_lists): the finer the sub-cells,
    the shorter each list a lane sums (and the larger the lists in memory);
  * waves per CU of the persistent grid (occupancy of the launch) and threads per block
    (waves sharing one block's LDS counters);
  * the finish/refill batch (how many lanes of a wave wait for each other before the
    divergent finish, bounce and refill code runs once for all of them: 1 serves each lane
    at once, 64 keeps 64-ray tiles together).
The tuner first picks the sub-cell count at the default launch shape, then times every
launch shape on the caller's own batch with HIP events on the current stream and keeps the
fastest.  Results are cached per (entry point, medium signature, batch-size bucket) for the
life of the process.
"""
from __future__ import annotations

import itertools
import math
from typing import Callable, Dict, Tuple

_CACHE: Dict[Tuple, Tuple[int, int, int, int]] = {}

SUBCELLS = (3, 4, 6, 8)
WAVES_PER_CU = (8, 12, 16, 24)
BLOCK_SIZES = (64, 128, 256)
REFILL = (8, 16, 32, 48)


def _key(dev, mode: str, n: int):
    m = dev.medium
    return (mode, m.cov_kind, bool(m.ns_multires), bool(m.single_realization), m.correlation_context,
            round(m.length_scale, 6), int(m.impulse_density), int(math.log2(max(n, 1))))


def _time(run: Callable[[], None], repeats: int) -> float:
    import torch
    stream = torch.cuda.current_stream()
    run()  # settle
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(repeats):
        run()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / repeats


def _apply(dev, choice, with_lists: bool):
    r, w, b, sub = choice
    if with_lists:
        dev.set_lists(sub, 2)
    dev.set_refill(r)
    dev.set_launch(w, b)


def tune(dev, mode: str, n: int, run: Callable[[], None], repeats: int = 1, verbose: bool = False,
         lists: bool = False, blocks=BLOCK_SIZES):
    """Pick (refill, waves_per_cu, block_size, sub-cells) for `run` (which launches once)
    and apply it to `dev`.  `lists`: the medium has an impulse table whose support lists
    are tuned too; `blocks`: the block sizes the entry point takes (the 1D kernels run
    256-thread blocks).  Returns the chosen tuple."""
    key = _key(dev, mode, n) + (lists,)
    if key in _CACHE:
        _apply(dev, _CACHE[key], lists)
        return _CACHE[key]
    best_sub = SUBCELLS[0]
    if lists:
        dev.set_refill(16)
        dev.set_launch(16, 128)
        best_ms = float("inf")
        for sub in SUBCELLS:
            dev.set_lists(sub, 2)
            ms = _time(run, repeats)
            if verbose:
                print(f"autotune {mode}: sub-cells {sub}: {ms:.3f} ms", flush=True)
            if ms < best_ms:
                best_sub, best_ms = sub, ms
        dev.set_lists(best_sub, 2)
    best, best_ms = None, float("inf")
    for r, w, b in itertools.product(REFILL, WAVES_PER_CU, blocks):
        dev.set_refill(r)
        dev.set_launch(w, b)
        ms = _time(run, repeats)
        if verbose:
            print(f"autotune {mode}: refill={r} waves/CU={w} block={b}: {ms:.3f} ms", flush=True)
        if ms < best_ms:
            best, best_ms = (r, w, b, best_sub), ms
    _CACHE[key] = best
    _apply(dev, best, lists)
    return best
