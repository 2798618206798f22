"""Autotuner for the batched persistent launch.

The persistent kernel has three launch-shape knobs:

* ``wave_shots`` — shots per wave group per work item.  A block takes one
  item at a time from the device-wide cursor and each of its wave groups
  runs its shots 64 at a time (one lane per shot), so larger items amortise
  the cursor atomic and keep a block on one program (its packed words stay
  hot in that CU's scalar cache) at the cost of a coarser tail;
* ``split_share`` — how finely heavy components are split over waves: a
  64-shot tile may take each of its waves at most 1/split_share of one
  wave's part of the batch; the components above that budget have their
  terms shared by 2 or 4 waves (``None``: no split).  Splitting shortens
  the tail of heavy tiles at the cost of a partial-sum exchange per graph;
* ``blocks_per_cu`` — 256-thread persistent blocks per CU.  Past the
  occupancy the kernel's registers allow, extra blocks only queue.

All depend on the mix (term counts, shot counts, parameter words), so the
tuner times every candidate on the actual batch (one untimed warm-up and the
median of ``reps`` hipGraph replays each) and keeps the fastest.  Results are
cached on disk per (device, kernel build, batch signature); ``KSIM_TUNE_CACHE``
names the cache file (default ``~/.cache/ksim_amd/batch_tune.json``).
"""
from __future__ import annotations

import hashlib
import json
import os
import time

WAVE_SHOTS = (64, 128, 256)
BLOCKS_PER_CU = (4, 5, 6, 8)
# term-split budget: a 64-shot tile may take each wave at most 1/share of
# its share of the batch (None: no split)
SPLIT_SHARES = (8, 32, None)


def cache_path() -> str:
    return os.environ.get("KSIM_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "ksim_amd",
                                       "batch_tune.json"))


def batch_signature(samplers, fs) -> str:
    """Key of a batch: per program its packed size and ops, shots and f width."""
    h = hashlib.sha256()
    for s, f in zip(samplers, fs):
        st = s.stats()
        h.update(f"{st['words']}:{st['terms']}:{st['ops_per_shot']}:"
                 f"{int(f.shape[0])}:{int(f.shape[1])};".encode())
    return h.hexdigest()[:32]


def library_key() -> str:
    """Hash of the built kernel library: tunings are kept per build, since
    a kernel change can move the best launch shape."""
    from . import _lib
    h = hashlib.sha256()
    with open(_lib.LIB_PATH, "rb") as fh:
        for block in iter(lambda: fh.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()[:16]


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_batch(batch, blocks: int, reps: int) -> float:
    """Median device time (s) of one hipGraph replay of ``batch``."""
    import torch
    batch.run(use_graph=True, blocks=blocks)  # capture + warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        batch.run(use_graph=True, blocks=blocks)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / 1e3)
    batch.check()
    times.sort()
    return times[len(times) // 2]


def split_budget(samplers, fs, cus: int, share) -> float:
    """Per-wave tile budget (algorithmic ops) for a split share: 1/share of
    one wave's part of the batch (5 blocks of 4 waves per CU); None turns
    the term split off."""
    if share is None:
        return 1e300
    total = sum(s.ops_per_shot * int(f.shape[0]) for s, f in zip(samplers, fs))
    return total / (cus * 5 * 4) / float(share)


def tune_batch(samplers, fs, seeds, reps: int = 3, use_cache: bool = True,
               wave_shots=WAVE_SHOTS, blocks_per_cu=BLOCKS_PER_CU, split_shares=SPLIT_SHARES,
               log=None) -> dict:
    """Pick (wave_shots, split share, blocks_per_cu) for this batch; returns
    ``{"wave_shots", "split_share", "tile_ops", "blocks_per_cu", "seconds",
    "trials", "cached"}``."""
    from .batch import BatchedSampler, device_info
    info = device_info()
    dev_key = f"cus{info['cus']}-clk{info['clock_khz']}-lib{library_key()}"
    key = f"{dev_key}/{batch_signature(samplers, fs)}"
    path = cache_path()
    table = _load(path) if use_cache else {}
    if key in table:
        return dict(table[key], cached=True)
    trials = []
    t0 = time.perf_counter()
    for ws in wave_shots:
        for share in split_shares:
            tops = split_budget(samplers, fs, info["cus"], share)
            b = BatchedSampler(samplers, fs, seeds, wave_shots=ws, tile_ops=tops)
            for bpc in blocks_per_cu:
                sec = time_batch(b, info["cus"] * bpc, reps)
                trials.append(dict(wave_shots=ws, split_share=share, tile_ops=tops,
                                   blocks_per_cu=bpc, seconds=sec))
                if log:
                    log(f"[tune] wave_shots={ws} split_share={share} blocks/CU={bpc}: "
                        f"{sec * 1e3:.2f} ms")
            del b
    best = min(trials, key=lambda r: r["seconds"])
    res = dict(wave_shots=best["wave_shots"], split_share=best["split_share"],
               tile_ops=best["tile_ops"], blocks_per_cu=best["blocks_per_cu"],
               seconds=best["seconds"], trials=trials,
               tune_s=time.perf_counter() - t0)
    if use_cache:
        table = _load(path)
        table[key] = res
        try:
            _store(path, table)
        except OSError:
            pass  # a read-only home only loses the cache
    return dict(res, cached=False)
