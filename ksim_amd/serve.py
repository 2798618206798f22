"""Serving: a request loop with continuous batching over the persistent kernel.

The reference has no serving path (it samples one program per call,
src/python/ksim/sample/sampler.py:57-67).  This loop keeps uploaded programs
resident on the device (a cache keyed by the caller's program key, like a
compile cache), admits requests while their HBM footprint fits the budget
and a slot is free, and runs the device work as a sequence of persistent
launches ("steps"), each over tiles of every active request:

* a request arriving while others run is admitted as soon as memory and a
  slot free up and its tiles join the next step, so its work overlaps the
  others' on the device;
* up to three steps are in flight on three streams, so the blocks of the
  next step start as those of the previous one drain (a tile of a heavy
  program can run for a large part of a second on one wave);
* a request completes when the step holding its last tile has finished and
  its output has been copied back; latency = completion - arrival.

Every shot of a request draws the per-shot stream keyed by (seed, component,
shot index), so a request's result is the same alone and in any mix.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import hashlib
import heapq
import math
import time
from collections import OrderedDict, deque

import numpy as np

from . import _lib
from .sampler import HipProgramSampler, blob_stats, program_blob


@dataclasses.dataclass
class Request:
    """One sampling request: ``prog`` (a FlatProgram, a program blob, or a
    sampler already built for it) under ``key``, its f matrix (shots, f_cols)
    uint8, a seed and an arrival time in seconds from the start of the run.
    With ``key=None`` the program is keyed by its content (the hash of its
    packed blob), so identical programs stay resident once."""
    key: object
    prog: object
    f: np.ndarray
    seed: int
    arrival: float = 0.0
    # filled by the loop
    footprint: int | None = None   # device bytes of its program, once known
    out_cols: int | None = None    # output columns of its program, once known
    out: np.ndarray | None = None
    admitted: float | None = None
    done: float | None = None

    @property
    def shots(self) -> int:
        return int(self.f.shape[0])

    @property
    def latency(self) -> float | None:
        return None if self.done is None else self.done - self.arrival


class _Active:
    def __init__(self, req, sampler, slot, d_f, d_out, tiles):
        self.req, self.sampler, self.slot = req, sampler, slot
        self.d_f, self.d_out = d_f, d_out
        self.tiles = tiles            # deque of (comp, shot0, count, split)
        self.outstanding = 0          # tiles launched, not yet retired
        self.bytes = d_f.numel() + d_out.numel()


class ServingLoop:
    """Continuous-batching request loop.

    hbm_budget      bytes of device memory the loop may hold (programs, f
                    matrices and outputs)
    max_slots       requests admitted at once
    step_ms         target duration of one persistent launch
    tile_shots      shots per wave group per work item (a multiple of 64);
                    an item of term split L covers tile_shots * 4 / L shots
    blocks          persistent blocks per launch (default: 6 per CU)
    fill_long       a step that holds long items also holds the work of
                    blocks / streams of the longest (see run)
    """

    def __init__(self, hbm_budget: int = 32 << 30, max_slots: int = 64, step_ms: float = 20.0,
                 tile_shots: int = 64, blocks: int | None = None, max_items: int = 1 << 16,
                 rate: float = 2.5e13, streams: int = 3, fill_long: bool = True):
        import torch
        from .batch import device_info
        torch.cuda.init()  # stream creation alone does not initialise torch's device state
        self.torch = torch
        self.lib = _lib.load()
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.step_s = step_ms / 1e3
        if tile_shots <= 0 or tile_shots % 64:
            raise ValueError("tile_shots must be a positive multiple of 64")
        self.tile_shots = int(tile_shots)
        self.max_items = int(max_items)
        self.blocks = int(blocks or device_info()["cus"] * 6)
        self.min_items = min(self.blocks, max_items)  # one item per persistent block
        h = C.c_void_p()
        _lib.check(self.lib.ksim_sched_create(self.max_slots, self.max_items, C.byref(h)))
        self._h = h
        if not 1 <= streams <= 4:
            raise ValueError("streams must be in 1..4 (the scheduler has four staging sets)")
        # default-priority streams: steps rotate over them; outputs return on
        # a copy stream
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.copy_stream = torch.cuda.Stream()
        self.stream = self.streams[0]
        self.programs: OrderedDict = OrderedDict()  # key -> [sampler, bytes, users]
        self.program_bytes = 0
        # algorithmic int64 ops per second of the whole device (sets the work
        # put into one step)
        self.rate = float(rate)
        self.fill_long = bool(fill_long)

    def _split(self, sampler, comp: int) -> int:
        """Term split of a request's tiles: a 64-shot tile should take each
        wave at most half of its share of one step's work."""
        target = self.rate * self.step_s / (self.blocks * 4) / 2
        L = C.c_int32(1)
        _lib.check(self.lib.ksim_program_split_hint(sampler.handle, int(comp), float(target),
                                                    C.byref(L), None))
        return int(L.value)

    # ------------------------------------------------------------ programs
    @staticmethod
    def content_key(blob: bytes) -> str:
        """Residency key of a program blob: programs with the same packed
        words share one resident copy, whoever submits them."""
        return "blob:" + hashlib.sha256(blob).hexdigest()[:32]

    def _resolve(self, r):
        """Give a request without a key its content key (packing a
        FlatProgram into a blob on the host once)."""
        if r.key is None:
            if not isinstance(r.prog, (bytes, bytearray, HipProgramSampler)):
                r.prog = program_blob(r.prog)
            if isinstance(r.prog, HipProgramSampler):
                r.prog = r.prog.to_bytes()
            r.key = self.content_key(bytes(r.prog))
        return r.key

    def _footprint(self, r) -> int:
        """Device bytes the program of request ``r`` takes once resident
        (its packed words), known before any upload: read from the blob
        header, from an existing sampler, or by packing a FlatProgram into a
        blob on the host here (the upload then starts from that blob).  Also
        records the program's output columns (``r.out_cols``)."""
        if r.footprint is None:  # once per request, however long it waits
            if isinstance(r.prog, HipProgramSampler):
                r.footprint = int(self.lib.ksim_program_device_bytes(r.prog.handle))
                r.out_cols = r.prog.num_outputs
            else:
                if not isinstance(r.prog, (bytes, bytearray)):
                    r.prog = program_blob(r.prog)
                st = blob_stats(bytes(r.prog))
                r.footprint = 4 * int(st["words"])
                r.out_cols = st["num_outputs"]
        return r.footprint

    def _need(self, r, ent) -> int:
        """Device bytes admitting ``r`` takes: its f matrix, its output rows
        (max(num_outputs, 1) bytes per shot, as allocated) and, for a
        program not resident yet, the program's words."""
        if ent is not None:
            out_cols = ent[0].num_outputs
            prog = 0
        else:
            prog = self._footprint(r)
            out_cols = r.out_cols
        return r.shots * (int(r.f.shape[1]) + max(int(out_cols), 1)) + prog

    def _program(self, key, prog):
        ent = self.programs.get(key)
        if ent is not None:
            self.programs.move_to_end(key)
            return ent
        if isinstance(prog, HipProgramSampler):
            sampler = prog
        elif isinstance(prog, (bytes, bytearray)):
            sampler = HipProgramSampler.from_bytes(bytes(prog))
        else:
            sampler = HipProgramSampler(prog)
        nbytes = int(self.lib.ksim_program_device_bytes(sampler.handle))
        ent = [sampler, nbytes, 0]
        self.programs[key] = ent
        self.program_bytes += nbytes
        return ent

    def _evict_for(self, need: int, used: int) -> int:
        """Drop idle programs (least recently used first) until `need` fits."""
        for key in list(self.programs.keys()):
            if used + need <= self.hbm_budget:
                break
            ent = self.programs[key]
            if ent[2] == 0:
                used -= ent[1]
                self.program_bytes -= ent[1]
                del self.programs[key]
        return used

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self.lib.ksim_sched_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- run
    def run(self, requests: list[Request], realtime: bool = True, progress: bool = False,
            max_wall_s: float | None = None) -> dict:
        """Serve ``requests`` (arrival times honoured when ``realtime``; all
        available at once otherwise).  Fills each request's ``out``,
        ``admitted`` and ``done`` and returns summary statistics.  With
        ``progress`` a status line goes to stderr every two seconds;
        ``max_wall_s`` bounds the run (RuntimeError with the loop state)."""
        import sys
        torch = self.torch
        lib = self.lib
        st = self.stream
        pending = sorted(requests, key=lambda r: r.arrival)
        pend_i = 0
        active: list[_Active] = []
        free_slots = list(range(self.max_slots))[::-1]
        inflight: deque = deque()      # (event, t_launch, est_ops, [(active, tiles)])
        copies: deque = deque()        # (event, active, host_out)
        used = self.program_bytes
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        done = 0
        busy_ops = 0.0
        n_steps = 0
        last_report = 0.0
        admitted_now = False
        with torch.cuda.stream(st):
            while done < len(requests):
                t = now()
                if progress and t - last_report > 2.0:
                    last_report = t
                    print(f"[serve] t={t:.1f}s done={done}/{len(requests)} admitted={pend_i} "
                          f"active={len(active)} inflight={len(inflight)} copies={len(copies)} "
                          f"steps={n_steps}", file=sys.stderr, flush=True)
                if max_wall_s is not None and t > max_wall_s:
                    raise RuntimeError(
                        f"serving loop exceeded {max_wall_s}s: done={done}/{len(requests)} "
                        f"active={len(active)} inflight={len(inflight)} steps={n_steps} "
                        f"tiles_left={sum(len(a.tiles) for a in active)}")
                # ---- admit released requests while memory and slots allow
                while pend_i < len(pending) and (not realtime or pending[pend_i].arrival <= t) \
                        and free_slots:
                    r = pending[pend_i]
                    ent = self.programs.get(self._resolve(r))
                    need = self._need(r, ent)
                    # requests whose output is still being copied back hold
                    # their buffers and program too
                    used = (self.program_bytes + sum(a.bytes for a in active)
                            + sum(c[1].bytes for c in copies))
                    if used + need > self.hbm_budget:
                        used = self._evict_for(need, used)
                        if used + need > self.hbm_budget and (active or copies):
                            break  # wait for memory to free up
                    ent = self._program(r.key, r.prog)
                    ent[2] += 1
                    sampler = ent[0]
                    d_f = torch.from_numpy(np.ascontiguousarray(r.f, dtype=np.uint8)).to(
                        "cuda", non_blocking=False)
                    d_out = torch.empty((r.shots, max(sampler.num_outputs, 1)),
                                        dtype=torch.uint8, device="cuda")
                    d_out.zero_()
                    slot = free_slots.pop()
                    admitted_now = True
                    _lib.check(lib.ksim_sched_set_slot(
                        self._h, slot, sampler.handle, d_f.data_ptr() if d_f.numel() else None,
                        int(r.f.shape[1]), r.shots, int(r.seed), d_out.data_ptr()))
                    tiles = deque()
                    for c in range(sampler.n_comp):
                        L = self._split(sampler, c)
                        span = self.tile_shots * (4 // L)
                        tiles.extend((c, s0, min(span, r.shots - s0), L)
                                     for s0 in range(0, r.shots, span))
                    a = _Active(r, sampler, slot, d_f, d_out, tiles)
                    r.admitted = now()
                    if not tiles:  # zero shots
                        r.out = np.zeros((0, sampler.num_outputs), np.uint8)
                        r.done = now()
                        done += 1
                        ent[2] -= 1
                        free_slots.append(slot)
                        _lib.check(lib.ksim_sched_clear_slot(self._h, slot))
                    else:
                        active.append(a)
                    pend_i += 1
                # ---- order the admissions (f upload, output zeroing on the first
                # stream) before any step that may read them on another stream
                if admitted_now:
                    adm = torch.cuda.Event()
                    adm.record(st)
                    for other in self.streams[1:]:
                        other.wait_event(adm)
                    admitted_now = False
                # ---- launch a step if a stream is free
                if len(inflight) < len(self.streams) and any(a.tiles for a in active):
                    budget = self.rate * self.step_s
                    items, taken, est, big = [], [], 0.0, 0.0
                    # round-robin over active requests, one item each per pass,
                    # until the step holds the estimated work of step_ms, at
                    # least one item per block of the persistent grid (an item
                    # is one block's work: fewer items leave SIMDs idle) and,
                    # when it holds long items, the work of a 1/streams share
                    # of the grid's blocks each running one of the longest:
                    # the steps in flight then keep every block busy while
                    # long items run (a step of a few long items among short
                    # ones holds the stream until they end, with most of its
                    # blocks idle)
                    share = self.blocks / len(self.streams) if self.fill_long else 0.0
                    def short():
                        return (est < max(budget, big * share)
                                or len(items) < self.min_items)
                    while short() and len(items) < self.max_items:
                        progressed = False
                        for a in active:
                            if not a.tiles:
                                continue
                            c, s0, cnt, L = a.tiles.popleft()
                            items.append((a.slot, c, s0, cnt, L))
                            taken.append(a)
                            a.outstanding += 1
                            w = a.sampler.ops_per_shot * cnt / max(a.sampler.n_comp, 1)
                            est += w
                            big = max(big, w)
                            progressed = True
                            if not short() or len(items) >= self.max_items:
                                break
                        if not progressed:
                            break
                    arr = np.asarray(items, dtype=np.int32).reshape(-1, 5)
                    sst = self.streams[n_steps % len(self.streams)]
                    _lib.check(lib.ksim_sched_launch(self._h, _lib.ptr(arr, C.c_int32), len(items),
                                                     self.blocks, sst.cuda_stream))
                    ev = torch.cuda.Event()
                    ev.record(sst)
                    inflight.append((ev, time.perf_counter(), est, taken))
                    n_steps += 1
                    continue
                # ---- retire finished steps (streams may finish out of order)
                progressed = False
                for step in [x for x in inflight if x[0].query()]:
                    inflight.remove(step)
                    ev, tl, est, taken = step
                    busy_ops += est
                    for a in taken:
                        a.outstanding -= 1
                        if not a.tiles and a.outstanding == 0:
                            host = torch.empty(tuple(a.d_out.shape), dtype=torch.uint8,
                                               pin_memory=True)
                            with torch.cuda.stream(self.copy_stream):
                                host.copy_(a.d_out, non_blocking=True)
                                cev = torch.cuda.Event()
                                cev.record(self.copy_stream)
                            copies.append((cev, a, host))
                            active.remove(a)
                    progressed = True
                while copies and copies[0][0].query():
                    cev, a, host = copies.popleft()
                    r = a.req
                    r.out = host.numpy()[:, :a.sampler.num_outputs].copy()
                    r.done = now()
                    done += 1
                    self.programs[r.key][2] -= 1
                    _lib.check(lib.ksim_sched_clear_slot(self._h, a.slot))
                    free_slots.append(a.slot)
                    progressed = True
                if progressed:
                    continue
                # ---- nothing to do right now: wait for the GPU or the next arrival
                if inflight:
                    time.sleep(0.0002)
                elif copies:
                    copies[0][0].synchronize()
                elif pend_i < len(pending):
                    dt = pending[pend_i].arrival - now()
                    if dt > 0:
                        time.sleep(min(dt, 0.01))
        wall = now()
        _lib.check(lib.ksim_sched_check(self._h, st.cuda_stream))
        torch.cuda.synchronize()
        lat = np.array([r.latency for r in requests], dtype=np.float64)
        shots = sum(r.shots for r in requests)
        return dict(requests=len(requests), shots=shots, wall_s=wall,
                    shots_per_s=shots / wall if wall > 0 else float("nan"),
                    p50_ms=float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
                    p99_ms=float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
                    mean_ms=float(lat.mean() * 1e3) if len(lat) else None,
                    steps=n_steps, programs_resident=len(self.programs))


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times (s) of n requests of a Poisson process with ``rate``/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
