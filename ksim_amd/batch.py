"""Batched sampling of many FlatPrograms in one persistent kernel.

The reference samples one component per kernel launch
(src/python/ksim/sample/sampler.py:60-64, src/cpp/ksim/kokkos_sim.cpp:214);
with many small programs the launches cannot fill the GPU.  Here every
(program, component, shot range) of the batch is one work item of one
256-thread block; items are ordered heaviest first and a fixed grid of
persistent blocks pulls them from a device-side cursor until the list is
drained.  The components of heavy programs get a term split: 2 or 4 waves
share each graph's terms on one 64-shot tile, so no single tile is long
enough to leave the rest of the grid waiting at the end.  The launch (cursor reset +
persistent kernel) can be captured once into a hipGraph and replayed.

Each program keeps its own resident f matrix, output buffer and seed, and
every shot draws the same per-shot stream it would in a plain launch, so a
program's result is the same alone and in any batch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .sampler import HipProgramSampler


def device_info() -> dict:
    info = np.zeros(4, np.int64)
    _lib.check(_lib.load().ksim_device_info(_lib.ptr(info, C.c_int64)))
    return dict(cus=int(info[0]), lds_per_cu=int(info[1]), clock_khz=int(info[2]),
                hbm_bytes=int(info[3]))


class BatchedSampler:
    """One persistent launch over several programs with resident inputs.

    ``samplers``: HipProgramSampler per program; ``f``: torch uint8 CUDA
    tensors (B_k, f_cols_k); ``seeds``: per-program seeds.  Outputs are
    allocated here (``self.out[k]``, (B_k, num_outputs_k) uint8).
    """

    def __init__(self, samplers: list[HipProgramSampler], f: list, seeds: list[int],
                 wave_shots: int = 128, blocks_per_cu: int = 6, tile_ops: float = 0.0):
        import torch
        from .sampler import _check_seed, device_u8_2d
        torch.cuda.init()
        if not len(samplers) == len(f) == len(seeds):
            raise ValueError("samplers, f and seeds must have the same length")
        for k, (s, fk) in enumerate(zip(samplers, f)):
            device_u8_2d(fk, f"f[{k}]")
            _check_seed(seeds[k], s.n_comp)
        self.torch = torch
        self.samplers = samplers
        self.f = f
        self.seeds = [int(x) for x in seeds]
        self.out = [torch.zeros((int(fk.shape[0]), max(s.num_outputs, 1)), dtype=torch.uint8,
                                device=fk.device) for s, fk in zip(samplers, f)]
        n = len(samplers)
        progs = (C.c_void_p * n)(*[s.handle for s in samplers])
        dfs = (C.c_void_p * n)(*[fk.data_ptr() if fk.numel() else None for fk in f])
        fcols = np.asarray([int(fk.shape[1]) for fk in f], np.int32)
        Bs = np.asarray([int(fk.shape[0]) for fk in f], np.int64)
        sd = (C.c_uint64 * n)(*[int(x) for x in seeds])
        douts = (C.c_void_p * n)(*[o.data_ptr() for o in self.out])
        h = C.c_void_p()
        _lib.check(_lib.load().ksim_batch_create(progs, n, dfs, _lib.ptr(fcols, C.c_int32),
                                                 _lib.ptr(Bs, C.c_int64), sd, douts,
                                                 int(wave_shots), float(tile_ops), C.byref(h)))
        self._h = h
        self.total_shots = int(Bs.sum())
        self.ops = float(sum(s.ops_per_shot * int(b) for s, b in zip(samplers, Bs)))
        # unavoidable HBM bytes: f rows in, output rows out
        self.bytes = float(sum((int(fk.shape[1]) + s.num_outputs) * int(b)
                               for s, fk, b in zip(samplers, f, Bs)))
        info = device_info()
        # persistent grid: enough 256-thread blocks for every CU to hold
        # as many as registers allow (5 waves per SIMD, 5 blocks per CU)
        self.wave_shots = int(wave_shots)
        self.tile_ops = float(tile_ops)
        self.blocks = info["cus"] * int(blocks_per_cu)

    def run(self, use_graph: bool = True, blocks: int | None = None, stream=None):
        st = stream if stream is not None else self.torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().ksim_batch_run(self._h, int(blocks or self.blocks),
                                              1 if use_graph else 0, st))

    def check(self):
        _lib.check(_lib.load().ksim_batch_check(self._h))

    @classmethod
    def tuned(cls, samplers, f, seeds, **kw):
        """A batch with the launch shape the autotuner picks for this mix
        (ksim_amd.autotune; cached per device and batch signature)."""
        from .autotune import tune_batch
        t = tune_batch(samplers, f, seeds, **kw)
        b = cls(samplers, f, seeds, wave_shots=t["wave_shots"], blocks_per_cu=t["blocks_per_cu"],
                tile_ops=t.get("tile_ops", 0.0))
        b.tune = t
        return b

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.load().ksim_batch_destroy(h)
            except Exception:
                pass


STATIC_MAX_PROGS = 256      # KSIM_STATIC_MAX_PROGS
STATIC_MAX_ITEMS = 65536    # KSIM_STATIC_MAX_ITEMS


def sample_static(samplers: list[HipProgramSampler], f: list, seeds: list[int], out: list,
                  wave_shots: int = 128, tile_ops: float = 0.0, blocks: int | None = None,
                  stream=None) -> None:
    """One batched launch that allocates no device memory
    (``ksim_batch_sample_static``): the caller's outputs ``out[k]``
    ((B_k, max(num_outputs_k, 1)) uint8 CUDA tensors) receive what a
    :class:`BatchedSampler` with the same arguments writes.  Blocks until
    done.  At most 256 programs and 65536 work items per call; calls that
    overlap on one GPU race on its fixed work-list array."""
    import torch
    from .sampler import _check_seed, device_u8_2d
    if not len(samplers) == len(f) == len(seeds) == len(out):
        raise ValueError("samplers, f, seeds and out must have the same length")
    for k, (s, fk, ok) in enumerate(zip(samplers, f, out)):
        device_u8_2d(fk, f"f[{k}]")
        device_u8_2d(ok, f"out[{k}]", (max(s.num_outputs, 1),))
        if ok.shape[0] != fk.shape[0]:
            raise ValueError(f"out[{k}] has {ok.shape[0]} rows, f[{k}] {fk.shape[0]}")
        _check_seed(seeds[k], s.n_comp)
    n = len(samplers)
    progs = (C.c_void_p * max(n, 1))(*[s.handle for s in samplers])
    dfs = (C.c_void_p * max(n, 1))(*[fk.data_ptr() if fk.numel() else None for fk in f])
    fcols = np.asarray([int(fk.shape[1]) for fk in f] or [0], np.int32)
    Bs = np.asarray([int(fk.shape[0]) for fk in f] or [0], np.int64)
    sd = (C.c_uint64 * max(n, 1))(*[int(x) for x in seeds])
    douts = (C.c_void_p * max(n, 1))(*[o.data_ptr() if o.numel() else None for o in out])
    nb = int(blocks or device_info()["cus"] * 6)
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().ksim_batch_sample_static(progs, n, dfs, _lib.ptr(fcols, C.c_int32),
                                                    _lib.ptr(Bs, C.c_int64), sd, douts,
                                                    int(wave_shots), float(tile_ops), nb, st))
