"""Device-side noise: the f bits drawn on the GPU.

:class:`NoiseModel` replaces ``ChannelSampler``
(``src/python/ksim/compile/channels.py:503-658``) with the same constructor
arguments — a list of channel probability arrays (channel i has 2^k_i
outcomes and produces e-bits ``[sum k_<i, sum k_<=i)``) and the
``error_transform`` matrix (num_f, num_e) over GF(2) — and the same
distribution of f rows: each channel fires with probability
``1 - probs[0]`` and picks outcome o with probability ``probs[o] / p_fire``,
flipping the f bits ``bits(o) @ error_transform[:, cols].T mod 2``
(``channels.py:601-620``).  The draws are made per shot on the device from
a SplitMix64 stream keyed by (seed, shot index), so

* ``sample(n)`` returns the (n, num_f) uint8 rows like ``ChannelSampler.sample``
  (successive calls continue the shot index, so they differ, like the
  reference's advancing generator);
* :meth:`ksim_amd.sampler.HipProgramSampler.sample_noisy` draws the same rows
  inside the sampling kernel: no f matrix in HBM and none on the host.

The reference's channel simplification passes (null bits, folding, merging,
absorbing) only speed its host sampler up and leave the f distribution
unchanged; here channels whose patterns flip nothing, or that cannot fire,
are dropped and the rest walked with geometric skips on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def channel_patterns(channel_probs, error_transform) -> tuple:
    """(probs concatenated, outcome counts, pattern rows) for ksim_noise_create:
    row (c, o) for outcome o >= 1 of channel c holds the f bits it flips."""
    T = np.asarray(error_transform, dtype=np.uint8) & 1
    if T.ndim != 2:
        T = T.reshape(0, 0) if T.size == 0 else T.reshape(T.shape[0], -1)
    num_f, num_e = T.shape
    probs, counts, rows = [], [], []
    e_off = 0
    for pr in channel_probs:
        pr = np.asarray(pr, dtype=np.float64).reshape(-1)
        n = pr.shape[0]
        k = int(round(np.log2(n))) if n > 0 else -1
        if n < 1 or (1 << k) != n:
            raise ValueError("channel probability arrays must have 2^k entries")
        if e_off + k > num_e:
            raise ValueError("channels produce more e bits than error_transform has columns")
        cols = T[:, e_off:e_off + k]  # (num_f, k)
        o = np.arange(1, n)
        bits = ((o[:, None] >> np.arange(k)) & 1).astype(np.uint8)  # (n-1, k)
        pat = (bits.astype(np.int64) @ cols.T.astype(np.int64)) % 2 if k else \
            np.zeros((0, num_f), np.int64)
        probs.append(pr)
        counts.append(n)
        rows.append(pat.astype(np.uint8).reshape(n - 1, num_f))
        e_off += k
    P = np.concatenate(probs) if probs else np.zeros(0, np.float64)
    R = np.concatenate(rows, axis=0) if rows else np.zeros((0, num_f), np.uint8)
    return (np.ascontiguousarray(P), np.asarray(counts, np.int32),
            np.ascontiguousarray(R, dtype=np.uint8), num_f)


class NoiseModel:
    """ChannelSampler on the device (see the module docstring)."""

    def __init__(self, channel_probs, error_transform, seed: int | None = None):
        lib = _lib.load()
        P, counts, rows, num_f = channel_patterns(channel_probs, error_transform)
        self.num_f = int(num_f)
        self.seed = int(seed if seed is not None else np.random.default_rng().integers(0, 2**30))
        self._next_shot = 0
        h = C.c_void_p()
        _lib.check(lib.ksim_noise_create(
            _lib.ptr(P, C.c_double) if P.size else None,
            _lib.ptr(counts, C.c_int32) if counts.size else None,
            _lib.ptr(rows, C.c_uint8) if rows.size else None,
            int(counts.size), self.num_f, C.byref(h)))
        self._h = h
        info = np.zeros(4, np.int64)
        _lib.check(lib.ksim_noise_info(h, _lib.ptr(info, C.c_int64)))
        self.n_channels, _, self.wf, self.n_cdf = (int(x) for x in info)

    @property
    def handle(self):
        return self._h

    def sample(self, num_samples: int = 1, shot_base: int | None = None) -> np.ndarray:
        """(num_samples, num_f) uint8 f rows; without ``shot_base`` the shot
        index continues from the previous call."""
        B = int(num_samples)
        if shot_base is None:
            shot_base = self._next_shot
            self._next_shot += B
        out = np.zeros((B, self.num_f), np.uint8)
        if out.size:
            _lib.check(_lib.load().ksim_noise_sample(self._h, B, int(shot_base), self.seed,
                                                     _lib.ptr(out, C.c_uint8)))
        return out

    def sample_device(self, d_out, shot_base: int = 0, stream=None):
        """Fill a resident (B, f_cols) uint8 CUDA tensor with f rows."""
        import torch
        from .sampler import device_u8_2d
        device_u8_2d(d_out, "d_out")
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().ksim_noise_sample_device(
            self._h, int(d_out.shape[0]), int(shot_base), self.seed,
            d_out.data_ptr() if d_out.numel() else None, int(d_out.shape[1]), st))

    def tables(self) -> dict:
        """The device sampling tables (for checking against the oracle)."""
        C_, cd = self.n_channels, self.n_cdf
        t = dict(skip=np.zeros(C_ + 1, np.uint64), acc=np.zeros(C_, np.uint64),
                 n_out=np.zeros(C_, np.uint32), off_cdf=np.zeros(C_, np.uint32),
                 off_pat=np.zeros(C_, np.uint32), cdf=np.zeros(cd, np.uint64),
                 pat=np.zeros(cd * self.wf, np.uint32), source=np.zeros(C_, np.int32))
        p = {k: (_lib.ptr(v, C.c_uint64) if v.dtype == np.uint64 else
                 _lib.ptr(v, C.c_uint32) if v.dtype == np.uint32 else _lib.ptr(v, C.c_int32))
             if v.size else None for k, v in t.items()}
        _lib.check(_lib.load().ksim_noise_tables(self._h, p["skip"], p["acc"], p["n_out"],
                                                  p["off_cdf"], p["off_pat"], p["cdf"],
                                                  p["pat"], p["source"]))
        t.update(C=C_, wf=self.wf, num_f=self.num_f)
        return t

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.load().ksim_noise_destroy(h)
            except Exception:
                pass


def sample_circuit_program(flat, channel_probs, error_transform, shots: int, seed: int = 42):
    """The sampling half of ``ksim.sample_circuit`` (src/python/ksim/run.py:64-77)
    on the GPU: noise and sampling in one kernel, returning (det, obs)."""
    from .sampler import HipProgramSampler, split_detectors
    s = HipProgramSampler(flat)
    if channel_probs:
        out = s.sample_noisy(NoiseModel(channel_probs, error_transform, seed=seed), shots, seed)
    else:
        n_f = error_transform.shape[0] if np.asarray(error_transform).size else 0
        out = s.sample(np.zeros((shots, max(n_f, s.min_num_f)), np.uint8), seed)
    return split_detectors(out, flat.num_detectors)
