"""Oracle for the device noise sampler.

TEST INFRASTRUCTURE ONLY (imported by tests/ and nowhere in ``ksim_amd``).

* :func:`walk_rows` restates the per-shot channel walk of the device noise
  sampler (``ksim_amd/csrc/device.hpp`` ``noise_walk`` / ``noise_stream``,
  SplitMix64 by S. Vigna) with Python integers, on the tables the library
  exports, so the device's f rows can be checked bit for bit.
* :func:`exact_marginals` / :func:`exact_pair_parity` give the closed-form
  distribution the reference's ChannelSampler
  (``src/python/ksim/compile/channels.py:503-658``) samples from: channels are
  independent, channel c fires outcome o with probability ``probs[o]`` and
  flips the f bits of that outcome's pattern, so for any set S of f bits the
  parity of S is the XOR of independent channel contributions and
  ``P(parity = 1) = (1 - prod_c (1 - 2 q_c)) / 2`` with ``q_c`` the
  probability that channel c flips an odd number of bits of S.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def mix64(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def noise_stream(seed: int, shot: int) -> int:
    return mix64((seed ^ mix64((shot + 0x632BE59BD9B4E019) & M64)) & M64)


def walk(tables: dict, x: int):
    """Fired (kept channel, outcome) pairs of one shot whose stream starts at x."""
    C = tables["C"]
    skip, acc, n_out, off_cdf = tables["skip"], tables["acc"], tables["n_out"], tables["off_cdf"]
    cdf = tables["cdf"]
    fired = []
    c = 0
    while c < C:
        x = (x + GAMMA) & M64
        u = mix64(x)
        lo, hi = 0, C - c
        while lo < hi:
            mid = (lo + hi + 1) >> 1
            if u < int(skip[mid]):
                lo = mid
            else:
                hi = mid - 1
        c += lo
        if c >= C:
            break
        x = (x + GAMMA) & M64
        if mix64(x) <= int(acc[c]):
            x = (x + GAMMA) & M64
            v = mix64(x)
            o = 0
            while o + 1 < int(n_out[c]) and v > int(cdf[int(off_cdf[c]) + o]):
                o += 1
            fired.append((c, o))
        c += 1
    return fired


def walk_rows(tables: dict, seed: int, shots: int, shot_base: int = 0) -> np.ndarray:
    """(shots, num_f) uint8 f rows the device draws for these shots."""
    nf, wf = tables["num_f"], tables["wf"]
    pat, off_pat = tables["pat"], tables["off_pat"]
    out = np.zeros((shots, nf), np.uint8)
    for s in range(shots):
        row = [0] * max(wf, 1)
        for c, o in walk(tables, noise_stream(seed, shot_base + s)):
            base = int(off_pat[c]) + o * wf
            for w in range(wf):
                row[w] ^= int(pat[base + w])
        for j in range(nf):
            out[s, j] = (row[j >> 5] >> (j & 31)) & 1
    return out


def flip_probability(probs, patterns, bits) -> float:
    """Probability that one channel flips an odd number of the f bits `bits`
    (probs: its 2^k outcome probabilities; patterns: (2^k - 1, num_f))."""
    q = 0.0
    for o in range(1, len(probs)):
        if int(np.sum(patterns[o - 1][list(bits)])) & 1:
            q += float(probs[o])
    return q


def parity_probability(channel_probs, pattern_rows, counts, bits) -> float:
    prod = 1.0
    r = 0
    for pr, n in zip(channel_probs, counts):
        q = flip_probability(pr, pattern_rows[r:r + n - 1], bits)
        prod *= 1.0 - 2.0 * q
        r += n - 1
    return (1.0 - prod) / 2.0


def exact_marginals(channel_probs, pattern_rows, counts, num_f) -> np.ndarray:
    return np.array([parity_probability(channel_probs, pattern_rows, counts, [j])
                     for j in range(num_f)])


def exact_pair_parity(channel_probs, pattern_rows, counts, i, j) -> float:
    return parity_probability(channel_probs, pattern_rows, counts, [i, j])
