"""Oracle checks of seeded sampling shared by the GPU tests: the shots at
given indices of a seeded run against the exact oracle fed the reference's
per-shot xorshift64 streams (component k keyed by seed + 0x9E37 k,
src/python/ksim/sample/sampler.py:63, src/cpp/ksim/kokkos_sim.cpp:215-217)."""
import numpy as np

from oracle import exact

SEED_STRIDE = 0x9E37


def seeded_columns(prog, seed, shots):
    """(len(shots), total outputs) uniforms of the given shot indices."""
    return np.concatenate([exact.kokkos_uniforms_at(seed + SEED_STRIDE * k, shots,
                                                    len(c.graphs) - 1)
                           for k, c in enumerate(prog.components)], axis=1)


def oracle_rows(prog, f, seed, idx, fast=False):
    """The oracle's rows for shots ``idx`` of a seeded run over f rows ``f``."""
    idx = np.asarray(idx, dtype=np.int64)
    U = seeded_columns(prog, seed, idx)
    if fast:
        from oracle import exact_fast
        return exact_fast.sample_program(prog, f[idx], U)
    return exact.sample_program(prog, f[idx], U)


def pick(B, n, seed):
    """Up to n shot indices of a run of B shots, the first and last among them."""
    if B <= n:
        return np.arange(B)
    rng = np.random.default_rng(seed)
    mid = rng.choice(np.arange(1, B - 1), size=n - 2, replace=False) if n > 2 else []
    return np.sort(np.concatenate([[0, B - 1], mid]).astype(np.int64))
