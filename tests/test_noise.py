"""Host side of the device noise sampler: its tables against the closed
forms, and the oracle's restated walk against the exact f distribution of the
reference's ChannelSampler (channels.py:503-658) — marginals and pair
parities at 5 sigma."""
import math

import numpy as np
import pytest

from ksim_amd.noise import NoiseModel, channel_patterns
from oracle import noise_ref


def model_a():
    # a 1-bit, a 2-bit (Pauli-like) and a 4-bit channel, one that cannot fire
    # and one whose outcomes flip nothing (both dropped)
    probs = [np.array([0.97, 0.03]),
             np.array([0.85, 0.05, 0.04, 0.06]),
             np.full(16, 0.2 / 15),
             np.array([1.0, 0.0]),
             np.array([0.5, 0.5])]
    probs[2][0] = 0.8
    T = np.zeros((6, 1 + 2 + 4 + 1 + 1), np.uint8)
    T[0, 0] = 1
    T[[1, 2], 1] = 1
    T[2, 2] = T[3, 2] = 1
    T[[3, 4], 3] = 1
    T[4, 4] = 1
    T[5, 5] = T[0, 6] = 1
    T[2, 7] = 1            # channel 3 (never fires)
    # column 8 (channel 4) stays zero: it flips nothing
    return probs, T


def test_patterns_follow_the_transform():
    probs, T = model_a()
    P, counts, rows, nf = channel_patterns(probs, T)
    assert nf == 6 and list(counts) == [2, 4, 16, 2, 2]
    # channel 1 outcome 3 = e1 ^ e2
    assert np.array_equal(rows[1 + 2], (T[:, 1] ^ T[:, 2]))


def test_tables_are_the_closed_forms():
    probs, T = model_a()
    nm = NoiseModel(probs, T, seed=1)
    t = nm.tables()
    assert t["C"] == 3 and list(t["source"]) == [0, 1, 2]
    p_fire = np.array([1.0 - probs[c][0] for c in range(3)])  # as channels.py:605
    p_max = p_fire.max()
    two64 = 2.0 ** 64
    for k in range(1, 4):
        assert math.isclose(float(t["skip"][k]) / two64, (1 - p_max) ** k, rel_tol=1e-15)
    for c in range(3):
        assert math.isclose((float(t["acc"][c]) + 1) / two64, p_fire[c] / p_max, rel_tol=1e-15)
        n = int(t["n_out"][c])
        cdf = t["cdf"][int(t["off_cdf"][c]):int(t["off_cdf"][c]) + n].astype(np.float64)
        q = np.cumsum(probs[c][1:]) / p_fire[c]
        assert int(t["cdf"][int(t["off_cdf"][c]) + n - 1]) == 2 ** 64 - 1
        assert np.allclose((cdf[:-1] + 1) / two64, q[:-1], rtol=1e-14)


def test_oracle_walk_has_the_exact_distribution():
    probs, T = model_a()
    nm = NoiseModel(probs, T, seed=7)
    t = nm.tables()
    B = 40_000
    rows = noise_ref.walk_rows(t, 7, B)
    P, counts, prow, nf = channel_patterns(probs, T)
    want = noise_ref.exact_marginals(probs, prow, counts, nf)
    got = rows.mean(axis=0)
    sig = np.sqrt(want * (1 - want) / B)
    assert np.all(np.abs(got - want) <= 5 * sig + 1e-12), (got, want)
    for i, j in [(0, 1), (1, 2), (2, 3), (0, 5), (3, 4)]:
        w = noise_ref.exact_pair_parity(probs, prow, counts, i, j)
        g = float(np.mean(rows[:, i] ^ rows[:, j]))
        assert abs(g - w) <= 5 * math.sqrt(w * (1 - w) / B) + 1e-12, (i, j, g, w)


def test_planted_table_error_is_caught():
    # the same statistics reject a walk with one channel's threshold doubled
    probs, T = model_a()
    t = NoiseModel(probs, T, seed=7).tables()
    t["acc"] = t["acc"].copy()
    t["acc"][0] = np.uint64(min(int(t["acc"][0]) * 2, 2 ** 64 - 1))
    B = 40_000
    rows = noise_ref.walk_rows(t, 7, B)
    P, counts, prow, nf = channel_patterns(probs, T)
    want = noise_ref.exact_marginals(probs, prow, counts, nf)
    sig = np.sqrt(want * (1 - want) / B)
    assert np.any(np.abs(rows.mean(axis=0) - want) > 5 * sig)


def test_rejects_bad_channels():
    with pytest.raises(ValueError):
        NoiseModel([np.array([0.5, 0.2, 0.3])], np.ones((1, 2), np.uint8))   # not 2^k
    with pytest.raises(ValueError):
        NoiseModel([np.array([0.5, 0.6])], np.ones((1, 1), np.uint8))        # sum != 1
    with pytest.raises(ValueError):
        NoiseModel([np.array([0.5, 0.25, 0.25, 0.0])], np.ones((1, 1), np.uint8))  # too few cols


def test_empty_model():
    nm = NoiseModel([], np.zeros((4, 0), np.uint8), seed=3)
    assert nm.n_channels == 0 and nm.num_f == 4
