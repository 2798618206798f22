"""Pin the CPU oracle (oracle/exact.py, oracle/numpy_ref.py) before trusting it.

* Z[w] multiply: against complex multiplication through the embedding, and
  against identities that fix the basis (w^8 = 1, w^3 = -conj w,
  (1+w)^4 = 2 w^2 (1+sqrt2)^2).
* term formula: H.T^k.H has |amplitude|^2 = sin^4(k pi/8) exactly, i.e.
  X + Y sqrt2 with X, Y rational — a dropped factor, a wrong sign of the
  phase or a wrong power of two changes X or Y.
* exact sum: value-preserving term splitting leaves every graph value
  unchanged (brute force over all parameter vectors).
* float64 path: numpy restatement bit-identical to the reference's own
  evaluate_flat / sample_flat outputs (golden vectors written from the
  reference module); the exact oracle's |amplitude| within 1e-13 relative
  of the reference's float sum and its fed-uniform streams identical.
"""
import cmath
import math
from fractions import Fraction

import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from oracle import exact, numpy_ref
from tests.golden.golden_io import load_program

W = cmath.exp(1j * math.pi / 4)


def embed(x):
    a, b, c, d = x
    return a + b * W + c * 1j + d * W.conjugate()


def test_zmul_matches_complex_product():
    rng = np.random.default_rng(1)
    for _ in range(500):
        x = tuple(int(v) for v in rng.integers(-50, 50, 4))
        y = tuple(int(v) for v in rng.integers(-50, 50, 4))
        assert abs(embed(exact.zmul(x, y)) - embed(x) * embed(y)) < 1e-9


def test_basis_identities():
    w = exact.UNIT[1]
    p = exact.ONE
    for k in range(8):
        assert p == exact.UNIT[k]
        p = exact.zmul(p, w)
    assert p == exact.ONE
    assert exact.UNIT[3] == (0, 0, 0, -1)  # w^3 = -conj(w)
    lam = (1, 1, 0, 0)
    l4 = exact.zmul(exact.zmul(lam, lam), exact.zmul(lam, lam))
    up = (1, 1, 0, 1)  # 1 + sqrt2
    rhs = exact.zmul(exact.zmul((2, 0, 0, 0), exact.UNIT[2]), exact.zmul(up, up))
    assert l4 == rhs


@pytest.mark.parametrize("k", range(8))
def test_htkh_exact_closed_form(k):
    prog = synth.htkh_program(k)
    d = graph_to_dict(prog.components[0].graphs[1])
    c, e = exact.graph_value(d, np.array([1]))
    X, Y = exact.exact_abs_fraction_sq(c, e)
    # sin^2(k pi/8) = (1 - cos(k pi/4))/2 and cos(k pi/4) in {+-1, 0, +-sqrt2/2}
    cosv = {0: (1, 0), 1: (0, Fraction(1, 2)), 2: (0, 0), 3: (0, Fraction(-1, 2)),
            4: (-1, 0), 5: (0, Fraction(-1, 2)), 6: (0, 0), 7: (0, Fraction(1, 2))}[k]
    s_a = Fraction(1 - cosv[0], 2)
    s_b = -cosv[1] / 2
    # (s_a + s_b sqrt2)^2
    want_X = s_a * s_a + 2 * s_b * s_b
    want_Y = 2 * s_a * s_b
    assert (X, Y) == (want_X, want_Y)
    assert exact.magnitude(c, e) == pytest.approx(math.sin(k * math.pi / 8) ** 2, abs=1e-15)


def test_noisy_htkh_flips_with_f():
    prog = synth.htkh_program(1, noisy_f=True)
    d = graph_to_dict(prog.components[0].graphs[1])
    p0 = exact.magnitude(*exact.graph_value(d, np.array([0, 1])))
    p1 = exact.magnitude(*exact.graph_value(d, np.array([1, 1])))
    assert p0 == pytest.approx(math.sin(math.pi / 8) ** 2)
    assert p1 == pytest.approx(math.cos(math.pi / 8) ** 2)


def test_ghz_marginals():
    gs = [graph_to_dict(g) for g in synth.ghz_graphs(4)]
    for i in range(1, 4):
        for m in range(2 ** i):
            bits = [(m >> j) & 1 for j in range(i)]
            v = exact.magnitude(*exact.graph_value(gs[i + 1], np.array(bits + [1])))
            want = 0.5 if all(b == 1 for b in bits) else 0.0
            assert v == want


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_term_split_preserves_values(seed):
    base = synth.shape_program(n_q=3, t=2, F=2, n_out=3, seed=40 + seed)
    split = synth.split_program(base, 6, seed=seed)
    for gb, gs in zip(base.components[0].graphs, split.components[0].graphs):
        assert gs.pf_phase_idx.shape[0] >= gb.pf_phase_idx.shape[0]
        P = gb.np_params.shape[2]
        db, ds = graph_to_dict(gb), graph_to_dict(gs)
        for m in range(2 ** P):
            bits = np.array([(m >> j) & 1 for j in range(P)])
            assert exact.graph_value(db, bits) == exact.graph_value(ds, bits)


def test_canonical_unique():
    assert exact.canonical((4, 8, 0, 12), -3) == ((1, 2, 0, 3), -1)
    assert exact.canonical((0, 0, 0, 0), 7) == ((0, 0, 0, 0), 0)
    assert exact.canonical((3, 2, 0, 0), 5) == ((3, 2, 0, 0), 5)


def test_magnitude_is_scale_invariant():
    rng = np.random.default_rng(3)
    for _ in range(200):
        c = tuple(int(v) for v in rng.integers(-(1 << 40), 1 << 40, 4))
        e = int(rng.integers(-50, 10))
        k = int(rng.integers(0, 20))
        assert exact.magnitude(c, e) == exact.magnitude(tuple(v << k for v in c), e - k)


def test_golden_numpy_restatement_bit_exact(golden_files):
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        for ci, comp in enumerate(prog.components):
            for gi, g in enumerate(comp.graphs):
                ref = d[f"eval_out_{ci}_{gi}"]
                mine = numpy_ref.evaluate_graph(g, d[f"eval_params_{ci}_{gi}"])
                assert np.array_equal(mine.view(np.int64), ref.view(np.int64)), (fn, ci, gi)
        f = d["f_params"]
        for seed in (5, 6):
            s = numpy_ref.sample_flat(prog, f, np.random.default_rng(seed))
            assert np.array_equal(s, d[f"sample_seed_{seed}"]), (fn, seed)


def test_golden_exact_oracle(golden_files):
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        for ci, comp in enumerate(prog.components):
            for gi, g in enumerate(comp.graphs):
                gd = graph_to_dict(g)
                pv = d[f"eval_params_{ci}_{gi}"]
                ref = np.abs(d[f"eval_out_{ci}_{gi}"])
                for r in range(pv.shape[0]):
                    m = exact.magnitude(*exact.graph_value(gd, pv[r]))
                    assert abs(m - ref[r]) <= 1e-13 * max(ref[r], 1e-300) or m == ref[r], (fn, ci, gi, r)
        f = d["f_params"]
        for seed in (5, 6):
            U = exact.flat_uniforms(prog, f.shape[0], np.random.default_rng(seed))
            assert np.array_equal(exact.sample_program(prog, f, U), d[f"sample_seed_{seed}"])


def test_kokkos_stream_properties():
    # restated from kokkos_sim.cpp:43-46,215-217; no vectors exist in the
    # reference, so check an independent uint64 formulation and uniformity
    U = exact.kokkos_uniforms(1234, 2000, 3)
    with np.errstate(over="ignore"):
        s = np.arange(2000, dtype=np.uint64)
        r = np.uint64(1234) ^ (s * np.uint64(0x9E3779B97F4A7C15))
        r ^= r >> np.uint64(30)
        r = r * np.uint64(0xBF58476D1CE4E5B9)
        r ^= r >> np.uint64(27)
        r[r == 0] = 1
        cols = []
        for _ in range(3):
            r ^= r << np.uint64(13)
            r ^= r >> np.uint64(7)
            r ^= r << np.uint64(17)
            cols.append((r >> np.uint64(11)).astype(np.float64) / 2.0 ** 53)
    assert np.array_equal(U, np.stack(cols, axis=1))
    assert abs(U.mean() - 0.5) < 0.02
    assert np.all((U >= 0) & (U < 1))
