"""Pin the C exact oracle (oracle/exact_fast.c) to the Python big-integer
oracle (oracle/exact.py), which tests/test_oracle.py pins to closed forms and
to the reference's golden vectors: the same canonical exact value for every
graph and parameter row tried, the same sampled shots, and a refusal (never a
wrapped value) for a term beyond its 512-bit range."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from oracle import exact, exact_fast
from tests.golden.golden_io import load_program


def _rows(rng, P, n):
    return rng.integers(0, 2, size=(n, P)).astype(np.uint8)


def _same_values(prog, rows, seed):
    rng = np.random.default_rng(seed)
    n = 0
    for comp in prog.components:
        F = len(comp.f_selection)
        for i, g in enumerate(comp.graphs):
            d = graph_to_dict(g)
            pv = _rows(rng, F + i, rows)
            got = exact_fast.graph_values(d, pv)
            for r in range(rows):
                assert got[r] == exact.graph_value(d, pv[r]), (i, r)
                n += 1
    return n


def test_golden_values_and_streams(golden_files):
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        for ci, comp in enumerate(prog.components):
            for gi, g in enumerate(comp.graphs):
                gd = graph_to_dict(g)
                pv = d[f"eval_params_{ci}_{gi}"]
                got = exact_fast.graph_values(gd, pv)
                assert got == [exact.graph_value(gd, r) for r in pv], (fn, ci, gi)
        f = d["f_params"]
        for seed in (5, 6):
            U = exact.flat_uniforms(prog, f.shape[0], np.random.default_rng(seed))
            assert np.array_equal(exact_fast.sample_program(prog, f, U), d[f"sample_seed_{seed}"])


@pytest.mark.parametrize("prog_fn", [
    lambda: synth.shape_program(n_q=8, t=8, F=6, n_out=5, seed=5, generic_ff=0.5, proj_frac=0.2),
    lambda: synth.shape_program(n_q=5, t=10, F=90, n_out=40, seed=6, proj_frac=0.0),
    lambda: synth.shape_program(n_q=6, t=36, F=72, n_out=2, seed=7, generic_ff=0.3),
    lambda: synth.shape_program(n_q=12, t=20, F=40, n_out=6, seed=8, n_components=2),
    lambda: synth.wide_exponent_program(seed=5),
    lambda: synth.split_program(synth.htkh_program(3), 20, seed=1),
])
def test_random_program_values(prog_fn):
    assert _same_values(prog_fn(), 4, 3) > 0


def test_sampled_shots_equal():
    prog = synth.shape_program(n_q=6, t=6, F=9, n_out=7, seed=41, n_components=2)
    B = 60
    f = synth.random_f(B, 9, 0.5, 3)
    U = np.random.default_rng(2).random((B, 7))
    assert np.array_equal(exact_fast.sample_program(prog, f, U), exact.sample_program(prog, f, U))


def test_refuses_values_beyond_its_range():
    # (1 + w)^900 has coefficients of about 570 bits after halving out every
    # factor 2: the C oracle must refuse it, never return a wrapped value
    g = synth.empty_graph(1, 0, Tnp=900)
    g.np_phases[0, :] = 1
    g.np_counts[0] = 900
    d = graph_to_dict(g)
    with pytest.raises(OverflowError):
        exact_fast.graph_values(d, np.zeros((1, 0), np.uint8))
    # 300 factors fit and agree with the big-integer oracle
    g.np_counts[0] = 300
    d = graph_to_dict(g)
    assert exact_fast.graph_values(d, np.zeros((1, 0), np.uint8))[0] == \
        exact.graph_value(d, np.zeros(0, np.int64))
