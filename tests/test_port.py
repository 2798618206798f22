"""The CPU baseline (oracle/kokkos_port.cpp, a restatement of the reference's
native sampler: u8 masks, a Z[w] multiply per factor, float64 term sum,
thread per shot) computes the reference algorithm: its shots equal the
exact oracle's under fed uniforms and under the seeded xorshift64 stream,
on every golden program and on closed-form programs with many terms.  The
bench times this baseline, so a bug here would misstate it."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from oracle import exact
from oracle.build_ref import PortComponent
from tests.golden.golden_io import load_program

SEED_STRIDE = 0x9E37  # per-component seed offset, src/python/ksim/sample/sampler.py:63


def port_sample(prog, f, seed=None, U=None):
    """The reference's KokkosProgramSampler.sample (sampler.py:57-67) over
    port components: per-component f columns, seed + 0x9E37 k (or fed
    uniform columns), concatenation, argsort(output_order)."""
    outs, col = [], 0
    for k, comp in enumerate(prog.components):
        n_out = len(comp.graphs) - 1
        pc = PortComponent([graph_to_dict(g) for g in comp.graphs], len(comp.f_selection))
        fs = np.ascontiguousarray(f[:, np.asarray(comp.f_selection, np.int64)], dtype=np.uint8)
        Uk = None if U is None else np.ascontiguousarray(U[:, col:col + n_out])
        s = 0 if seed is None else seed + SEED_STRIDE * k
        outs.append(pc.sample(fs, s, uniforms=Uk)[:, :n_out])
        col += n_out
    return np.concatenate(outs, axis=1)[:, np.argsort(prog.output_order)]


def kokkos_columns(prog, B, seed):
    return np.concatenate([exact.kokkos_uniforms(seed + SEED_STRIDE * k, B, len(c.graphs) - 1)
                           for k, c in enumerate(prog.components)], axis=1)


def test_port_fed_uniforms_equal_oracle(golden_files):
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        f = d["f_params"]
        U = np.random.default_rng(11).random((f.shape[0], sum(len(c.graphs) - 1
                                                             for c in prog.components)))
        assert np.array_equal(port_sample(prog, f, U=U), exact.sample_program(prog, f, U)), fn


def test_port_seeded_stream_equals_oracle(golden_files):
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        f = d["f_params"]
        for seed in (0, 43, 2 ** 40 + 5):
            want = exact.sample_program(prog, f, kokkos_columns(prog, f.shape[0], seed))
            assert np.array_equal(port_sample(prog, f, seed=seed), want), (fn, seed)


@pytest.mark.parametrize("prog_fn", [
    lambda: synth.split_program(synth.htkh_program(3), 20, seed=1),
    lambda: synth.split_program(synth.single_program(synth.ghz_graphs(6)), 12, seed=2),
    lambda: synth.product_program([synth.htkh_program(1), synth.htkh_program(3, noisy_f=True),
                                   synth.htht_h_program()]),
])
def test_port_closed_form_programs(prog_fn):
    prog = prog_fn()
    F = max([int(np.max(c.f_selection)) + 1 for c in prog.components if len(c.f_selection)],
            default=0)
    B = 400
    f = synth.random_f(B, F, 0.5, 3)
    want = exact.sample_program(prog, f, kokkos_columns(prog, B, 77))
    assert np.array_equal(port_sample(prog, f, seed=77), want)


def test_port_known_answer():
    # H.T.H.T.H, seed 43, 1M shots: the reference's published 0.25034
    # (docs/benchmarks.md:124-125) through the baseline itself
    prog = synth.htht_h_program()
    out = port_sample(prog, np.zeros((1_000_000, 0), np.uint8), seed=43)
    assert round(float(out[:, 0].mean()), 5) == 0.25034
