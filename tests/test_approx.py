"""Approximate (complex64) floatfactors on the CPU: the oracle's float path
(oracle/exact.py graph_float: exact terms cast to float64, times their
factors, summed in term order as src/cpp/ksim/zx_eval.hpp:168-178) is pinned
by closed forms of rotations outside Clifford+T, rejects planted mistakes,
agrees with the numpy restatement of the reference's evaluate_flat
(src/python/ksim/sample/evaluate.py:63-65) within 1e-12, and the CPU
baseline port samples the same shots."""
import math

import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from oracle import exact, numpy_ref

THETAS = (0.3, 1.1, -0.7, 2.9, -2.2)


def f32(x):
    return float(np.float32(x))


@pytest.mark.parametrize("theta", THETAS)
def test_rz_closed_form(theta):
    # P(1) = (1 - cos theta) / 2; with the stored complex64 factors the
    # float64 sum is exactly 1/2 - fl32(cos theta)/2 (imaginary parts cancel)
    d = graph_to_dict(synth.rz_program(theta).components[0].graphs[1])
    p1 = exact.graph_magnitude(d, [1])
    assert p1 == abs(0.5 - f32(math.cos(theta)) / 2)
    assert abs(p1 - math.sin(theta / 2) ** 2) < 1e-7
    assert abs(exact.graph_magnitude(d, [0]) - math.cos(theta / 2) ** 2) < 1e-7


def _phase_p1():
    return {th: exact.graph_magnitude(
        graph_to_dict(synth.phase_program(th).components[0].graphs[1]), [1]) for th in THETAS}


def test_phase_closed_form():
    for th, p1 in _phase_p1().items():
        assert abs(p1 - abs(math.cos((th + math.pi / 4) / 2))) < 3e-7, th


@pytest.mark.parametrize("planted", ["conjugate", "swap", "drop_w"])
def test_phase_closed_form_rejects_planted_errors(planted, monkeypatch):
    cast = exact.term_float
    fac = exact.approx_factor
    if planted == "conjugate":
        monkeypatch.setattr(exact, "approx_factor", lambda d, t: (fac(d, t)[0], -fac(d, t)[1]))
    elif planted == "swap":
        monkeypatch.setattr(exact, "term_float", lambda c, e: cast(c, e)[::-1])
    else:  # the cast drops the w coefficient
        monkeypatch.setattr(exact, "term_float",
                            lambda c, e: cast((c[0], 0, c[2], 0), e))
    bad = [abs(p1 - abs(math.cos((th + math.pi / 4) / 2))) for th, p1 in _phase_p1().items()]
    assert max(bad) > 1e-3


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_float_path_equals_the_reference_formula(seed):
    prog = synth.with_approx_factors(
        synth.shape_program(n_q=6, t=8, F=6, n_out=5, seed=40 + seed, n_components=2), seed,
        frac=0.7)
    rng = np.random.default_rng(seed)
    n = 0
    for comp in prog.components:
        F = len(comp.f_selection)
        for i, g in enumerate(comp.graphs):
            if not g.pf_has_approx:
                continue
            d = graph_to_dict(g)
            pv = rng.integers(0, 2, size=(12, F + i)).astype(np.uint8)
            ref = numpy_ref.evaluate_graph(g, pv)
            for r in range(12):
                sr, si = exact.graph_float(d, pv[r])
                z = complex(sr, si)
                assert abs(z - ref[r]) <= 1e-12 * max(1.0, abs(ref[r])), (i, r)
                n += 1
    assert n > 0


def test_port_samples_approximate_programs_like_the_oracle():
    from tests.test_port import kokkos_columns, port_sample
    prog = synth.product_program([synth.rz_program(0.9), synth.phase_program(-0.4),
                                  synth.htkh_program(3)])
    B = 3000
    f = np.zeros((B, 0), np.uint8)
    want = exact.sample_program(prog, f, kokkos_columns(prog, B, 12))
    assert np.array_equal(port_sample(prog, f, seed=12), want)
