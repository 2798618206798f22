"""GPU: graphs with approximate (complex64) floatfactors, the rotations
outside Clifford+T (src/python/ksim/compile/pipeline.py:362-372).  Their
terms are exact in Z[w], cast to float64, multiplied by their factors and
summed in float64 in term order (src/cpp/ksim/zx_eval.hpp:168-178).  Bars:
|amplitude| bit-identical to the oracle's float path (oracle/exact.py
graph_float) and within 1e-12 of the numpy restatement of the reference's
evaluate_flat; shot streams identical to the oracle's under fed uniforms
and the seeded stream, at every term split and in the batched and serving
paths; closed-form statistics at 5 sigma."""
import math

import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from ksim_amd.sampler import Component, HipProgramSampler
from oracle import exact, numpy_ref
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def approx_programs():
    return {
        "rz": synth.rz_program(0.9),
        "phase": synth.phase_program(-1.3),
        "shape": synth.with_approx_factors(
            synth.shape_program(n_q=6, t=8, F=6, n_out=5, seed=40, n_components=2), 1, frac=0.7),
        "wide_rows": synth.with_approx_factors(synth.wide_exponent_program(seed=5), 2, frac=1.0),
        "mixed": synth.product_program([synth.rz_program(0.4), synth.htkh_program(3),
                                        synth.with_approx_factors(
                                            synth.shape_program(n_q=5, t=6, F=4, n_out=4,
                                                                seed=3), 3)]),
        "wide_params": synth.with_approx_factors(
            synth.shape_program(n_q=6, t=4, F=300, n_out=3, seed=6), 4),
    }


@pytest.mark.parametrize("name", list(approx_programs()))
def test_eval_abs_equals_oracle(name):
    prog = approx_programs()[name]
    assert not HipProgramSampler(prog).exact
    rng = np.random.default_rng(1)
    for comp in prog.components:
        gds = [graph_to_dict(g) for g in comp.graphs]
        kc = Component(gds, len(comp.f_selection))
        F = len(comp.f_selection)
        for i, g in enumerate(comp.graphs):
            pv = rng.integers(0, 2, size=(24, F + i)).astype(np.uint8)
            got = kc.eval_abs(i, pv)
            want = np.array([exact.graph_magnitude(gds[i], r) for r in pv])
            assert np.array_equal(got, want), (name, i)
            if name == "wide_rows":
                # (1+sqrt2)^n' up to n' = 60 has coefficients beyond 2^63:
                # the reference's int64 products wrap there (exact_scalar.py
                # has no overflow check), the exact path does not
                continue
            ref = np.abs(numpy_ref.evaluate_graph(g, pv))
            assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, float(np.max(ref))), (name, i)


@pytest.mark.parametrize("name", list(approx_programs()))
def test_streams_equal_oracle(name):
    prog = approx_programs()[name]
    s = HipProgramSampler(prog)
    B = 400
    f = synth.random_f(B, max(s.min_num_f, 1), 0.5, 3)
    U = np.random.default_rng(4).random((B, s.tot_out))
    want_u = exact.sample_program(prog, f[:64], U[:64])
    idx = pick(B, 10, 5)
    want_s = oracle_rows(prog, f, 88, idx)
    ref = None
    for L in (0, 1, 2, 4):
        s.set_term_split(L)
        assert np.array_equal(s.sample_uniforms(f, U)[:64], want_u), (name, L)
        got = s.sample(f, 88)
        assert np.array_equal(got[idx], want_s), (name, L)
        if ref is None:
            ref = got
        assert np.array_equal(got, ref), (name, L)


def test_eval_exact_refuses_approximate_graphs():
    s = HipProgramSampler(synth.rz_program(0.5))
    with pytest.raises(NotImplementedError, match="approximate"):
        s.eval_exact(0, 1, np.ones((2, 1), np.uint8))
    # its exact graph 0 still evaluates exactly
    coeffs, e = s.eval_exact(0, 0, np.zeros((1, 0), np.uint8))
    assert exact.canonical(tuple(int(x) for x in coeffs[0]), int(e[0])) == ((1, 0, 0, 0), 0)


@pytest.mark.parametrize("theta", [0.9, 2.3])
def test_rz_statistics(theta):
    n = 200_000
    p = math.sin(theta / 2) ** 2
    got = HipProgramSampler(synth.rz_program(theta)).sample(np.zeros((n, 0), np.uint8), 17)
    assert abs(got[:, 0].mean() - p) < 5 * math.sqrt(p * (1 - p) / n)


def test_approximate_programs_batched_and_served():
    import torch
    from ksim_amd.batch import BatchedSampler, sample_static
    from ksim_amd.serve import Request, ServingLoop
    progs = list(approx_programs().values())[:5] + [synth.shape_program(n_q=6, t=4, F=5,
                                                                         n_out=6, seed=9)]
    samplers = [HipProgramSampler(p) for p in progs]
    shots = [300, 129, 700, 64, 1000, 257]
    fs = [synth.random_f(B, max(s.min_num_f, 1), 0.3, k)
          for k, (B, s) in enumerate(zip(shots, samplers))]
    seeds = [900 + k for k in range(len(progs))]
    alone = [s.sample(f, sd) for s, f, sd in zip(samplers, fs, seeds)]
    for k, (p, f) in enumerate(zip(progs, fs)):
        idx = pick(f.shape[0], 4, k)
        assert np.array_equal(alone[k][idx], oracle_rows(p, f, seeds[k], idx)), k
    dfs = [torch.from_numpy(f).cuda() for f in fs]
    for tile_ops in (0.0, 1.0):
        log(f"approx batch tile_ops={tile_ops}")
        b = BatchedSampler(samplers, dfs, seeds, wave_shots=64, tile_ops=tile_ops)
        b.run(use_graph=True)
        torch.cuda.synchronize()
        b.check()
        for k, s in enumerate(samplers):
            assert np.array_equal(b.out[k].cpu().numpy()[:, :s.num_outputs], alone[k]), k
        outs = [torch.zeros((f.shape[0], max(s.num_outputs, 1)), dtype=torch.uint8, device="cuda")
                for s, f in zip(samplers, dfs)]
        sample_static(samplers, dfs, seeds, outs, wave_shots=64, tile_ops=tile_ops)
        for k, s in enumerate(samplers):
            assert np.array_equal(outs[k].cpu().numpy()[:, :s.num_outputs], alone[k]), k
    log("approx serving")
    reqs = [Request(key=k, prog=samplers[k], f=fs[k], seed=seeds[k]) for k in range(len(progs))]
    loop = ServingLoop(max_slots=4, step_ms=0.5, hbm_budget=1 << 30)
    loop.run(reqs, realtime=False, max_wall_s=60)
    loop.close()
    for k, r in enumerate(reqs):
        assert np.array_equal(r.out, alone[k]), k
