"""GPU: components wider than the register words (F + n_out > 256
parameters, up to 4096), whose parameter words the kernels keep in LDS,
against the oracle: exact graph values, eval_abs, fed-uniform and seeded
streams through every plain entry point and term split, in-kernel noise,
packed rows, the batched, serving and no-allocation paths, and seeded random
programs around the 256-parameter boundary.  The reference walks any
parameter count (src/cpp/ksim/zx_eval.hpp:86-97)."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from ksim_amd.sampler import Component, HipProgramSampler, unpack_outputs
from oracle import exact, exact_fast
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def _exact_graphs(prog, s, rows, seed, graphs=None):
    rng = np.random.default_rng(seed)
    for c, comp in enumerate(prog.components):
        F = len(comp.f_selection)
        idx = graphs if graphs is not None else range(len(comp.graphs))
        for i in idx:
            if i >= len(comp.graphs):
                continue
            d = graph_to_dict(comp.graphs[i])
            pv = rng.integers(0, 2, size=(rows, F + i)).astype(np.uint8)
            coeffs, e = s.eval_exact(c, i, pv)
            want = exact_fast.graph_values(d, pv)
            for r in range(rows):
                assert exact.canonical(tuple(int(x) for x in coeffs[r]), int(e[r])) == want[r], \
                    (c, i, r)


@pytest.mark.parametrize("F,n_out,n_comp", [(250, 7, 1), (300, 3, 1), (1000, 2, 1),
                                            (520, 40, 3)])
def test_wide_program_against_oracle(F, n_out, n_comp):
    log(f"wide: F={F} n_out={n_out} n_comp={n_comp}")
    prog = synth.shape_program(n_q=6, t=4, F=F, n_out=n_out, seed=F + n_out,
                               n_components=n_comp)
    s = HipProgramSampler(prog)
    _exact_graphs(prog, s, 4, 1)
    B = 700
    f = synth.random_f(B, s.min_num_f, 0.5, 2)
    U = np.random.default_rng(3).random((B, s.tot_out))
    nb = 40
    want_u = exact_fast.sample_program(prog, f[:nb], U[:nb])
    idx = pick(B, 8, 4)
    want_s = oracle_rows(prog, f, 77, idx, fast=True)
    ref = None
    for L in (0, 1, 2, 4):
        log(f"wide: term split {L}")
        s.set_term_split(L)
        got_u = s.sample_uniforms(f, U)
        assert np.array_equal(got_u[:nb], want_u), L
        got = s.sample(f, 77)
        assert np.array_equal(got[idx], want_s), L
        if ref is None:
            ref = got
        assert np.array_equal(got, ref), L
        assert np.array_equal(unpack_outputs(s.sample_packed(f, 77), s.num_outputs), ref), L


def test_component_900_f_bits_120_outputs():
    # one component of 900 f bits and 120 outputs (1020 parameters, 32
    # words): samples bit-exactly against the oracle
    prog = synth.shape_program(n_q=8, t=4, F=900, n_out=120, seed=9)
    comp = prog.components[0]
    gds = [graph_to_dict(g) for g in comp.graphs]
    log("wide 900/120: building")
    kc = Component(gds, 900)
    s = HipProgramSampler(prog)
    B = 512
    f = synth.random_f(B, 900, 0.5, 5)
    idx = pick(B, 6, 5)
    log("wide 900/120: Component.sample against the oracle")
    got = kc.sample(f, 31)
    U = exact.kokkos_uniforms_at(31, idx, 120)
    want = exact_fast.sample_component(gds, f[idx], U)
    assert np.array_equal(got[idx], want)
    log("wide 900/120: program sample, fed uniforms")
    Uf = np.random.default_rng(6).random((B, 120))
    got_u = s.sample_uniforms(f, Uf)
    assert np.array_equal(got_u[idx], exact_fast.sample_program(prog, f[idx], Uf[idx]))
    log("wide 900/120: eval_abs and exact values")
    for gi in (0, 1, 60, 120):
        pv = np.random.default_rng(gi).integers(0, 2, size=(5, 900 + gi)).astype(np.uint8)
        vals = exact_fast.graph_values(gds[gi], pv)
        assert np.array_equal(kc.eval_abs(gi, pv), np.array([exact.magnitude(*v) for v in vals]))
    _exact_graphs(prog, s, 2, 7, graphs=(0, 119, 120))


def test_wide_noise_in_kernel_equals_rows():
    from tests.test_gpu_random import random_noise
    rng = np.random.default_rng(8)
    prog = synth.shape_program(n_q=6, t=2, F=400, n_out=9, seed=21, n_components=2)
    s = HipProgramSampler(prog)
    nm = random_noise(s.min_num_f + 3, rng)
    rows = nm.sample(900, shot_base=0)
    for L in (1, 4):
        s.set_term_split(L)
        assert np.array_equal(s.sample_noisy(nm, 900, 9), s.sample(rows, 9)), L


@pytest.mark.parametrize("seed", [31, 32])
def test_random_programs_around_256_parameters(seed):
    rng = np.random.default_rng(seed)
    for _ in range(10):
        F = int(rng.choice([200, 224, 250, 255, 256, 257, 288, 300, 513, 700]))
        n_out = int(rng.choice([1, 2, 5, 31, 33]))
        nc = int(rng.integers(1, 3))
        prog = synth.shape_program(n_q=int(rng.integers(4, 8)), t=int(rng.integers(0, 7)), F=F,
                                   n_out=max(n_out, nc), seed=int(rng.integers(1 << 30)),
                                   n_components=nc)
        if rng.random() < 0.4:
            for c in prog.components:
                c.f_selection = rng.permutation(np.asarray(c.f_selection))
        s = HipProgramSampler(prog)
        B = int(rng.integers(1, 300))
        f = synth.random_f(B, s.min_num_f, float(rng.choice([1e-3, 0.5])), B)
        idx = pick(B, 4, B)
        want = oracle_rows(prog, f, 5, idx, fast=True)
        for L in (1, 4):
            s.set_term_split(L)
            assert np.array_equal(s.sample(f, 5)[idx], want), (F, n_out, nc, L)


def test_wide_programs_batched_served_and_without_allocation():
    import torch
    from ksim_amd.batch import BatchedSampler, sample_static
    from ksim_amd.serve import Request, ServingLoop
    progs = [synth.shape_program(n_q=6, t=4, F=300, n_out=6, seed=41),
             synth.shape_program(n_q=5, t=4, F=3, n_out=5, seed=1),
             synth.shape_program(n_q=8, t=10, F=10, n_out=9, seed=2, n_components=2),
             synth.shape_program(n_q=6, t=2, F=600, n_out=3, seed=42, n_components=2)]
    samplers = [HipProgramSampler(p) for p in progs]
    shots = [400, 700, 130, 257]
    fs = [synth.random_f(B, max(s.min_num_f, 1), 0.3, k)
          for k, (B, s) in enumerate(zip(shots, samplers))]
    seeds = [600 + k for k in range(4)]
    alone = [s.sample(f, sd) for s, f, sd in zip(samplers, fs, seeds)]
    for k, (p, f) in enumerate(zip(progs, fs)):
        idx = pick(f.shape[0], 3, k)
        assert np.array_equal(alone[k][idx], oracle_rows(p, f, seeds[k], idx, fast=True)), k
    dfs = [torch.from_numpy(f).cuda() for f in fs]
    for tile_ops in (0.0, 1.0):
        log(f"wide batch tile_ops={tile_ops}")
        b = BatchedSampler(samplers, dfs, seeds, wave_shots=64, tile_ops=tile_ops)
        for use_graph in (False, True):
            for o in b.out:
                o.fill_(7)
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            b.check()
            for k, s in enumerate(samplers):
                assert np.array_equal(b.out[k].cpu().numpy()[:, :s.num_outputs], alone[k]), k
        outs = [torch.full((f.shape[0], max(s.num_outputs, 1)), 7, dtype=torch.uint8,
                           device="cuda") for s, f in zip(samplers, dfs)]
        sample_static(samplers, dfs, seeds, outs, wave_shots=64, tile_ops=tile_ops)
        for k, s in enumerate(samplers):
            assert np.array_equal(outs[k].cpu().numpy()[:, :s.num_outputs], alone[k]), k
    log("wide serving")
    reqs = [Request(key=k, prog=samplers[k], f=fs[k], seed=seeds[k]) for k in range(4)]
    loop = ServingLoop(max_slots=4, step_ms=0.5, hbm_budget=1 << 30)
    loop.run(reqs, realtime=False, max_wall_s=60)
    loop.close()
    for k, r in enumerate(reqs):
        assert np.array_equal(r.out, alone[k]), k
