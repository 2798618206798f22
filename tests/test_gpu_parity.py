"""GPU parity: the HIP sampler against the exact oracle and the reference's
golden vectors.  Bars: exact Z[w] values equal (canonical forms), float64
|amplitude| bit-identical to the oracle's, shot streams bit-identical when
both sides get the same uniforms."""
import math

import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from ksim_amd.sampler import Component, HipProgramSampler, pack_words, sample_flat
from oracle import exact
from tests.golden.golden_io import load_program

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def gpu_exact_matches(prog, rows, seed):
    s = HipProgramSampler(prog)
    rng = np.random.default_rng(seed)
    n = 0
    for c, comp in enumerate(prog.components):
        for i, g in enumerate(comp.graphs):
            P = g.np_params.shape[2]
            pv = rng.integers(0, 2, size=(rows, P)).astype(np.uint8)
            coeffs, e = s.eval_exact(c, i, pv)
            d = graph_to_dict(g)
            for r in range(rows):
                want = exact.graph_value(d, pv[r])
                got = exact.canonical(tuple(int(x) for x in coeffs[r]), int(e[r]))
                assert got == want, (c, i, r)
                n += 1
    return n


def test_eval_exact_golden(golden_files):
    for fn in golden_files:
        log(f"eval_exact {fn}")
        gpu_exact_matches(load_program(np.load(fn)), 16, 1)


@pytest.mark.parametrize("shape", [dict(n_q=8, t=8, F=6, n_out=5, generic_ff=0.5, proj_frac=0.2),
                                   dict(n_q=6, t=4, F=40, n_out=8),
                                   dict(n_q=5, t=10, F=90, n_out=60, proj_frac=0.0),
                                   dict(n_q=9, t=0, F=100, n_out=29),
                                   dict(n_q=6, t=36, F=72, n_out=2, generic_ff=0.3)])
def test_eval_exact_shapes(shape):
    log(f"eval_exact shape {shape}")
    prog = synth.shape_program(seed=5, **shape)
    assert gpu_exact_matches(prog, 8, 2) > 0


def test_wide_table_rows():
    # |n'| up to ~60: terms beyond the single-digit rows add digit by digit
    log("eval_exact wide rows")
    prog = synth.wide_exponent_program(seed=5)
    s = HipProgramSampler(prog)
    st = s.stats()
    assert st["wide_terms"] > 0 and st["n2_bound"] > 46
    assert gpu_exact_matches(prog, 64, 3) > 0
    log("streams wide rows")
    B = 256
    f = synth.random_f(B, 6, 0.5, 9)
    U = np.random.default_rng(9).random((B, prog.num_outputs))
    want = exact.sample_program(prog, f, U)
    # 128-bit sums through the plain kernel and the split kernel
    for L in (1, 2, 4, 0):
        s.set_term_split(L)
        assert np.array_equal(s.sample_uniforms(f, U), want), L


def test_mixed_widths_program():
    # one launch holding a component with 128-bit graphs and components
    # whose graphs are single-term or 64-bit: every graph takes its own path
    log("mixed widths")
    prog = synth.product_program([synth.wide_exponent_program(n_out=2, seed=3),
                                  synth.shape_program(n_q=6, t=6, F=4, n_out=3, seed=8),
                                  synth.htkh_program(3)])
    s = HipProgramSampler(prog)
    assert s.stats()["wide_terms"] > 0
    B = 200
    f = synth.random_f(B, s.min_num_f, 0.5, 4)
    U = np.random.default_rng(4).random((B, s.tot_out))
    want = exact.sample_program(prog, f, U)
    for L in (1, 4):
        s.set_term_split(L)
        assert np.array_equal(s.sample_uniforms(f, U), want), L


def test_eval_abs_matches_oracle_and_reference(golden_files):
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        if len(prog.components) != 1:
            continue
        comp = prog.components[0]
        gds = [graph_to_dict(g) for g in comp.graphs]
        kc = Component(gds, len(comp.f_selection))
        assert kc.n_out == len(gds) - 1
        for gi, g in enumerate(comp.graphs):
            pv = np.ascontiguousarray(d[f"eval_params_0_{gi}"])
            got = kc.eval_abs(gi, pv)
            want = np.array([exact.magnitude(*exact.graph_value(gds[gi], r)) for r in pv])
            assert np.array_equal(got, want), (fn, gi)
            ref = np.abs(d[f"eval_out_0_{gi}"])
            # the reference's own amplitude precision gate is 1e-9 (test_nonclifford.py:160)
            assert np.max(np.abs(got - ref)) < 1e-12, (fn, gi)


def test_fed_uniform_streams_golden(golden_files):
    for fn in golden_files:
        log(f"streams {fn}")
        d = np.load(fn)
        prog = load_program(d)
        f = d["f_params"]
        s = HipProgramSampler(prog)
        for seed in (5, 6):
            # sample_flat's own stream (PCG64 draws in its order): reference output
            got = sample_flat(prog, f, np.random.default_rng(seed), sampler=s)
            assert np.array_equal(got, d[f"sample_seed_{seed}"]), (fn, seed)


def test_run_sampler_cpu_backend_is_the_reference_numpy_stream(golden_files):
    # run.py:89-90: the "cpu" backend is sample_flat(flat, f, default_rng(seed));
    # the golden files hold the reference's own output of exactly that call
    from ksim_amd.sampler import run_sampler
    for fn in golden_files:
        d = np.load(fn)
        prog = load_program(d)
        for seed in (5, 6):
            got = run_sampler(prog, d["f_params"], seed, backend="cpu")
            assert np.array_equal(got, d[f"sample_seed_{seed}"]), (fn, seed)


@pytest.mark.parametrize("seed", [0, 1])
def test_fed_uniform_streams_shape(seed):
    prog = synth.shape_program(n_q=6, t=6, F=8, n_out=6, seed=20 + seed, n_components=2)
    B = 96
    f = synth.random_f(B, 8, 0.5, seed)
    U = np.random.default_rng(seed).random((B, 6))
    got = HipProgramSampler(prog).sample_uniforms(f, U)
    want = exact.sample_program(prog, f, U)
    assert np.array_equal(got, want)


def test_seeded_stream_is_the_reference_kokkos_stream():
    prog = synth.product_program([synth.htkh_program(1), synth.htkh_program(3, noisy_f=True),
                                  synth.shape_program(n_q=5, t=4, F=3, n_out=4, seed=3)])
    B = 300
    num_f = sum(len(c.f_selection) for c in prog.components)
    f = synth.random_f(B, num_f, 0.4, 9)
    seed = 1234
    got = HipProgramSampler(prog).sample(f, seed)
    # reference: component k uses seed + 0x9E37 k, xorshift64 per shot
    cols = []
    for k, comp in enumerate(prog.components):
        cols.append(exact.kokkos_uniforms(seed + 0x9E37 * k, B, len(comp.graphs) - 1))
    U = np.concatenate(cols, axis=1)
    want = exact.sample_program(prog, f, U)
    assert np.array_equal(got, want)


def test_seeded_known_answer_htht_h():
    # The reference's GPU sampler on H.T.H.T.H (P(1) = 1/4 exactly) measured
    # 0.25034 at 1,000,000 shots with seed 42 + 1 on component 0
    # (docs/benchmarks.md:82,124-125, benchmarks/bench_xtim.py:146,164-167).
    # The same seeded call here gives the same count of ones, shot for shot
    # the oracle's stream, through every entry point the reference's
    # callers use.
    from ksim_amd.sampler import run_sampler
    B, seed = 1_000_000, 43
    prog = synth.htht_h_program()
    f = np.zeros((B, 0), np.uint8)
    want = exact.kokkos_uniforms(seed, B, 1)[:, 0] < 0.25
    log("known answer: HipProgramSampler")
    got = HipProgramSampler(prog).sample(f, seed)[:, 0]
    assert round(float(got.mean()), 5) == 0.25034
    assert np.array_equal(got.astype(bool), want)
    log("known answer: run_sampler(backend='gpu')")
    assert np.array_equal(run_sampler(prog, f, seed, backend="gpu")[:, 0], got)
    log("known answer: Component")
    kc = Component([graph_to_dict(g) for g in prog.components[0].graphs], 0)
    assert np.array_equal(kc.sample(f, seed)[:, 0], got)
    # the same with the terms split over more terms and 2 / 4 waves per tile
    log("known answer: split terms, waves per tile")
    s = HipProgramSampler(synth.split_program(prog, 6, seed=3))
    for L in (1, 2, 4):
        s.set_term_split(L)
        assert np.array_equal(s.sample(f[:200_000], seed)[:, 0], got[:200_000]), L


def test_component_sample_equals_oracle_stream():
    # Component.sample (kokkos_sim.cpp:188-240): f_sel rows in, (B, n_out)
    # bits out, drawn from the xorshift64 stream of its seed
    prog = synth.shape_program(n_q=6, t=6, F=9, n_out=7, seed=41)
    comp = prog.components[0]
    gds = [graph_to_dict(g) for g in comp.graphs]
    kc = Component(gds, 9)
    B = 512
    f = synth.random_f(B, 9, 0.5, 8)
    for seed in (0, 99, 2 ** 63 + 17):
        U = exact.kokkos_uniforms(seed, B, kc.n_out)
        want = exact.sample_component(gds, f.astype(np.int64), U)
        assert np.array_equal(kc.sample(f, seed), want), seed


def test_component_api_edges():
    prog = synth.htkh_program(1, noisy_f=True)
    gds = [graph_to_dict(g) for g in prog.components[0].graphs]
    kc = Component(gds, 1)
    with pytest.raises(ValueError, match="column count"):
        kc.sample(np.zeros((4, 2), np.uint8), 1)
    with pytest.raises(TypeError):
        kc.sample(np.zeros((4, 1), np.int64), 1)
    out = kc.sample(np.zeros((0, 1), np.uint8), 1)
    assert out.shape == (0, 1)
    # n_out = 0: (B, 1) of zeros
    k0 = Component([graph_to_dict(synth.constant_graph(0))], 0)
    assert k0.n_out == 0
    assert np.array_equal(k0.sample(np.zeros((5, 0), np.uint8), 3), np.zeros((5, 1), np.uint8))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_htkh_statistics(k):
    """P(1) = sin^2(k pi/8) within 5 sigma at 20k shots, seed 11 — the known
    answer of the reference's tests/test_nonclifford.py:105-121."""
    prog = synth.split_program(synth.htkh_program(k), 5, seed=k)
    n = 20_000
    exact_p = math.sin(k * math.pi / 8) ** 2
    sigma = math.sqrt(exact_p * (1 - exact_p) / n)
    p1 = HipProgramSampler(prog).sample(np.zeros((n, 0), np.uint8), seed=11)[:, 0].mean()
    assert abs(p1 - exact_p) < 5 * sigma


def test_ghz_and_noise_statistics():
    prog = synth.product_program([synth.single_program(synth.ghz_graphs(6)),
                                  synth.htkh_program(2, noisy_f=True)])
    n = 40_000
    f = synth.random_f(n, 1, 0.25, 4)
    out = HipProgramSampler(prog).sample(f, seed=7)
    # output_order reversed: the GHZ block is the last 6 columns
    ghz = out[:, 1:]
    assert np.all(ghz == ghz[:, :1])
    assert abs(ghz[:, 0].mean() - 0.5) < 5 * math.sqrt(0.25 / n)
    p = 0.75 * math.sin(2 * math.pi / 8) ** 2 + 0.25 * math.cos(2 * math.pi / 8) ** 2
    assert abs(out[:, 0].mean() - p) < 5 * math.sqrt(p * (1 - p) / n)


def test_shot_range_split_is_identical():
    import torch
    from ksim_amd import _lib
    import ctypes as C
    prog = synth.shape_program(n_q=6, t=6, F=5, n_out=7, seed=31)
    s = HipProgramSampler(prog)
    B = 1000
    f = torch.from_numpy(synth.random_f(B, 5, 0.5, 2)).cuda()
    full = torch.zeros((B, 7), dtype=torch.uint8, device="cuda")
    parts = torch.zeros((B, 7), dtype=torch.uint8, device="cuda")
    lib = _lib.load()
    _lib.check(lib.ksim_program_sample_device(s.handle, f.data_ptr(), B, 5, 0, 99, None,
                                              full.data_ptr(), None))
    cut = 437
    _lib.check(lib.ksim_program_sample_device(s.handle, f.data_ptr(), cut, 5, 0, 99, None,
                                              parts.data_ptr(), None))
    _lib.check(lib.ksim_program_sample_device(s.handle, f[cut:].data_ptr(), B - cut, 5, cut, 99,
                                              None, parts[cut:].data_ptr(), None))
    _lib.check(lib.ksim_program_check(s.handle, None))
    torch.cuda.synchronize()
    assert torch.equal(full, parts)
    assert torch.equal(full.cpu(), torch.from_numpy(s.sample(f.cpu().numpy(), 99)))


def test_program_blob_round_trip():
    from ksim_amd.sampler import program_blob
    prog = synth.shape_program(n_q=7, t=6, F=9, n_out=6, seed=11, n_components=2)
    s = HipProgramSampler(prog)
    blob = s.to_bytes()
    assert blob == program_blob(prog)
    t = HipProgramSampler.from_bytes(blob)
    assert (t.n_comp, t.num_outputs, t.tot_out, t.min_num_f) == \
        (s.n_comp, s.num_outputs, s.tot_out, s.min_num_f)
    assert t.ops_per_shot == s.ops_per_shot
    f = synth.random_f(777, 9, 0.5, 4)
    assert np.array_equal(t.sample(f, 21), s.sample(f, 21))
    with pytest.raises(ValueError):
        HipProgramSampler.from_bytes(blob[:-4])


@pytest.mark.parametrize("name,make", [
    ("blob_v4_two_comp.bin",
     lambda: synth.shape_program(n_q=6, t=6, F=8, n_out=4, seed=3, n_components=2)),
    ("blob_v4_wide.bin", synth.wide_exponent_program)])
def test_version_4_blobs_sample_as_the_oracle(name, make):
    # blobs of the version-4 packer hold every LIN record with its own delta
    # (no sign-record groups): uploaded as they are, they give the oracle's
    # shot stream and the same rows as the current packing of the program
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", name), "rb") as fh:
        blob = fh.read()
    prog = make()
    old = HipProgramSampler.from_bytes(blob)
    new = HipProgramSampler(prog)
    assert old.to_bytes() != new.to_bytes()
    B = 3000
    f = synth.random_f(B, old.min_num_f, 0.5, 8)
    U = np.random.default_rng(9).random((B, old.tot_out))
    got = old.sample_uniforms(f, U)
    assert np.array_equal(got, new.sample_uniforms(f, U))
    assert np.array_equal(got[:300], exact.sample_program(prog, f[:300], U[:300]))
    assert np.array_equal(old.sample(f, 5), new.sample(f, 5))


@pytest.mark.parametrize("n_sign", [1, 31, 32, 33, 64, 65, 97])
def test_sign_record_groups(n_sign):
    # sign-record groups of every size 1-32, several per chunk, mixed with LIN
    # and PAIR records: exact values and the fed-uniform stream as the oracle
    from tests.test_packer import sign_record_program
    prog = sign_record_program(n_sign, seed=n_sign)
    assert gpu_exact_matches(prog, rows=16, seed=n_sign) > 0
    s = HipProgramSampler(prog)
    B = 500
    f = synth.random_f(B, s.min_num_f, 0.5, n_sign)
    U = np.random.default_rng(n_sign).random((B, s.tot_out))
    assert np.array_equal(s.sample_uniforms(f, U), exact.sample_program(prog, f, U))


@pytest.mark.parametrize("split", [1, 2, 4])
def test_parity_groups(split):
    # a program on the 128-bit kernels carries its sign records, PPX pairs
    # and projectors as nibble parity tables read through LDS: exact values
    # (the eval kernel's buffer), and the fed-uniform stream through the
    # plain kernel (its own buffer) and the term-split kernel (half of the
    # exchange area, alternating with 64-bit exchanges) as the oracle
    from tests.test_packer import table_graphs
    prog = synth.shape_program(n_q=22, t=6, F=12, n_out=22, seed=4062)
    assert len(table_graphs(pack_words(prog))) >= 10
    if split == 1:
        assert gpu_exact_matches(prog, rows=8, seed=11) > 0
    s = HipProgramSampler(prog)
    s.set_term_split(split)
    B = 300  # a ragged last tile
    f = synth.random_f(B, s.min_num_f, 0.5, 30 + split)
    U = np.random.default_rng(40 + split).random((B, s.tot_out))
    assert np.array_equal(s.sample_uniforms(f, U), exact.sample_program(prog, f, U))


@pytest.mark.parametrize("which", ["single", "multi", "permuted", "wide"])
def test_packed_rows_equal_byte_rows(which):
    from ksim_amd.sampler import unpack_outputs
    if which == "single":
        prog = synth.shape_program(n_q=6, t=4, F=8, n_out=7, seed=2)
    elif which == "multi":
        prog = synth.shape_program(n_q=6, t=4, F=8, n_out=20, seed=3, n_components=3)
    elif which == "permuted":
        prog = synth.product_program([synth.htkh_program(1), synth.htkh_program(3, noisy_f=True),
                                      synth.shape_program(n_q=5, t=2, F=3, n_out=4, seed=4)])
    else:  # more than 64 outputs: several row words, contiguous columns
        prog = synth.shape_program(n_q=4, t=0, F=10, n_out=70, seed=5)
    s = HipProgramSampler(prog)
    B = 333
    f = synth.random_f(B, max(s.min_num_f, 1), 0.4, 8)
    want = s.sample(f, 99)
    packed = s.sample_packed(f, 99)
    assert packed.shape == (B, (s.num_outputs + 31) // 32)
    assert np.array_equal(unpack_outputs(packed, s.num_outputs), want)


@pytest.mark.parametrize("n_components", [1, 2])
def test_plain_term_split_is_identical(n_components):
    # 1, 2 or 4 waves per 64-shot tile (and the automatic choice): the same
    # bits through every entry point, byte rows, packed rows, fed uniforms
    # and device-drawn noise, on a shot count that ends in a partial tile
    from ksim_amd.noise import NoiseModel
    from tests.test_gpu_noise import model_b
    prog = synth.shape_program(n_q=10, t=14, F=20, n_out=9, seed=41, n_components=n_components)
    s = HipProgramSampler(prog)
    B = 1000
    f = synth.random_f(B, s.min_num_f, 0.4, 3)
    U = np.random.default_rng(5).random((B, s.tot_out))
    probs, T = model_b(num_f=s.min_num_f, n_ch=60, p=0.02, seed=2)
    nm = NoiseModel(probs, T, seed=9)
    got = {}
    for L in (1, 2, 4, 0):
        s.set_term_split(L)
        got[L] = (s.sample(f, 77), s.sample_packed(f, 77), s.sample_uniforms(f, U),
                  s.sample_noisy(nm, B, 78))
    for L in (2, 4, 0):
        for k in range(4):
            assert np.array_equal(got[L][k], got[1][k]), (L, k)
    # and the fed-uniform stream is the oracle's (first 300 shots: the
    # big-integer oracle is slow)
    assert np.array_equal(got[4][2][:300], exact.sample_program(prog, f[:300], U[:300]))


def test_automatic_split_past_the_resident_slots():
    # shot counts whose tiles fill the resident wave slots 1.02x and 1.17x
    # over, where the automatic choice weighs how the waves fall into
    # rounds: still the bits of one wave per tile
    import torch
    from ksim_amd import _lib
    lib = _lib.load()
    prog = synth.shape_program(n_q=30, t=24, F=48, n_out=30, seed=924)
    s = HipProgramSampler(prog)
    for B in (64 * 5200 - 17, 64 * 6000):
        f = torch.from_numpy(synth.random_f(B, 48, 1e-3, 24)).cuda()
        rows = {}
        for L in (1, 4, 0):
            s.set_term_split(L)
            out = torch.full((B, s.num_outputs), 7, dtype=torch.uint8, device="cuda")
            _lib.check(lib.ksim_program_sample_device(s.handle, f.data_ptr(), B, 48, 0, 31,
                                                      None, out.data_ptr(), None))
            s.check()
            rows[L] = out.cpu().numpy()
        s.set_term_split(0)
        assert np.array_equal(rows[0], rows[1]), B
        assert np.array_equal(rows[4], rows[1]), B


def permute_selections(prog, seed):
    """The same program reading its f columns in a shuffled order."""
    rng = np.random.default_rng(seed)
    for c in prog.components:
        c.f_selection = rng.permutation(np.asarray(c.f_selection))
    return prog


@pytest.mark.parametrize("n_components,extra_cols,n_out,permuted",
                         [(1, 0, 6, False), (2, 3, 40, False), (2, 300, 40, False),
                          (1, 0, 6, True), (2, 3, 40, True), (2, 300, 40, True)])
def test_f_row_reads_agree(n_components, extra_cols, n_out, permuted):
    # a 16-byte aligned f with rows of at most 32 * WL bytes is read by the
    # whole wave through LDS; any other f lane by lane: the same bits, on
    # row widths that are not multiples of 4 or 16 (extra_cols), a second
    # component whose f bits start mid-row, shot counts that end inside a
    # tile and inside a 16-byte piece, (300 extra columns) rows too long to
    # stage, and f columns selected in a shuffled order
    import torch
    from ksim_amd import _lib
    prog = synth.shape_program(n_q=8, t=8, F=37, n_out=n_out, seed=51, n_components=n_components)
    if permuted:
        prog = permute_selections(prog, 52)
    s = HipProgramSampler(prog)
    num_f = s.min_num_f + extra_cols
    lib = _lib.load()
    for B in (1, 63, 65, 1000, 1031):
        f = synth.random_f(B, num_f, 0.4, B)
        rows = {}
        for shift in (0, 4, 1):  # byte offset of f from a 256-byte aligned allocation
            buf = torch.zeros(B * num_f + 16, dtype=torch.uint8, device="cuda")
            buf[shift:shift + B * num_f] = torch.from_numpy(f.reshape(-1)).cuda()
            out = torch.full((B, s.num_outputs), 7, dtype=torch.uint8, device="cuda")
            _lib.check(lib.ksim_program_sample_device(s.handle, buf.data_ptr() + shift, B, num_f,
                                                      0, 99, None, out.data_ptr(), None))
            s.check()
            rows[shift] = out.cpu().numpy()
        assert np.array_equal(rows[4], rows[0]), B
        assert np.array_equal(rows[1], rows[0]), B
        assert np.array_equal(rows[0], s.sample(f, 99)), B
    # fed uniforms against the oracle on a short run
    B = 70
    f = synth.random_f(B, num_f, 0.4, 5)
    U = np.random.default_rng(6).random((B, s.tot_out))
    assert np.array_equal(s.sample_uniforms(f, U), exact.sample_program(prog, f, U))


def test_widths_change_and_a_graph_without_terms():
    # the sampler runs the graphs a mask width at a time; here the width
    # steps from 1 to 2 words between outputs 0 and 1 (F = 31), and output 1
    # comes from a graph without terms (probability 0)
    g = synth.shape_program(n_q=6, t=4, F=31, n_out=4, seed=61).components[0].graphs
    graphs = [g[0], g[1], synth.empty_graph(0, 0), g[3], g[4]]
    prog = synth.single_program(graphs, F=31)
    s = HipProgramSampler(prog)
    B = 200
    f = synth.random_f(B, 31, 0.3, 7)
    U = np.random.default_rng(8).random((B, s.tot_out))
    got = s.sample_uniforms(f, U)
    assert np.array_equal(got, exact.sample_program(prog, f, U))
    assert not got[:, 1].any()
    assert np.array_equal(s.sample(f, 5)[:, 1], np.zeros(B, np.uint8))


def _boundary_uniforms(prog, f, rng):
    """Fed uniforms that sit on the rounding boundary of every draw: along
    the oracle's own path, u is fl(p1 / prev) itself (bit 0), the double
    just below it (bit 1), the double just above (bit 0) or a plain draw."""
    from ksim_amd.program import graph_to_dict
    B = f.shape[0]
    cols = []
    for comp in prog.components:
        gds = [graph_to_dict(g) for g in comp.graphs]
        f_sel = f[:, np.asarray(comp.f_selection, dtype=np.int64)].astype(np.int64)
        U = np.zeros((B, len(gds) - 1))
        for s in range(B):
            prev = exact.magnitude(*exact.graph_value(gds[0], f_sel[s]))
            m = []
            for i in range(len(gds) - 1):
                p1 = exact.magnitude(*exact.graph_value(gds[i + 1], exact.params_row(f_sel[s], m, i)))
                with np.errstate(divide="ignore", invalid="ignore"):
                    q = np.float64(p1) / np.float64(prev)
                k = rng.integers(4)
                if not np.isfinite(q) or k == 3:
                    u = rng.random()
                else:
                    u = (q, np.nextafter(q, -np.inf), np.nextafter(q, np.inf))[k]
                U[s, i] = u
                bit = 1 if u < q else 0
                m.append(bit)
                prev = p1 if bit else prev - p1
        cols.append(U)
    return np.concatenate(cols, axis=1)


@pytest.mark.parametrize("seed", [0, 1])
def test_draws_at_the_rounding_boundary(seed):
    # bit = u < fl(p1 / prev): uniforms placed on the rounding boundary of
    # every draw (u = fl(q) and its neighbours) give the reference's bits,
    # one wave per tile and four
    rng = np.random.default_rng(seed)
    progs = [synth.shape_program(n_q=5, t=4, F=6, n_out=6, seed=70 + seed),
             synth.shape_program(n_q=6, t=2, F=40, n_out=30, seed=72 + seed, n_components=2),
             synth.htkh_program(3)]
    for prog in progs:
        s = HipProgramSampler(prog)
        f = synth.random_f(64, max(s.min_num_f, 1), 0.3, 9 + seed)
        U = _boundary_uniforms(prog, f, rng)
        want = exact.sample_program(prog, f, U)
        assert np.array_equal(s.sample_uniforms(f, U), want)
        s.set_term_split(4)
        assert np.array_equal(s.sample_uniforms(f, U), want)


def _kokkos_uniforms_at(seed, shots, n_out):
    """The per-shot xorshift64 draws (exact.kokkos_uniforms) of the given
    shot indices rather than of shots 0..B-1."""
    m64 = (1 << 64) - 1
    out = np.zeros((len(shots), n_out))
    for j, s in enumerate(shots):
        r = exact.kokkos_seed(seed, int(s))
        for i in range(n_out):
            r ^= (r << 13) & m64
            r ^= r >> 7
            r ^= (r << 17) & m64
            out[j, i] = (r >> 11) * (1.0 / 9007199254740992.0)
    return out


@pytest.mark.parametrize("name", ["config2", "config3"])
def test_full_size_runs(name):
    # the benchmark programs at their full 1M shots: every launch
    # decomposition (term split, shot ranges of any alignment, packed rows)
    # gives the same rows, and shots drawn from anywhere in the run equal
    # the oracle's under the same seeded stream
    import torch
    from ksim_amd import _lib
    from ksim_amd.sampler import unpack_outputs
    prog = getattr(synth, f"{name}_program")()
    s = HipProgramSampler(prog)
    B, F, seed = 1_000_000, s.min_num_f, 4242
    f_np = synth.random_f(B, F, 1e-3, 11)
    f = torch.from_numpy(f_np).cuda()
    lib = _lib.load()

    def run(L, cuts=(0, B)):
        s.set_term_split(L)
        out = torch.full((B, s.num_outputs), 7, dtype=torch.uint8, device="cuda")
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            _lib.check(lib.ksim_program_sample_device(s.handle, f[lo:hi].data_ptr(), hi - lo, F,
                                                      lo, seed, None, out[lo:hi].data_ptr(),
                                                      None))
        s.check()
        return out

    ref = run(0)
    for L in (1, 4):
        assert torch.equal(run(L), ref), L
    assert torch.equal(run(0, (0, 1, 333_333, 700_001, B)), ref)
    s.set_term_split(0)
    packed = torch.zeros((B, s.out_words), dtype=torch.int32, device="cuda")
    _lib.check(lib.ksim_program_sample_device_packed(s.handle, f.data_ptr(), B, F, 0, seed, None,
                                                     packed.data_ptr(), None))
    s.check()
    ref_np = ref.cpu().numpy()
    assert np.array_equal(unpack_outputs(packed.cpu().numpy().view(np.uint32), s.num_outputs),
                          ref_np)
    idx = np.sort(np.random.default_rng(3).choice(B, 12, replace=False))
    cols = [_kokkos_uniforms_at(seed + 0x9E37 * k, idx, len(c.graphs) - 1)
            for k, c in enumerate(prog.components)]
    want = exact.sample_program(prog, f_np[idx], np.concatenate(cols, axis=1))
    assert np.array_equal(ref_np[idx], want)
