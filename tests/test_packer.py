"""Host logic without a GPU: the C++ packer (ksim_program_pack) rewritten into
the kernel's record format must evaluate, by the kernel's own arithmetic
(tests/devfmt.py), to the exact oracle's value for every graph — including
the special cases the packer folds (zero masks, projectors, phase pairs that
can vanish, generic floatfactors, G = 0, P = 0, several mask words).  Also
the error behaviour of program creation."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import FlatComponent, FlatProgram, graph_to_dict
from ksim_amd.sampler import lookup_tables, pack_words
from oracle import exact
from tests.devfmt import (CH, GF_TABLE, GH_DWORDS, GH_FLAGS, PH_DWORDS, TH, TH_DWORDS, Packed,
                          chunk_walk)
from tests.golden.golden_io import load_program


@pytest.fixture(scope="module")
def tables():
    return lookup_tables()


def check_program(prog, tables, rows=10, seed=0, all_rows=False):
    pk = Packed(pack_words(prog), tables)
    rng = np.random.default_rng(seed)
    n = 0
    for c, comp in enumerate(prog.components):
        for i, g in enumerate(comp.graphs):
            d = graph_to_dict(g)
            P = g.np_params.shape[2] if g.np_params.ndim == 3 else 0
            if all_rows and P <= 8:
                it = [np.array([(m >> j) & 1 for j in range(P)]) for m in range(2 ** P)]
            else:
                it = [rng.integers(0, 2, size=P) for _ in range(rows)]
            for bits in it:
                want = exact.graph_value(d, bits)
                got = exact.canonical(*pk.eval_graph(c, i, bits))
                assert got == want, (c, i, list(bits))
                n += 1
    return n


def test_tables_are_the_closed_forms(tables):
    tab, pit, W, PM, D, G = tables
    lam, up, um = (1, 1, 0, 0), (1, 1, 0, 1), (-1, 1, 0, 1)
    for n in (-W, -40, -23, -22, -3, 0, 5, 22, 23, 47, W):
        for r in range(4):
            for s in range(8):
                x = exact.UNIT[s]
                for _ in range(r):
                    x = exact.zmul(x, lam)
                for _ in range(abs(n)):
                    x = exact.zmul(x, up if n > 0 else um)
                gi = ((n + W) * 4 + r) * 8 + s
                digits = tab[:, gi, :].astype(np.int64)
                assert np.all(np.abs(digits) < 2 ** 30)
                got = tuple(sum(int(digits[d, k]) << (G * d) for d in range(D)) for k in range(4))
                assert got == x, (n, r, s)
                if abs(n) <= 22:  # central rows are a single digit
                    assert not digits[1:].any()
    # pi^a pibar^b
    pi, pib = (1, 1, 0, -1), (1, -1, 0, 1)  # 1 +- i sqrt2, i sqrt2 = w - conj w
    for a in (0, 1, 7, PM - 1):
        for b in (0, 3, PM - 1):
            x = (1, 0, 0, 0)
            for _ in range(a):
                x = exact.zmul(x, pi)
            for _ in range(b):
                x = exact.zmul(x, pib)
            assert tuple(int(v) for v in pit[a * PM + b]) == x
    # every entry lies in Z[i sqrt2], (P0, P1, 0, -P1): the kernels multiply
    # by P0 and P1 only
    pit = np.asarray(pit).reshape(PM * PM, 4).astype(np.int64)
    assert not pit[:, 2].any() and np.array_equal(pit[:, 3], -pit[:, 1])


def test_golden_programs(tables, golden_files):
    for fn in golden_files:
        check_program(load_program(np.load(fn)), tables, all_rows=True)


@pytest.mark.parametrize("seed", range(4))
def test_shape_programs(tables, seed):
    prog = synth.shape_program(n_q=8, t=8, F=6, n_out=5, seed=seed, generic_ff=0.5,
                               proj_frac=0.2)
    assert check_program(prog, tables, rows=6, seed=seed) > 0


def test_wide_params(tables):
    # P up to 100: four 32-bit mask words
    prog = synth.shape_program(n_q=6, t=4, F=90, n_out=10, seed=7)
    check_program(prog, tables, rows=4)


@pytest.mark.parametrize("F,n_out", [(250, 7), (300, 3), (1000, 2)])
def test_components_wider_than_the_register_words(tables, F, n_out):
    # F + n_out past 256 parameters: masks of 9 to 32 words, the layout the
    # kernels that keep parameter words in LDS walk
    prog = synth.shape_program(n_q=6, t=4, F=F, n_out=n_out, seed=F)
    assert check_program(prog, tables, rows=3, seed=1) > 0
    st = {}
    w = pack_words(prog, st)
    assert int(w[4]) == (F + n_out + 31) // 32  # PH_WMAX


def test_rejects_more_than_4096_parameters():
    from ksim_amd import _lib
    import ctypes as C
    from ksim_amd.sampler import ProgramArgs
    prog = synth.shape_program(n_q=2, t=0, F=4090, n_out=7, seed=1)
    pa = ProgramArgs(prog)
    n = C.c_int64()
    with pytest.raises(ValueError, match="4096"):
        _lib.check(_lib.load().ksim_program_pack(*pa.args(), None, 0, C.byref(n), None))


def test_all_phase_pairs(tables):
    # one term per (alpha, beta), masks on both, one, or neither parity
    G = 64 * 4
    P = 3
    g = synth.empty_graph(G, P, Tph=1)
    t = 0
    for a in range(8):
        for b in range(8):
            for kind in range(4):
                g.ph_alpha[t, 0] = a
                g.ph_beta[t, 0] = b
                g.ph_counts[t] = 1
                if kind in (0, 1):
                    g.ph_alpha_p[t, 0, 0] = 1
                if kind in (0, 2):
                    g.ph_beta_p[t, 0, 1] = 1
                if kind == 3:
                    g.ph_alpha_p[t, 0, 2] = 1
                    g.ph_beta_p[t, 0, 2] = 1
                g.pf_power2[t] = -(t % 5)
                t += 1
    prog = synth.single_program([synth.constant_graph(2), g], F=2)
    check_program(prog, tables, all_rows=True)


def test_node_phases_and_projectors(tables):
    G, P = 16, 4
    g = synth.empty_graph(G, P, Tnp=3, Thp=2, Tpp=2)
    rng = np.random.default_rng(5)
    g.np_phases[:] = rng.integers(0, 8, size=(G, 3))
    g.np_params[:] = rng.integers(0, 2, size=(G, 3, P))
    g.np_counts[:] = rng.integers(0, 4, size=G)
    g.hp_coeffs[:] = rng.integers(0, 8, size=(G, 2))
    g.hp_params[:] = rng.integers(0, 2, size=(G, 2, P))
    g.pp_psi_c[:] = rng.integers(0, 2, size=(G, 2))
    g.pp_phi_c[:] = rng.integers(0, 2, size=(G, 2))
    g.pp_psi_p[:] = rng.integers(0, 2, size=(G, 2, P))
    g.pp_phi_p[:] = rng.integers(0, 2, size=(G, 2, P))
    g.pp_phi_p[3] = g.pp_psi_p[3]          # shared pi-product parity
    g.pf_phase_idx[:] = rng.integers(0, 8, size=G)
    g.pf_floatfactor[:] = rng.integers(-3, 4, size=(G, 4))
    g.pf_power2[:] = rng.integers(-6, 3, size=G)
    prog = synth.single_program([synth.constant_graph(3), g], F=3)
    check_program(prog, tables, all_rows=True)


def sign_record_program(n_sign: int, seed: int, P: int = 40):
    """One graph whose terms hold n_sign node phases with phi odd or 2 mod 4
    (sign records), a few half-pi factors with odd coefficients (LIN records
    with their own delta) and phase pairs, after a constant graph 0."""
    G, Thp, Tph = 2, 3, 3
    rng = np.random.default_rng(seed)
    g = synth.empty_graph(G, P, Tnp=n_sign, Thp=Thp, Tph=Tph)
    g.np_phases[:] = rng.choice([1, 2, 3, 5, 6, 7], size=(G, n_sign))
    g.np_params[:] = rng.random((G, n_sign, P)) < 0.2
    g.np_counts[:] = n_sign
    g.hp_coeffs[:] = rng.choice([1, 3, 5, 7], size=(G, Thp))
    g.hp_params[:] = rng.random((G, Thp, P)) < 0.2
    g.ph_alpha[:] = rng.integers(0, 8, size=(G, Tph))
    g.ph_beta[:] = rng.integers(0, 8, size=(G, Tph))
    g.ph_alpha_p[:] = rng.random((G, Tph, P)) < 0.2
    g.ph_beta_p[:] = rng.random((G, Tph, P)) < 0.2
    g.ph_counts[:] = Tph
    g.pf_power2[:] = rng.integers(-4, 1, size=G)
    return synth.single_program([synth.constant_graph(P - 1), g], F=P - 1)


def chunk_shapes(words):
    """(n_sign, n_lin, n_pair) of every record chunk of the program's last graph."""
    w = [int(x) for x in words]
    off = w[PH_DWORDS]
    ng = w[off + CH["NGRAPHS"]]
    gh = w[w[off + CH["GRAPH"]] + (ng - 1) * GH_DWORDS:][:GH_DWORDS]
    W32, th, out = gh[1], gh[3], []
    table = bool(gh[GH_FLAGS] & GF_TABLE)
    for _ in range(gh[0]):
        h = w[th:th + TH_DWORDS]
        rp = th + TH_DWORDS + h[TH["NPROJ"]] * (W32 + 1) + h[TH["NPPX"]] * (2 * W32 + 2)
        for _ in range(h[TH["NCHUNK"]]):
            npm, nl, npr, _, g, rp = chunk_walk(w, rp, W32, table)
            if table:  # sign records: the forms marked in S+, S-, N+, N- of each group
                gs = [w[g + k * (8 + 128 * W32):][:4] for k in range(npm)]
                npm = sum(bin(a | b | c | d).count("1") for a, b, c, d in gs)
            out.append((npm, nl, npr))
        th += TH_DWORDS + h[TH["SIZE"]]
    return out


def table_graphs(words):
    """Indices of the graphs of component 0 that carry parity tables."""
    w = [int(x) for x in words]
    off = w[PH_DWORDS]
    gb = w[off + CH["GRAPH"]]
    return [i for i in range(w[off + CH["NGRAPHS"]])
            if w[gb + i * GH_DWORDS + GH_FLAGS] & GF_TABLE]


def test_parity_tables_evaluate_as_the_oracle(tables):
    # a program with graphs that need 128-bit sums (it runs the 128-bit
    # kernels) carries the sign groups of its multi-term graphs as nibble
    # parity tables: the XOR of one entry per nibble of the parameter words
    # is the group's parity word, so every graph evaluates to the oracle's
    # exact value (tables 16-byte aligned, checked by the interpreter)
    prog = synth.shape_program(n_q=22, t=6, F=12, n_out=22, seed=4062)
    words = pack_words(prog)
    assert len(table_graphs(words)) >= 10
    shapes = chunk_shapes(words)
    assert sum(m for m, _, _ in shapes) > 0
    assert check_program(prog, tables, rows=6, seed=7) > 0


def test_parity_tables_only_where_the_kernels_read_them(tables):
    # the 64-bit kernels hold a 1 KiB buffer per wave: an all-64-bit program
    # carries tables in its graphs of at most 2 mask words (and they
    # evaluate as the oracle), none in wider graphs; a program whose
    # component needs more than 8 lane words keeps every mask
    narrow = synth.shape_program(n_q=20, t=12, F=24, n_out=8, seed=12)
    words = pack_words(narrow)
    w = [int(x) for x in words]
    gb = w[w[PH_DWORDS] + CH["GRAPH"]]
    tg = table_graphs(words)
    assert tg and all(w[gb + i * GH_DWORDS + 1] <= 2 for i in tg)
    assert check_program(narrow, tables, rows=4, seed=3) > 0
    assert table_graphs(pack_words(synth.shape_program(n_q=20, t=12, F=70, n_out=8, seed=12))) == []
    a = synth.shape_program(n_q=22, t=6, F=12, n_out=22, seed=4062)
    assert table_graphs(pack_words(a))
    # the same component beside one of 257 parameters (9 lane words): the
    # program runs the kernels that keep parameter words in LDS
    b = synth.shape_program(n_q=6, t=4, F=250, n_out=7, seed=250)
    ca, cb = a.components[0], b.components[0]
    n_a = len(ca.graphs) - 1
    both = FlatProgram(
        components=[ca, FlatComponent(tuple(range(n_a, n_a + len(cb.graphs) - 1)),
                                      np.asarray(cb.f_selection), cb.graphs)],
        output_order=np.arange(n_a + len(cb.graphs) - 1),
        num_outputs=n_a + len(cb.graphs) - 1, num_detectors=0)
    assert table_graphs(pack_words(both)) == []


@pytest.mark.parametrize("n_sign", [1, 31, 32, 33, 63, 64, 65, 97])
def test_sign_record_groups(tables, n_sign):
    # node phases pack into sign-record groups of up to 32 (one funnel-shift
    # word each), chunks of at most 63 records mixing them with LIN and PAIR
    # records; every group size from 1 to 32 and several groups per chunk
    # evaluate to the oracle's exact values
    prog = sign_record_program(n_sign, seed=n_sign)
    shapes = chunk_shapes(pack_words(prog))
    # (one-sided phase pairs may add sign records of their own)
    assert sum(m for m, _, _ in shapes) >= 2 * n_sign
    assert all(m + l + p <= 63 for m, l, p in shapes)
    assert any(m > 0 and l + p > 0 for m, l, p in shapes) or n_sign >= 63
    check_program(prog, tables, rows=24, seed=n_sign)


def test_empty_graph_and_zero_terms(tables):
    g0 = synth.empty_graph(0, 0)
    g1 = synth.empty_graph(2, 1, Tnp=1)
    g1.np_phases[:, 0] = 4        # constant zero factor when the mask is empty
    prog = synth.single_program([g0, g1], F=0)
    check_program(prog, tables, all_rows=True)


@pytest.mark.parametrize("prog_fn", [
    lambda: synth.rz_program(0.7),
    lambda: synth.phase_program(-1.3),
    lambda: synth.with_approx_factors(synth.shape_program(n_q=6, t=6, F=8, n_out=6, seed=3), 4),
    lambda: synth.with_approx_factors(synth.wide_exponent_program(seed=5), 5, frac=1.0),
    lambda: synth.with_approx_factors(
        synth.shape_program(n_q=6, t=4, F=300, n_out=3, seed=6, n_components=2), 6),
])
def test_approximate_floatfactors(tables, prog_fn):
    # graphs with complex64 factors: the packed terms (kept in the
    # reference's order, factors in the graph's table) give, by the
    # kernels' float path, the oracle's float64 magnitude bit for bit
    prog = prog_fn()
    pk = Packed(pack_words(prog), tables)
    rng = np.random.default_rng(2)
    n = 0
    for c, comp in enumerate(prog.components):
        for i, g in enumerate(comp.graphs):
            d = graph_to_dict(g)
            P = g.np_params.shape[2] if g.np_params.ndim == 3 else 0
            for _ in range(4):
                bits = rng.integers(0, 2, size=P)
                assert pk.magnitude(c, i, bits) == exact.graph_magnitude(d, bits), (c, i)
                n += 1
    assert n > 0


def test_rejects_wrong_param_count():
    g0 = synth.constant_graph(0)
    g1 = synth.empty_graph(1, 3, Tnp=1)   # graph 1 must have P = F + 1 = 1
    prog = synth.single_program([g0, g1], F=0)
    with pytest.raises(ValueError, match="expected F\\+i"):
        pack_words(prog)


def test_rejects_bad_counts():
    prog = synth.htkh_program(1)
    prog.components[0].graphs[1].np_counts[0] = 5
    with pytest.raises(ValueError, match="np_counts"):
        pack_words(prog)


def test_multi_component_layout(tables):
    prog = synth.product_program([synth.htkh_program(1), synth.ghz_program(3)
                                  if hasattr(synth, "ghz_program") else
                                  synth.single_program(synth.ghz_graphs(3)),
                                  synth.htkh_program(3, noisy_f=True)])
    check_program(prog, tables, all_rows=True)


@pytest.mark.parametrize("t", [32, 40])
def test_high_t_count(tables, t):
    # many phase pairs per term: pi exponents up to ~t/8, wide term counts
    prog = synth.shape_program(n_q=6, t=t, F=2 * t, n_out=2, seed=t)
    pk = Packed(pack_words(prog), tables)
    rng = np.random.default_rng(t)
    comp = prog.components[0]
    for i in (0, 2):
        g = comp.graphs[i]
        d = graph_to_dict(g)
        bits = rng.integers(0, 2, size=g.np_params.shape[2])
        assert exact.canonical(*pk.eval_graph(0, i, bits)) == exact.graph_value(d, bits)


def test_wide_rows(tables):
    # n' = N + 2 (V >> 2) far beyond the single-digit rows: such terms add as
    # several base-2^31 digits
    prog = synth.wide_exponent_program(seed=5)
    st = {}
    pk = Packed(pack_words(prog, st), tables)
    assert st["wide_terms"] > 0 and st["n2_bound"] > 40
    rng = np.random.default_rng(5)
    comp = prog.components[0]
    for i, g in enumerate(comp.graphs):
        d = graph_to_dict(g)
        for _ in range(4):
            bits = rng.integers(0, 2, size=g.np_params.shape[2])
            assert exact.canonical(*pk.eval_graph(0, i, bits)) == exact.graph_value(d, bits)
    assert pk.max_n2 > 46  # the third digit plane is used too


def test_rejects_beyond_the_widest_row():
    prog = synth.wide_exponent_program(n_out=1, G=2, T=300, seed=1)
    with pytest.raises(ValueError):
        pack_words(prog)


def test_pack_stats_ops_match_the_formula():
    from ksim_amd.sampler import algorithmic_ops
    for prog in (synth.shape_program(n_q=8, t=8, F=6, n_out=5, seed=1),
                 synth.config1_program(), synth.htkh_program(3, noisy_f=True)):
        st = {}
        w = pack_words(prog, st)
        assert st["words"] == len(w)
        assert st["ops_per_shot"] == algorithmic_ops(prog)
        assert 0 < st["sum_bits"] <= 128


def test_graph_widths_follow_their_position():
    # graph i is ceil((F + i) / 32) mask words wide, and so is a graph
    # without terms (whose own P is not checked): the sampler keeps graph
    # i's trying bit, parameter F + i - 1, in the graph's last mask word
    F = 31
    graphs = [synth.constant_graph(F), synth.constant_graph(F + 1), synth.empty_graph(0, 0),
              synth.constant_graph(F + 3)]
    w = pack_words(synth.single_program(graphs, F=F))
    c0 = int(w[8])
    gh = int(w[c0 + 7])
    assert [int(w[gh + 16 * i + 1]) for i in range(4)] == [1, 1, 2, 2]
