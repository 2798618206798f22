"""FlatProgram generators: closed-form programs and shape-matched workloads.

``ksim.compile`` needs stim and pyzx_param, which are not available, so every
workload here is a FlatProgram built directly in the IR of
``src/python/ksim/program.py:29-53``.  Two kinds:

* **closed-form programs** whose output distribution is known exactly
  (used for statistical tests): H.T^k.H (P(1) = sin^2(k pi/8), the known
  answer of tests/test_nonclifford.py:87-121 in the reference), a noisy
  variant with one f-bit, GHZ-like parity programs, products of independent
  components, and an exact value-preserving term split that raises the term
  count without changing any amplitude.
* **shape-matched programs** for throughput: term counts, family widths,
  mask densities and phase distributions follow a stated shape model
  (:func:`shape_program`); the values are not probabilities.

The stim text of each benchmark circuit is produced by the ``*_circuit``
functions so the provenance of a shape is recorded.
"""
from __future__ import annotations

import math

import numpy as np

from .program import FlatComponent, FlatProgram, FlatScalarGraph


# --------------------------------------------------------------------------
# building blocks
# --------------------------------------------------------------------------
def empty_graph(G: int, P: int, Tnp: int = 0, Thp: int = 0, Tpp: int = 0,
                Tph: int = 0) -> FlatScalarGraph:
    """All-padding graph: every term is 1 * 2^0 (ff = 1, no live factors)."""
    u8 = np.uint8
    ff = np.zeros((G, 4), np.int32)
    ff[:, 0] = 1
    return FlatScalarGraph(
        np_phases=np.zeros((G, Tnp), u8), np_params=np.zeros((G, Tnp, P), u8),
        np_counts=np.zeros(G, np.int32),
        hp_coeffs=np.zeros((G, Thp), u8), hp_params=np.zeros((G, Thp, P), u8),
        pp_psi_c=np.zeros((G, Tpp), u8), pp_psi_p=np.zeros((G, Tpp, P), u8),
        pp_phi_c=np.zeros((G, Tpp), u8), pp_phi_p=np.zeros((G, Tpp, P), u8),
        ph_alpha=np.zeros((G, Tph), u8), ph_alpha_p=np.zeros((G, Tph, P), u8),
        ph_beta=np.zeros((G, Tph), u8), ph_beta_p=np.zeros((G, Tph, P), u8),
        ph_counts=np.zeros(G, np.int32),
        pf_phase_idx=np.zeros(G, u8), pf_floatfactor=ff,
        pf_power2=np.zeros(G, np.int32),
        pf_approx=np.ones(G, np.complex64), pf_has_approx=False)


def constant_graph(P: int, power2: int = 0) -> FlatScalarGraph:
    """One term equal to 2^power2 for every parameter value."""
    g = empty_graph(1, P)
    g.pf_power2[0] = power2
    return g


def single_program(graphs: list, F: int = 0, num_detectors: int | None = None) -> FlatProgram:
    n_out = len(graphs) - 1
    comp = FlatComponent(tuple(range(n_out)), np.arange(F, dtype=np.int32), graphs)
    return FlatProgram([comp], np.arange(n_out, dtype=np.int32), n_out,
                       n_out if num_detectors is None else num_detectors)


def htkh_graphs(k: int, noisy_f: bool = False) -> list:
    """H.T^k.H on one qubit, measured: graph0 = 1, graph1 = sin^2(k pi/8).

    graph1 has one term with node factors (1 + w^{4 par + k})(1 + w^{4 par - k})
    over the trying bit and power2 = -2; with par = 1 this is
    |1 - w^k|^2 / 4 = sin^2(k pi / 8).  With ``noisy_f`` one f-bit enters both
    masks, so f = 1 flips the parity and P(1 | f=1) = cos^2(k pi/8).
    """
    P = 2 if noisy_f else 1
    F = 1 if noisy_f else 0
    g0 = constant_graph(F)
    g1 = empty_graph(1, P, Tnp=2)
    g1.np_phases[0] = [k % 8, (8 - k) % 8]
    g1.np_params[0, :, :] = 1
    g1.np_counts[0] = 2
    g1.pf_power2[0] = -2
    return [g0, g1]


def htkh_program(k: int, noisy_f: bool = False) -> FlatProgram:
    return single_program(htkh_graphs(k, noisy_f), F=1 if noisy_f else 0)


def htht_h_graphs() -> list:
    """H.T.H.T.H on one qubit, measured: P(1) = 1/4 exactly (the reference's
    class-boundary case, docs/benchmarks.md:124-125).

    graph0 = 1; graph1 has two terms, 2^-2 and 2^-2 (1 + w^{4 par}) over the
    trying bit, so it is 3/4 at par = 0 and 1/4 at par = 1."""
    g0 = constant_graph(0)
    g1 = empty_graph(2, 1, Tnp=1)
    g1.np_phases[1, 0] = 0
    g1.np_params[1, 0, 0] = 1
    g1.np_counts[1] = 1
    g1.pf_power2[:] = -2
    return [g0, g1]


def htht_h_program() -> FlatProgram:
    return single_program(htht_h_graphs(), F=0)


def rz_graphs(theta: float) -> list:
    """H.RZ(theta).H on one qubit, measured: P(1) = sin^2(theta / 2), a
    rotation outside Clifford+T, so its phases are approximate complex64
    floatfactors (the reference's compiler moves any phase whose
    denominator is not 1, 2 or 4 there, compile/pipeline.py:362-372).

    graph1 (param: the trying bit m) has three terms: 1/2,
    (1/4) (-1)^m e^{i theta} and (1/4) (-1)^m e^{-i theta}, the (-1)^m a
    half-pi record of coefficient 4; at m = 1 the sum is
    (1 - cos theta) / 2."""
    g0 = constant_graph(0)
    g1 = empty_graph(3, 1, Thp=1)
    g1.pf_power2[:] = [-1, -2, -2]
    g1.hp_coeffs[1:, 0] = 4
    g1.hp_params[1:, 0, 0] = 1
    g1.pf_approx = np.array([1.0, np.exp(1j * theta), np.exp(-1j * theta)], np.complex64)
    g1.pf_has_approx = True
    return [g0, g1]


def rz_program(theta: float) -> FlatProgram:
    return single_program(rz_graphs(theta), F=0)


def phase_graphs(theta: float) -> list:
    """One output with P(1) = |1 + w e^{i theta}| / 2 = |cos((theta + pi/4)
    / 2)| (graph0 = 1): a complex sum whose magnitude depends on the sign of
    theta and on w, so a conjugated factor, a dropped w or a swapped real
    and imaginary part changes it.  graph1 has two terms, 1/2 and
    (1/2) w^5 (-1)^m e^{i theta} (w^5 (-1)^m = w at m = 1, the (-1)^m a
    half-pi record of coefficient 4 on the trying bit)."""
    g0 = constant_graph(0)
    g1 = empty_graph(2, 1, Thp=1)
    g1.pf_power2[:] = [-1, -1]
    g1.hp_coeffs[1, 0] = 4
    g1.hp_params[1, 0, 0] = 1
    g1.pf_phase_idx[1] = 5
    g1.pf_approx = np.array([1.0, np.exp(1j * theta)], np.complex64)
    g1.pf_has_approx = True
    return [g0, g1]


def phase_program(theta: float) -> FlatProgram:
    return single_program(phase_graphs(theta), F=0)


def with_approx_factors(prog: FlatProgram, seed: int, frac: float = 0.5) -> FlatProgram:
    """``prog`` with random unit complex64 factors on the terms of about
    ``frac`` of its graphs (pf_has_approx set on those): the shape of a
    circuit with arbitrary-angle rotations.  Values are not probabilities."""
    rng = np.random.default_rng(seed)
    comps = []
    for c in prog.components:
        gs = []
        for g in c.graphs:
            g = _copy_graph(g)
            if rng.random() < frac:
                G = g.pf_phase_idx.shape[0]
                g.pf_approx = np.exp(1j * rng.uniform(0, 2 * np.pi, size=G)).astype(np.complex64)
                g.pf_has_approx = True
            gs.append(g)
        comps.append(FlatComponent(c.output_indices, c.f_selection, gs))
    return FlatProgram(comps, prog.output_order, prog.num_outputs, prog.num_detectors)


def ghz_graphs(n: int) -> list:
    """n outputs that are all equal, each value 0/1 with probability 1/2.

    graph i+1 (params m_0..m_{i-1}, trying) = 1/2 * prod_{j<i} [m_j = trying],
    each indicator a node factor (1 + w^{4 par(m_j + trying)}) = 2 or 0.
    """
    graphs = [constant_graph(0)]
    graphs.append(constant_graph(1, power2=-1))
    for i in range(1, n):
        P = i + 1
        g = empty_graph(1, P, Tnp=i)
        for j in range(i):
            g.np_params[0, j, j] = 1
            g.np_params[0, j, i] = 1
        g.np_counts[0] = i
        g.pf_power2[0] = -1 - i
        graphs.append(g)
    return graphs


def product_program(programs: list) -> FlatProgram:
    """Independent components side by side; f-bits and outputs concatenated,
    outputs interleaved through a non-trivial ``output_order``."""
    comps = []
    f_off = 0
    out_off = 0
    for p in programs:
        for c in p.components:
            F = len(c.f_selection)
            n = len(c.graphs) - 1
            comps.append(FlatComponent(tuple(range(out_off, out_off + n)),
                                       np.asarray(c.f_selection, np.int32) + f_off,
                                       list(c.graphs)))
            f_off += F
            out_off += n
    order = np.arange(out_off, dtype=np.int32)[::-1].copy()
    return FlatProgram(comps, order, out_off, out_off)


def _copy_graph(g: FlatScalarGraph) -> FlatScalarGraph:
    return FlatScalarGraph(**{k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
                              for k, v in g.__dict__.items()})


def split_terms(g: FlatScalarGraph, rounds: int, rng: np.random.Generator) -> FlatScalarGraph:
    """Exact value-preserving term splitting.

    A term containing a live node factor (1 + w^{4 par + phi}) equals
    R + w^phi (-1)^par R, where R is the term without that factor; the
    (-1)^par is a half-pi record with coefficient 4 on the same mask.  Each
    round splits one random eligible term, so G grows by one and the graph's
    value is unchanged for every parameter vector.
    """
    for _ in range(rounds):
        cand = [t for t in range(g.pf_phase_idx.shape[0]) if g.np_counts[t] > 0]
        if not cand:
            break
        t = int(rng.choice(cand))
        j = int(rng.integers(0, g.np_counts[t]))
        G, Tnp = g.np_phases.shape
        P = g.np_params.shape[2]
        phi = int(g.np_phases[t, j])
        mask = g.np_params[t, j].copy()
        # R: drop factor j (swap with the last live factor)
        last = int(g.np_counts[t]) - 1
        for arr in (g.np_phases, g.np_params):
            arr[t, j], arr[t, last] = arr[t, last].copy(), arr[t, j].copy()
        g.np_counts[t] = last
        # append a copy of term t with w^phi and a (-1)^par half-pi record
        def app(a, row):
            return np.concatenate([a, row[None]], axis=0)
        Thp = g.hp_coeffs.shape[1]
        hp_c = np.concatenate([g.hp_coeffs, np.zeros((G, 1), np.uint8)], axis=1)
        hp_p = np.concatenate([g.hp_params, np.zeros((G, 1, P), np.uint8)], axis=1)
        new_hc = hp_c[t].copy(); new_hc[Thp] = 4
        new_hp = hp_p[t].copy(); new_hp[Thp] = mask
        g.hp_coeffs = app(hp_c, new_hc)
        g.hp_params = app(hp_p, new_hp)
        for name in ("np_phases", "np_params", "np_counts", "pp_psi_c", "pp_psi_p",
                     "pp_phi_c", "pp_phi_p", "ph_alpha", "ph_alpha_p", "ph_beta",
                     "ph_beta_p", "ph_counts", "pf_floatfactor", "pf_power2",
                     "pf_approx"):
            a = getattr(g, name)
            setattr(g, name, app(a, a[t].copy()))
        pf = np.concatenate([g.pf_phase_idx, [(int(g.pf_phase_idx[t]) + phi) % 8]]).astype(np.uint8)
        g.pf_phase_idx = pf
    return g


def split_program(prog: FlatProgram, rounds: int, seed: int = 0) -> FlatProgram:
    rng = np.random.default_rng(seed)
    comps = []
    for c in prog.components:
        gs = [split_terms(_copy_graph(g), rounds, rng) for g in c.graphs]
        comps.append(FlatComponent(c.output_indices, c.f_selection, gs))
    return FlatProgram(comps, prog.output_order, prog.num_outputs, prog.num_detectors)


# --------------------------------------------------------------------------
# shape model
# --------------------------------------------------------------------------
def terms_for_t(t: int, rng: np.random.Generator, jitter: bool = True) -> int:
    """Stabilizer-rank term count 2^{0.228 t} (cat5 scaling), jittered by
    2^{U(-1,1)} per graph; 1 for a Clifford graph."""
    if t <= 0:
        return 1
    g = 2.0 ** (0.228 * t)
    if jitter:
        g *= 2.0 ** rng.uniform(-1.0, 1.0)
    return max(1, int(round(g)))


def _bernoulli_u8(rng: np.random.Generator, shape, p: float) -> np.ndarray:
    """{0,1} uint8 array with P(1) = p, drawn one byte per element."""
    thr = int(round(p * 256))
    return (rng.integers(0, 256, size=shape, dtype=np.uint8) < thr).astype(np.uint8)


_PURE_FF = np.array([[1, 0, 0, 0], [-1, 0, 0, 0], [0, 1, 0, 1], [2, 0, 0, 0],
                     [1, 1, 0, 0], [0, 0, 1, 0], [1, 0, 1, 0]], np.int32)


def shape_graph(G: int, P: int, n_q: int, t: int, rng: np.random.Generator,
                generic_ff: float = 0.1, proj_frac: float = 0.03,
                density: float = 0.125) -> FlatScalarGraph:
    """One graph under the shape model.

    T_np = n_q + t, T_hp = T_pp = ceil(n_q/4), T_ph = floor(t/2) + 1;
    live counts uniform in [T/2, T]; mask bits Bernoulli(density) over P;
    node phases odd or +-2 (never a zero factor) except a ``proj_frac``
    fraction in {0, 4} (parity projectors); half-pi coefficients in {2, 6};
    phase-pair phases uniform mod 8; prefactor phase uniform, floatfactor
    a small Z[omega] element (a ``generic_ff`` fraction drawn from
    [-2, 2]^4), power2 = a per-graph base uniform in [-n_q, 0] plus a per-term
    jitter in [-3, 0] (the terms of one decomposition share a scale).
    """
    Tnp = n_q + t
    Thp = Tpp = max(1, math.ceil(n_q / 4))
    Tph = t // 2 + 1
    g = empty_graph(G, P, Tnp, Thp, Tpp, Tph)
    if P > 0:
        for name, T in (("np_params", Tnp), ("hp_params", Thp), ("pp_psi_p", Tpp),
                        ("pp_phi_p", Tpp), ("ph_alpha_p", Tph), ("ph_beta_p", Tph)):
            getattr(g, name)[:] = _bernoulli_u8(rng, (G, T, P), density)
    live_ph = np.array([1, 2, 3, 5, 6, 7], np.uint8)
    ph = rng.choice(live_ph, size=(G, Tnp))
    proj = rng.random((G, Tnp)) < proj_frac
    ph[proj] = rng.choice(np.array([0, 4], np.uint8), size=int(proj.sum()))
    g.np_phases[:] = ph
    g.np_counts[:] = rng.integers(Tnp // 2, Tnp + 1, size=G)
    g.hp_coeffs[:] = rng.choice(np.array([2, 6], np.uint8), size=(G, Thp))
    g.pp_psi_c[:] = rng.integers(0, 2, size=(G, Tpp))
    g.pp_phi_c[:] = rng.integers(0, 2, size=(G, Tpp))
    g.ph_alpha[:] = rng.integers(0, 8, size=(G, Tph))
    g.ph_beta[:] = rng.integers(0, 8, size=(G, Tph))
    g.ph_counts[:] = rng.integers(Tph // 2, Tph + 1, size=G)
    g.pf_phase_idx[:] = rng.integers(0, 8, size=G)
    ff = _PURE_FF[rng.integers(0, len(_PURE_FF), size=G)].copy()
    gen = rng.random(G) < generic_ff
    if gen.any():
        r = rng.integers(-2, 3, size=(int(gen.sum()), 4)).astype(np.int32)
        r[np.all(r == 0, axis=1), 0] = 1
        ff[gen] = r
    g.pf_floatfactor[:] = ff
    g.pf_power2[:] = int(rng.integers(-n_q, 1)) + rng.integers(-3, 1, size=G)
    return g


def shape_program(n_q: int, t: int, F: int, n_out: int, seed: int,
                  n_components: int = 1, **kw) -> FlatProgram:
    """A shape-matched program: ``n_components`` components sharing the
    outputs and f-bits evenly; graph i of a component has P = F_c + i params
    (i = 0 is the normaliser with P = F_c)."""
    rng = np.random.default_rng(seed)
    comps = []
    f_base = 0
    out_base = 0
    for c in range(n_components):
        Fc = F // n_components + (1 if c < F % n_components else 0)
        nc = n_out // n_components + (1 if c < n_out % n_components else 0)
        graphs = []
        for i in range(nc + 1):
            P = Fc + i
            graphs.append(shape_graph(terms_for_t(t, rng), P, n_q, t, rng, **kw))
        comps.append(FlatComponent(tuple(range(out_base, out_base + nc)),
                                   np.arange(f_base, f_base + Fc, dtype=np.int32), graphs))
        f_base += Fc
        out_base += nc
    order = rng.permutation(n_out).astype(np.int32)
    return FlatProgram(comps, order, n_out, n_out)


def wide_exponent_program(n_out: int = 3, G: int = 6, T: int = 110, F: int = 6,
                          seed: int = 0) -> FlatProgram:
    """Terms whose (1 + sqrt2) exponent n' = N + 2 floor(V/4) reaches about
    +-T/2, far beyond the single-digit table rows.  Every node factor
    1 + w^k with odd k carries one power of (1 + w); k in {1, 7} adds no
    (1 + sqrt2) power and k in {3, 5} adds (1 + sqrt2)^-1, so a term whose
    ~T live phases come from one pair sits near n' = +T/2 or -T/2.  Sparse
    masks (a parity flip turns k into k + 4) spread the lanes around it, and
    power2 rescales each term to magnitude ~1.  The terms of one graph share
    a pair (odd graphs {3, 5}), so one graph's terms need similar scales."""
    rng = np.random.default_rng(seed)
    graphs = []
    pairs = (np.array([1, 7], np.uint8), np.array([3, 5], np.uint8))
    for i in range(n_out + 1):
        P = F + i
        g = empty_graph(G, P, Tnp=T)
        g.np_phases[:] = rng.choice(pairs[i % 2], size=(G, T))
        if P:
            g.np_params[:] = _bernoulli_u8(rng, (G, T, P), 0.05)
        g.np_counts[:] = rng.integers(T - 10, T + 1, size=G)
        g.pf_phase_idx[:] = rng.integers(0, 8, size=G)
        # |1 + w^k| = 1.848 (k = 1, 7) or 0.765 (k = 3, 5)
        mag = math.log2(1.8478) if i % 2 == 0 else math.log2(0.7654)
        g.pf_power2[:] = (-np.round(g.np_counts * mag)).astype(np.int32) \
            + rng.integers(-2, 1, size=G).astype(np.int32)
        graphs.append(g)
    return single_program(graphs, F=F)


def random_f(B: int, F: int, p: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.random((B, F)) < p).astype(np.uint8)


# --------------------------------------------------------------------------
# the five benchmark workloads
# --------------------------------------------------------------------------
def config1_program(seed: int = 1) -> FlatProgram:
    """5-qubit random Clifford+T (40 {H,S,CX} + 4 T), all measured, F = 0."""
    return shape_program(n_q=5, t=4, F=0, n_out=5, seed=seed)


def config2_program(seed: int = 2) -> FlatProgram:
    """d=5 repetition-code memory, 5 rounds, 9 qubits: Clifford (G = 1),
    ~29 outputs and ~100 f-bits."""
    return shape_program(n_q=9, t=0, F=100, n_out=29, seed=seed)


def config2_noise(seed: int = 102, num_f: int = 100, p: float = 1e-3):
    """Noise of the d=5 repetition-code memory (5 rounds): DEPOLARIZE2(p)
    after each of the 40 CX (4-bit channels, 15 equally likely Pauli pairs)
    and X_ERROR(p) on the 25 measured qubits: 185 e-bits mapped onto
    ``num_f`` f-bits by a seeded sparse GF(2) transform (each e-bit flips 1-3
    f-bits), as ``(channel_probs, error_transform)``."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(40):
        pr = np.full(16, p / 15.0)
        pr[0] = 1.0 - p
        probs.append(pr)
    for _ in range(25):
        probs.append(np.array([1.0 - p, p]))
    num_e = 40 * 4 + 25
    T = np.zeros((num_f, num_e), np.uint8)
    for e in range(num_e):
        T[rng.choice(num_f, size=int(rng.integers(1, 4)), replace=False), e] = 1
    return probs, T


def config3_program(seed: int = 3) -> FlatProgram:
    """30-qubit brickwork encoder, 12 T + DEPOLARIZE1 after each T:
    n_out = 30, F = 24, G ~ 2^{0.228*12} ~ 7."""
    return shape_program(n_q=30, t=12, F=24, n_out=30, seed=seed)


def config4_specs(seed: int = 4, n_programs: int = 64) -> list[dict]:
    """64 programs: q ~ U[16,60], t ~ U[4,24], F = 2t, shots ~ U[1k,100k]."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n_programs):
        q = int(rng.integers(16, 61))
        t = int(rng.integers(4, 25))
        shots = int(rng.integers(1000, 100001))
        specs.append(dict(n_q=q, t=t, F=2 * t, n_out=q, shots=shots, seed=1000 * seed + k))
    return specs


def config5_spec(rng: np.random.Generator, t_max: int = 40, q_max: int = 85) -> dict:
    """One serving request: config-4 generator extended to t <= 40, q <= 85,
    shots log-uniform in [100, 1M]."""
    q = int(rng.integers(16, q_max + 1))
    t = int(rng.integers(4, t_max + 1))
    shots = int(round(math.exp(rng.uniform(math.log(100), math.log(1_000_000)))))
    return dict(n_q=q, t=t, F=2 * t, n_out=q, shots=shots, seed=int(rng.integers(1 << 30)))


def program_from_spec(spec: dict) -> FlatProgram:
    return shape_program(spec["n_q"], spec["t"], spec["F"], spec["n_out"], spec["seed"])


# --------------------------------------------------------------------------
# stim text of the benchmark circuits (provenance of the shapes)
# --------------------------------------------------------------------------
def config1_circuit(seed: int = 1) -> str:
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(40):
        g = rng.choice(["H", "S", "CX"])
        if g == "CX":
            a, b = rng.choice(5, size=2, replace=False)
            ops.append(f"CX {a} {b}")
        else:
            ops.append(f"{g} {int(rng.integers(5))}")
    for pos in sorted(rng.integers(0, len(ops) + 1, size=4), reverse=True):
        ops.insert(int(pos), f"T {int(rng.integers(5))}")
    return "\n".join(ops + ["M 0 1 2 3 4"]) + "\n"


def config2_circuit(d: int = 5, rounds: int = 5, p: float = 1e-3) -> str:
    data = list(range(0, 2 * d - 1, 2))
    anc = list(range(1, 2 * d - 1, 2))
    lines = []
    for r in range(rounds):
        for a in anc:
            lines += [f"CX {a - 1} {a}", f"DEPOLARIZE2({p}) {a - 1} {a}",
                      f"CX {a + 1} {a}", f"DEPOLARIZE2({p}) {a + 1} {a}"]
        lines.append(f"X_ERROR({p}) " + " ".join(map(str, anc)))
        lines.append("MR " + " ".join(map(str, anc)))
    lines.append(f"X_ERROR({p}) " + " ".join(map(str, data)))
    lines.append("M " + " ".join(map(str, data)))
    return "\n".join(lines) + "\n"


def config3_circuit(n: int = 30, depth: int = 20, t_count: int = 12, seed: int = 3,
                    p: float = 1e-3) -> str:
    rng = np.random.default_rng(seed)
    t_at = {}
    for _ in range(t_count):
        t_at.setdefault(int(rng.integers(depth)), []).append(int(rng.integers(n)))
    lines = []
    for layer in range(depth):
        off = layer % 2
        for q in range(off, n - 1, 2):
            lines.append(f"CX {q} {q + 1}")
        for q in range(n):
            if rng.random() < 0.5:
                lines.append(f"{'H' if rng.random() < 0.5 else 'S'} {q}")
        for q in t_at.get(layer, []):
            lines += [f"T {q}", f"DEPOLARIZE1({p}) {q}"]
    lines.append("M " + " ".join(map(str, range(n))))
    return "\n".join(lines) + "\n"
