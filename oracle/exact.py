"""Exact big-integer oracle for the FlatProgram sampler path.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``ksim_amd`` package imports this
module; it is the checker that ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` compare the HIP path against.

What it restates (reference files under src/):

* ``UNIT`` / :func:`zmul` — the Z[omega] basis (1, omega, i, conj(omega)) and
  multiply of ``src/cpp/ksim/zx_eval.hpp:24-47`` and
  ``src/python/ksim/sample/exact_scalar.py:19-35``.
* :func:`term_value` — one stabilizer-rank term, family by family, as in
  ``zx_eval.hpp:102-165`` and ``src/python/ksim/sample/evaluate.py:32-61``,
  but with Python integers so nothing can wrap.
* :func:`graph_value` — the exact sum over terms (the reference sums in
  float64 at ``zx_eval.hpp:168-178`` / ``evaluate.py:63-66``; here the sum is
  exact and returned in canonical form).
* :func:`magnitude` — the float64 |amplitude| the samplers compare with,
  computed from the exact value by one fixed sequence of correctly rounded
  IEEE operations (the HIP kernel runs the same sequence).
* :func:`graph_float` / :func:`graph_magnitude` — graphs with approximate
  (complex64) floatfactors: exact terms cast to float64, times their
  factors, summed in term order (``zx_eval.hpp:168-178``).
* :func:`sample_component` / :func:`sample_program` — the autoregressive
  recurrence of ``evaluate.py:69-87`` and ``src/cpp/ksim/kokkos_sim.cpp:214-230``
  driven by given uniforms, in ``sample_flat``'s column order.
* :func:`kokkos_uniforms` — the per-shot xorshift64 stream and seed mix of
  ``kokkos_sim.cpp:43-46,215-217``.

Pinning (see tests/test_oracle.py): the multiply against exact complex
arithmetic in Q(sqrt2)(i); the term formula against brute-force closed forms
(H.T^k.H gives sin^2(k pi/8)); term splitting invariance; the float64 value
against the reference's own ``evaluate_flat`` on golden vectors; the kokkos
stream against the reference's published H.T.H.T.H figure (P(1) = 0.25034 at
seed 43, 1,000,000 shots).
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

# omega^k in the basis (1, omega, i, conj(omega)); omega^3 = -conj(omega).
UNIT = ((1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, -1),
        (-1, 0, 0, 0), (0, -1, 0, 0), (0, 0, -1, 0), (0, 0, 0, 1))
ONE = (1, 0, 0, 0)
ZERO = (0, 0, 0, 0)

# double nearest cos(pi/4), the constant zx_eval.hpp:100 uses
SQ = 0.70710678118654752440


def zmul(x, y):
    """Z[omega] product (zx_eval.hpp:27-33)."""
    a1, b1, c1, d1 = x
    a2, b2, c2, d2 = y
    return (a1 * a2 + b1 * d2 - c1 * c2 + d1 * b2,
            a1 * b2 + b1 * a2 + c1 * d2 + d1 * c2,
            a1 * c2 + b1 * b2 + c1 * a2 - d1 * d2,
            a1 * d2 - b1 * c2 - c1 * b2 + d1 * a2)


def zadd(x, y):
    return tuple(p + q for p, q in zip(x, y))


def zshift(x, k):
    """x * 2**k for k >= 0."""
    return tuple(c << k for c in x)


def canonical(c, e):
    """Unique form of c * 2**e: zero is ((0,0,0,0), 0); otherwise halve while
    all four coefficients are even."""
    if all(v == 0 for v in c):
        return ZERO, 0
    while all(v % 2 == 0 for v in c):
        c = tuple(v // 2 for v in c)
        e += 1
    return tuple(c), e


def _parity(mask_row, params):
    """<mask, params> mod 2 (zx_eval.hpp:91-97, exact_scalar.py:64-66)."""
    return int(np.dot(mask_row.astype(np.int64), params) & 1)


def term_value(d: dict, t: int, params: np.ndarray):
    """Exact value of term t of graph dict d: (coeffs, exponent) with the
    value coeffs * 2**exponent, or None when a factor is zero.

    Families (zx_eval.hpp:106-165, evaluate.py:32-61):
      NodePhases  prod_{j<np_counts[t]} (1 + w^{(4 par + phase) mod 8})
      HalfPi      w^{sum (par * coeff) mod 8}
      PiProducts  (-1)^{sum psi * phi}, psi = const ^ par
      PhasePairs  prod_{j<ph_counts[t]} (1 + w^a + w^b - w^{a+b}),
                  a = alpha + 4 par_a, b = beta + 4 par_b
      prefactor   w^{pf_phase_idx} * pf_floatfactor * 2^{pf_power2}
    """
    acc = ONE
    np_ph = d["np_phases"]; np_pa = d["np_params"]
    for j in range(int(d["np_counts"][t])):
        k = (4 * _parity(np_pa[t, j], params) + int(np_ph[t, j])) & 7
        u = UNIT[k]
        acc = zmul(acc, (u[0] + 1, u[1], u[2], u[3]))
    hp_c = d["hp_coeffs"]; hp_p = d["hp_params"]
    tot = 0
    for j in range(hp_c.shape[1] if hp_c.ndim == 2 else 0):
        tot += (_parity(hp_p[t, j], params) * int(hp_c[t, j])) & 7
    acc = zmul(acc, UNIT[tot & 7])
    e = 0
    psc, psp = d["pp_psi_c"], d["pp_psi_p"]
    phc, php = d["pp_phi_c"], d["pp_phi_p"]
    for j in range(psc.shape[1] if psc.ndim == 2 else 0):
        psi = (int(psc[t, j]) + _parity(psp[t, j], params)) & 1
        phi = (int(phc[t, j]) + _parity(php[t, j], params)) & 1
        e ^= psi & phi
    if e:
        acc = tuple(-v for v in acc)
    al, alp = d["ph_alpha"], d["ph_alpha_p"]
    be, bep = d["ph_beta"], d["ph_beta_p"]
    for j in range(int(d["ph_counts"][t])):
        a = (int(al[t, j]) + 4 * _parity(alp[t, j], params)) & 7
        b = (int(be[t, j]) + 4 * _parity(bep[t, j], params)) & 7
        u, v, w = UNIT[a], UNIT[b], UNIT[(a + b) & 7]
        acc = zmul(acc, (1 + u[0] + v[0] - w[0], u[1] + v[1] - w[1],
                         u[2] + v[2] - w[2], u[3] + v[3] - w[3]))
    acc = zmul(acc, UNIT[int(d["pf_phase_idx"][t]) & 7])
    ff = tuple(int(x) for x in d["pf_floatfactor"][t])
    acc = zmul(acc, ff)
    if all(v == 0 for v in acc):
        return None
    return acc, int(d["pf_power2"][t])


def graph_value(d: dict, params) -> tuple[tuple, int]:
    """Exact amplitude of graph d at params, canonical (coeffs, exponent).

    G = 0 gives zero (evaluate.py:27-28)."""
    params = np.asarray(params, dtype=np.int64).reshape(-1)
    G = int(np.asarray(d["pf_phase_idx"]).shape[0])
    terms = []
    for t in range(G):
        tv = term_value(d, t, params)
        if tv is not None:
            terms.append(tv)
    if not terms:
        return ZERO, 0
    emin = min(e for _, e in terms)
    s = ZERO
    for c, e in terms:
        s = zadd(s, zshift(c, e - emin))
    return canonical(s, emin)


def int_to_double(x: int) -> float:
    """Correctly rounded (round-half-even) conversion; Python's int->float is."""
    return float(x)


def magnitude(c, e) -> float:
    """float64 |c * 2**e| by the fixed sequence the HIP kernel also runs:

        s = b + d, t = b - d                       (exact)
        k = max(0, L - 62), L = max bit length of |a|, |c|, |s|, |t|
        a', c', s', t' = a >> k, c >> k, s >> k, t >> k   (floor)
        re = dbl(a') + dbl(s') * SQ
        im = dbl(c') + dbl(t') * SQ
        |z| = ldexp(sqrt(re*re + im*im), e + k)

    dbl() is the correctly rounded integer->double conversion, every other
    step one IEEE-754 double operation with no fused multiply-add.  The
    re/im formula is the cast of zx_eval.hpp:168-169, applied to the exact
    sum.  The result does not depend on the representation (scaling c by 2
    and e by -1 leaves every step unchanged) and the truncation to 62 bits
    costs less than 2^-61 relative.
    """
    a, b, cc, d = c
    s, t = b + d, b - d
    L = max(abs(a).bit_length(), abs(cc).bit_length(), abs(s).bit_length(), abs(t).bit_length())
    k = max(0, L - 62)
    a, cc, s, t = a >> k, cc >> k, s >> k, t >> k
    re = int_to_double(a) + int_to_double(s) * SQ
    im = int_to_double(cc) + int_to_double(t) * SQ
    return math.ldexp(math.sqrt(re * re + im * im), e + k)


def term_float(c, e) -> tuple[float, float]:
    """float64 (re, im) of the exact value c * 2**e by the first steps of
    :func:`magnitude` (same 62-bit window): re = dbl(a') + dbl(s') * SQ,
    im = dbl(c') + dbl(t') * SQ, scaled by 2**(e + k).  For coefficients
    below 2**53 this is the reference's cast of one term
    (zx_eval.hpp:168-171), whatever the representation."""
    a, b, cc, d = c
    s, t = b + d, b - d
    L = max(abs(a).bit_length(), abs(cc).bit_length(), abs(s).bit_length(), abs(t).bit_length())
    k = max(0, L - 62)
    a, cc, s, t = a >> k, cc >> k, s >> k, t >> k
    re = int_to_double(a) + int_to_double(s) * SQ
    im = int_to_double(cc) + int_to_double(t) * SQ
    return math.ldexp(re, e + k), math.ldexp(im, e + k)


def approx_factor(d: dict, t: int) -> tuple[float, float]:
    """The complex64 factor of term t as two float64 values (the float32
    pairs of pf_approx_f32, src/python/ksim/sample/sampler.py:22,41)."""
    f = np.asarray(d["pf_approx_f32"], dtype=np.float32).reshape(-1, 2)
    return float(f[t, 0]), float(f[t, 1])


def graph_float(d: dict, params) -> tuple[float, float]:
    """float64 amplitude of a graph with approximate floatfactors
    (pf_has_approx): each term's exact value cast by :func:`term_float`,
    multiplied by its factor (re*ar - im*ai, re*ai + im*ar) and added in
    term order, as zx_eval.hpp:168-178 does; zero terms add nothing."""
    params = np.asarray(params, dtype=np.int64).reshape(-1)
    G = int(np.asarray(d["pf_phase_idx"]).shape[0])
    sr = si = 0.0
    for t in range(G):
        tv = term_value(d, t, params)
        if tv is None:
            continue
        re, im = term_float(*tv)
        ar, ai = approx_factor(d, t)
        sr += re * ar - im * ai
        si += re * ai + im * ar
    return sr, si


def graph_magnitude(d: dict, params) -> float:
    """The float64 |amplitude| the samplers decide on: :func:`magnitude` of
    the exact sum, or, for a graph with approximate floatfactors,
    sqrt(re*re + im*im) of :func:`graph_float` (kokkos_sim.cpp:220,224)."""
    if d["pf_has_approx"]:
        sr, si = graph_float(d, params)
        return math.sqrt(sr * sr + si * si)
    return magnitude(*graph_value(d, params))


def exact_abs_fraction_sq(c, e) -> tuple[Fraction, Fraction]:
    """|c*2^e|^2 = X + Y*sqrt(2) exactly, as Fractions (X, Y)."""
    a, b, cc, d = c
    X = a * a + b * b + cc * cc + d * d
    Y = a * (b + d) + cc * (b - d)
    s = Fraction(2) ** (2 * e)
    return Fraction(X) * s, Fraction(Y) * s


def params_row(f_row, m_bits, i, trying=True):
    """[f | m_0..m_{i-1} | trying] (zx_eval.hpp:65, evaluate.py:79-80)."""
    row = list(int(x) for x in f_row) + list(int(x) for x in m_bits[:i])
    if trying:
        row.append(1)
    return np.asarray(row, dtype=np.int64)


def sample_component(graph_dicts, f_sel: np.ndarray, uniforms: np.ndarray,
                     trace: list | None = None) -> np.ndarray:
    """Autoregressive sampling of one component with given uniforms.

    f_sel (B, F) {0,1}; uniforms (B, n_out) float64.  For each shot:
    prev = |graph0|; for output i: p1 = |graph i+1 at [f|m_<i|1]|,
    bit = u < p1/prev, prev = p1 if bit else prev - p1
    (evaluate.py:76-85, kokkos_sim.cpp:219-229).  When ``trace`` is a list,
    each (shot, graph, canonical exact value) is appended to it.
    """
    B = f_sel.shape[0]
    n_out = len(graph_dicts) - 1
    out = np.zeros((B, max(n_out, 0)), dtype=np.uint8)

    def prob(d, params, s, i):
        if d["pf_has_approx"]:
            v = graph_float(d, params)
            if trace is not None:
                trace.append((s, i, v))
            return math.sqrt(v[0] * v[0] + v[1] * v[1])
        v = graph_value(d, params)
        if trace is not None:
            trace.append((s, i, v))
        return magnitude(*v)

    with np.errstate(divide="ignore", invalid="ignore"):
        for s in range(B):
            f_row = f_sel[s]
            prev = prob(graph_dicts[0], np.asarray(f_row, dtype=np.int64), s, 0)
            m = []
            for i in range(n_out):
                p1 = prob(graph_dicts[i + 1], params_row(f_row, m, i), s, i + 1)
                ratio = np.float64(p1) / np.float64(prev)
                bit = 1 if float(uniforms[s, i]) < ratio else 0
                m.append(bit)
                prev = p1 if bit else prev - p1
            out[s, :] = m
    return out


def sample_program(prog, f_params: np.ndarray, uniforms: np.ndarray,
                   graph_dicts=None) -> np.ndarray:
    """Whole-program sampling with a (B, total_outputs) uniforms matrix whose
    columns are component-major then output-major, the order in which
    ``sample_flat`` draws ``rng.random(B)`` (evaluate.py:73-82).  Returns the
    (B, num_outputs) bits permuted by argsort(output_order) (evaluate.py:86-87).
    """
    from ksim_amd.program import graph_to_dict
    B = f_params.shape[0]
    outs = []
    col = 0
    for k, comp in enumerate(prog.components):
        gds = graph_dicts[k] if graph_dicts is not None else [graph_to_dict(g) for g in comp.graphs]
        n_out = len(gds) - 1
        f_sel = f_params[:, np.asarray(comp.f_selection, dtype=np.int64)].astype(np.int64)
        outs.append(sample_component(gds, f_sel, uniforms[:, col:col + n_out]))
        col += n_out
    combined = np.concatenate(outs, axis=1) if outs else np.zeros((B, 0), np.uint8)
    return combined[:, np.argsort(prog.output_order)]


def flat_uniforms(prog, B: int, rng: np.random.Generator) -> np.ndarray:
    """The uniforms ``sample_flat`` draws from ``rng``: one rng.random(B) per
    output, component-major (evaluate.py:73-82)."""
    cols = []
    for comp in prog.components:
        for _ in range(len(comp.graphs) - 1):
            cols.append(rng.random(B))
    if not cols:
        return np.zeros((B, 0))
    return np.stack(cols, axis=1)


_M64 = (1 << 64) - 1


def kokkos_seed(seed: int, s: int) -> int:
    """Per-shot xorshift64 state (kokkos_sim.cpp:215-217)."""
    r = (seed ^ ((s * 0x9E3779B97F4A7C15) & _M64)) & _M64
    r ^= r >> 30
    r = (r * 0xBF58476D1CE4E5B9) & _M64
    r ^= r >> 27
    return r if r else 1


def _xorshift_stream(seed: int, shots, n_out: int, shifts=(13, 7, 17), mix=(30, 27),
                     mul=(0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9)) -> np.ndarray:
    """The xorshift64 draws of the given shots with the step shifts, the
    seed-mix shifts and multipliers as parameters (the reference's by
    default); tests plant other values here to show the known answer
    rejects them."""
    s = np.asarray(shots, dtype=np.uint64).reshape(-1)
    out = np.zeros((s.shape[0], n_out), dtype=np.float64)
    a, b, c = (np.uint64(x) for x in shifts)
    with np.errstate(over="ignore"):
        r = np.uint64(seed & _M64) ^ (s * np.uint64(mul[0]))
        r ^= r >> np.uint64(mix[0])
        r = r * np.uint64(mul[1])
        r ^= r >> np.uint64(mix[1])
        r[r == 0] = np.uint64(1)
        for i in range(n_out):
            r ^= r << a
            r ^= r >> b
            r ^= r << c
            out[:, i] = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return out


def kokkos_uniforms_at(seed: int, shots, n_out: int) -> np.ndarray:
    """(len(shots), n_out) uniforms of the xorshift64 stream of the given
    shot indices (kokkos_sim.cpp:43-46,215-217): the state of shot s is
    kokkos_seed(seed, s); each draw is s ^= s<<13; s ^= s>>7; s ^= s<<17;
    u = (s>>11) * 2^-53.  Vectorised over shots in uint64 arithmetic
    (wrapping multiplies are the reference's uint64 semantics).

    Pinned by the reference's published H.T.H.T.H figure: its GPU sampler
    (seed 42 + 1, component 0, 1,000,000 shots, docs/benchmarks.md:82,124-125,
    benchmarks/bench_xtim.py:146,164-167) measured P(1) = 0.25034 against the
    exact 1/4, i.e. the fraction of shots whose first uniform is below 1/4."""
    return _xorshift_stream(seed, shots, n_out)


def kokkos_uniforms(seed: int, B: int, n_out: int) -> np.ndarray:
    """(B, n_out) uniforms of shots 0..B-1 (kokkos_uniforms_at)."""
    return kokkos_uniforms_at(seed, np.arange(B, dtype=np.uint64), n_out)
