"""float64 numpy restatement of the reference's CPU sampler.

TEST INFRASTRUCTURE ONLY (the checker and the numpy CPU baseline; never part
of the product path).

Restates, with the same sequence of numpy operations so that the float64
results agree bit for bit:

* ``evaluate_flat`` — ``src/python/ksim/sample/evaluate.py:26-66``: int64
  Z[omega] products per family with one halving step per multiply
  (``exact_scalar.py:38-55``), the cross-family chain without halving
  (``evaluate.py:58-61``), the complex cast of ``exact_scalar.py:58-61`` and a
  float sum over terms;
* ``sample_flat`` — ``evaluate.py:69-87``: one ``rng.random(B)`` per output,
  component-major, ``bit = u < p1/prev``, ``prev = where(bit, p1, prev-p1)``.

Pinned by the golden vectors in tests/golden (written from the reference's
own module) and, for values, by the exact oracle in ``oracle/exact.py``.
"""
from __future__ import annotations

import numpy as np

UNIT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1],
                 [-1, 0, 0, 0], [0, -1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]],
                dtype=np.int64)
ONE_PLUS = UNIT.copy()
ONE_PLUS[:, 0] += 1
ID4 = np.array([1, 0, 0, 0], dtype=np.int64)
OMEGA = np.exp(1j * np.pi / 4)


def _zmul(x, y):
    a1, b1, c1, d1 = (x[..., k] for k in range(4))
    a2, b2, c2, d2 = (y[..., k] for k in range(4))
    return np.stack([a1 * a2 + b1 * d2 - c1 * c2 + d1 * b2,
                     a1 * b2 + b1 * a2 + c1 * d2 + d1 * c2,
                     a1 * c2 + b1 * b2 + c1 * a2 - d1 * d2,
                     a1 * d2 - b1 * c2 - c1 * b2 + d1 * a2], axis=-1)


def _halve_once(power, c):
    ev = np.all(c % 2 == 0, axis=-1) & np.any(c != 0, axis=-1)
    return power + ev.astype(power.dtype), np.where(ev[..., None], c // 2, c)


def _chain(vals, axis):
    """Product along ``axis`` with one halving step after each multiply."""
    v = np.moveaxis(vals, axis, 0)
    if v.shape[0] == 0:
        shp = v.shape[1:]
        return np.broadcast_to(ID4, shp).copy(), np.zeros(shp[:-1], np.int64)
    acc = v[0].astype(np.int64)
    pw = np.zeros(acc.shape[:-1], np.int64)
    for k in range(1, v.shape[0]):
        acc = _zmul(acc, v[k].astype(np.int64))
        pw, acc = _halve_once(pw, acc)
    return acc, pw


def _cast(c, power):
    z = c[..., 0] + c[..., 1] * OMEGA + c[..., 2] * 1j + c[..., 3] * np.conj(OMEGA)
    return z * np.exp2(power.astype(np.float64))


def _par(masks, vals):
    return (np.einsum("gtp,bp->bgt", masks.astype(np.int64),
                      vals.astype(np.int64)) % 2).astype(np.int64)


def evaluate_graph(g, vals: np.ndarray) -> np.ndarray:
    """complex128 (B,) amplitude of one graph (object with the FlatScalarGraph
    fields) at parameter rows ``vals`` (B, P)."""
    if g.pf_phase_idx.shape[0] == 0:
        return np.zeros(vals.shape[0], dtype=np.complex128)
    B = vals.shape[0]
    Gn = g.pf_phase_idx.shape[0]

    r = _par(g.np_params, vals)
    tv = ONE_PLUS[(4 * r + g.np_phases[None]) % 8]
    live = np.arange(g.np_phases.shape[1])[None, :] < g.np_counts[:, None]
    tv = np.where(live[None, ..., None], tv, ID4)
    c_np, p_np = _chain(tv, axis=2)

    r = _par(g.hp_params, vals)
    c_hp = UNIT[np.sum((r * g.hp_coeffs[None]) % 8, axis=-1) % 8]

    psi = (g.pp_psi_c[None] + _par(g.pp_psi_p, vals)) % 2
    phi = (g.pp_phi_c[None] + _par(g.pp_phi_p, vals)) % 2
    sg = np.sum(psi * phi, axis=-1) % 2
    c_pp = (1 - 2 * sg)[..., None] * ID4

    ra = _par(g.ph_alpha_p, vals)
    rb = _par(g.ph_beta_p, vals)
    a = (g.ph_alpha[None] + 4 * ra) % 8
    b = (g.ph_beta[None] + 4 * rb) % 8
    tv = ID4 + UNIT[a] + UNIT[b] - UNIT[(a + b) % 8]
    live = np.arange(g.ph_alpha.shape[1])[None, :] < g.ph_counts[:, None]
    tv = np.where(live[None, ..., None], tv, ID4)
    c_ph, p_ph = _chain(tv, axis=2)

    c_st = UNIT[g.pf_phase_idx]
    c_ff = g.pf_floatfactor.astype(np.int64)
    coeff = _zmul(_zmul(_zmul(c_np, c_hp), _zmul(c_pp, c_ph)),
                  _zmul(np.broadcast_to(c_st, (B, Gn, 4)),
                        np.broadcast_to(c_ff, (B, Gn, 4))))
    z = _cast(coeff, p_np + p_ph)
    if g.pf_has_approx:
        return np.sum(z * g.pf_approx[None] * np.exp2(g.pf_power2)[None], axis=-1)
    return np.sum(z * np.exp2(g.pf_power2.astype(np.float64))[None], axis=-1)


def sample_flat(prog, f_params: np.ndarray, rng: np.random.Generator,
                uniforms: np.ndarray | None = None) -> np.ndarray:
    """The reference's numpy sampler; ``uniforms`` (B, total_outputs), when
    given, replaces the rng draws column by column in the same order."""
    B = f_params.shape[0]
    outs = []
    col = 0
    for comp in prog.components:
        n_out = len(comp.graphs) - 1
        f_sel = f_params[:, comp.f_selection].astype(np.int64)
        m = np.zeros((B, n_out), dtype=np.int64)
        with np.errstate(divide="ignore", invalid="ignore"):
            prev = np.abs(evaluate_graph(comp.graphs[0], f_sel))
            for i, g in enumerate(comp.graphs[1:]):
                params = np.hstack([f_sel, m[:, :i], np.ones((B, 1), np.int64)])
                p1 = np.abs(evaluate_graph(g, params))
                u = rng.random(B) if uniforms is None else uniforms[:, col]
                col += 1
                bits = u < (p1 / prev)
                m[:, i] = bits
                prev = np.where(bits, p1, prev - p1)
        outs.append(m)
    combined = np.concatenate(outs, axis=1) if outs else np.zeros((B, 0), int)
    return combined[:, np.argsort(prog.output_order)]


def channel_sample(channel_probs, error_transform, num_samples: int,
                   rng: np.random.Generator) -> np.ndarray:
    """The reference's host noise sampler, restated: ``ChannelSampler.sample``
    (src/python/ksim/compile/channels.py:622-658) with its per-channel data
    (``_precompute_sparse``, :575-620) built without the simplification
    passes, which leave the f distribution unchanged.  Geometric skips along
    the shot axis, ``searchsorted`` on the conditional CDF, XOR of the
    outcome's pattern into the rows.  Used as the CPU baseline of noise
    generation (the stream differs from the reference's only in that
    simplified channels consume the generator in a different order)."""
    T = np.asarray(error_transform, dtype=np.uint8)
    num_f = T.shape[0]
    result = np.zeros((num_samples, num_f), dtype=np.uint8)
    e_off = 0
    for probs in channel_probs:
        probs = np.asarray(probs, dtype=np.float64)
        n = len(probs)
        k = int(np.log2(n))
        cols = T[:, e_off:e_off + k]
        e_off += k
        p_fire = 1.0 - float(probs[0])
        if p_fire <= 1e-15 or n <= 1:
            continue
        cond_cdf = np.cumsum(probs[1:] / p_fire, dtype=np.float64)
        cond_cdf /= cond_cdf[-1]
        outcomes = np.arange(1, n)
        bits = ((outcomes[:, None] >> np.arange(k)) & 1).astype(np.uint8)
        pats = (bits @ cols.T % 2).astype(np.uint8)
        expected = num_samples * p_fire
        sigma = np.sqrt(expected * (1.0 - p_fire))
        n_draws = int(expected + 7.0 * sigma) + 100
        positions = np.cumsum(rng.geometric(p_fire, size=n_draws)) - 1
        positions = positions[positions < num_samples]
        if len(positions) == 0:
            continue
        idx = np.searchsorted(cond_cdf, rng.uniform(size=len(positions)))
        result[positions] ^= pats[idx]
    return result
