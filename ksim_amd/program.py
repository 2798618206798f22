"""FlatProgram IR, re-declared with the reference's names, fields and dtypes.

The three dataclasses are the data contract of the sampler path and are kept
field-for-field identical to ``src/python/ksim/program.py:29-53`` so that a
``FlatProgram`` built by ``ksim.compile`` (or by the reference's
``flatten_program``) can be handed to :class:`ksim_amd.HipProgramSampler`
unchanged.

Per scalar graph (one per autoregressive step; ``G`` stabilizer-rank terms,
``P`` parity inputs ``[f (F) | m prefix (i) | trying]``):

==============  ============================================  =====
field           shape                                         dtype
==============  ============================================  =====
np_phases       (G, T_np)      node phases mod 8               u8
np_params       (G, T_np, P)   node parity masks               u8
np_counts       (G,)           live node factors per term      i32
hp_coeffs       (G, T_hp)      half-pi coefficients            u8
hp_params       (G, T_hp, P)                                   u8
pp_psi_c/phi_c  (G, T_pp)      pi-product constant bits        u8
pp_psi_p/phi_p  (G, T_pp, P)                                   u8
ph_alpha/beta   (G, T_ph)      phase-pair phases mod 8         u8
ph_alpha_p/b_p  (G, T_ph, P)                                   u8
ph_counts       (G,)           live phase pairs per term       i32
pf_phase_idx    (G,)           prefactor omega exponent        u8
pf_floatfactor  (G, 4)         prefactor Z[omega] coefficients i32
pf_power2       (G,)           prefactor power of two          i32
pf_approx       (G,)           approximate factor              c64
pf_has_approx   bool
==============  ============================================  =====

:func:`validate_graph` enforces the shapes and the dtype coercions that the
reference applies before upload (``src/python/ksim/sample/sampler.py:21-43``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class FlatScalarGraph:
    np_phases: np.ndarray;  np_params: np.ndarray;  np_counts: np.ndarray
    hp_coeffs: np.ndarray;  hp_params: np.ndarray
    pp_psi_c: np.ndarray;   pp_psi_p: np.ndarray
    pp_phi_c: np.ndarray;   pp_phi_p: np.ndarray
    ph_alpha: np.ndarray;   ph_alpha_p: np.ndarray
    ph_beta: np.ndarray;    ph_beta_p: np.ndarray;  ph_counts: np.ndarray
    pf_phase_idx: np.ndarray; pf_floatfactor: np.ndarray
    pf_power2: np.ndarray;    pf_approx: np.ndarray; pf_has_approx: bool


@dataclass
class FlatComponent:
    output_indices: tuple
    f_selection: np.ndarray
    graphs: list = field(default_factory=list)


@dataclass
class FlatProgram:
    components: list
    output_order: np.ndarray
    num_outputs: int
    num_detectors: int


# Order of the array fields of one graph, as the reference's ``_graph_to_dict``
# packs them (src/python/ksim/sample/sampler.py:21-43).
GRAPH_U8_3D = ("np_params", "hp_params", "pp_psi_p", "pp_phi_p",
               "ph_alpha_p", "ph_beta_p")


def graph_to_dict(g: FlatScalarGraph) -> dict:
    """Coerce one graph to the dict of C-contiguous arrays the native
    component constructor takes.

    Same keys, dtypes and the complex64-as-float32-pairs view as the
    reference's ``_graph_to_dict`` (src/python/ksim/sample/sampler.py:21-43);
    arrays that already have the right dtype and layout are not copied.
    """
    apx = np.ascontiguousarray(g.pf_approx, dtype=np.complex64)
    return dict(
        np_phases=np.ascontiguousarray(g.np_phases, dtype=np.uint8),
        np_params=np.ascontiguousarray(g.np_params, dtype=np.uint8),
        np_counts=np.ascontiguousarray(g.np_counts, dtype=np.int32),
        hp_coeffs=np.ascontiguousarray(g.hp_coeffs, dtype=np.uint8),
        hp_params=np.ascontiguousarray(g.hp_params, dtype=np.uint8),
        pp_psi_c=np.ascontiguousarray(g.pp_psi_c, dtype=np.uint8),
        pp_psi_p=np.ascontiguousarray(g.pp_psi_p, dtype=np.uint8),
        pp_phi_c=np.ascontiguousarray(g.pp_phi_c, dtype=np.uint8),
        pp_phi_p=np.ascontiguousarray(g.pp_phi_p, dtype=np.uint8),
        ph_alpha=np.ascontiguousarray(g.ph_alpha, dtype=np.uint8),
        ph_alpha_p=np.ascontiguousarray(g.ph_alpha_p, dtype=np.uint8),
        ph_beta=np.ascontiguousarray(g.ph_beta, dtype=np.uint8),
        ph_beta_p=np.ascontiguousarray(g.ph_beta_p, dtype=np.uint8),
        ph_counts=np.ascontiguousarray(g.ph_counts, dtype=np.int32),
        pf_phase_idx=np.ascontiguousarray(g.pf_phase_idx, dtype=np.uint8),
        pf_floatfactor=np.ascontiguousarray(g.pf_floatfactor, dtype=np.int32),
        pf_power2=np.ascontiguousarray(g.pf_power2, dtype=np.int32),
        pf_approx_f32=apx.view(np.float32),
        pf_has_approx=bool(g.pf_has_approx),
    )


def graph_shape(d: dict) -> tuple[int, int, int, int, int, int]:
    """(G, P, T_np, T_hp, T_pp, T_ph) read from array shapes.

    Mirrors how the reference's native constructor reads its per-graph meta
    (src/cpp/ksim/kokkos_sim.cpp:93-103): a missing trailing dimension
    counts as 1, and P is taken from ``np_params``.
    """
    def dim(key, k):
        a = np.asarray(d[key])
        return int(a.shape[k]) if k < a.ndim else 1
    return (dim("pf_phase_idx", 0), dim("np_params", 2), dim("np_phases", 1),
            dim("hp_coeffs", 1), dim("pp_psi_c", 1), dim("ph_alpha", 1))


def validate_graph(d: dict) -> None:
    """Check that a graph dict is internally consistent; raise ValueError."""
    G, P, Tnp, Thp, Tpp, Tph = graph_shape(d)
    want = {
        "np_phases": (G, Tnp), "np_params": (G, Tnp, P), "np_counts": (G,),
        "hp_coeffs": (G, Thp), "hp_params": (G, Thp, P),
        "pp_psi_c": (G, Tpp), "pp_psi_p": (G, Tpp, P),
        "pp_phi_c": (G, Tpp), "pp_phi_p": (G, Tpp, P),
        "ph_alpha": (G, Tph), "ph_alpha_p": (G, Tph, P),
        "ph_beta": (G, Tph), "ph_beta_p": (G, Tph, P), "ph_counts": (G,),
        "pf_phase_idx": (G,), "pf_floatfactor": (G, 4), "pf_power2": (G,),
    }
    for k, shp in want.items():
        a = np.asarray(d[k])
        if a.size == 0 and 0 in shp:
            continue
        if a.shape != shp:
            raise ValueError(f"graph field {k!r}: shape {a.shape}, expected {shp}")
    if np.any(np.asarray(d["np_counts"]) > Tnp) or np.any(np.asarray(d["np_counts"]) < 0):
        raise ValueError("np_counts out of range")
    if np.any(np.asarray(d["ph_counts"]) > Tph) or np.any(np.asarray(d["ph_counts"]) < 0):
        raise ValueError("ph_counts out of range")
