"""The exact oracle for large programs: graph values from oracle/exact_fast.c.

TEST INFRASTRUCTURE ONLY (a checker for tests/, never imported by
``ksim_amd``).  ``graph_values`` returns the same canonical exact values as
:func:`oracle.exact.graph_value` (the C code restates the same family
products of src/cpp/ksim/zx_eval.hpp:102-165 in 512-bit integers and refuses
anything that would leave that range); ``sample_program`` runs the
autoregressive recurrence of src/python/ksim/sample/evaluate.py:69-87 on
those values with :func:`oracle.exact.magnitude`, vectorised over shots, and
equals :func:`oracle.exact.sample_program`.  tests/test_oracle.py checks both
equalities on the golden programs and on random shape programs.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import exact

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libexact_fast.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build_ref
            build_ref.build_exact_fast()
        from ksim_amd import _lib as kl
        lib = C.CDLL(LIB)
        lib.xf_graph_create.restype = C.c_void_p
        lib.xf_graph_create.argtypes = [C.POINTER(kl.GraphView)]
        lib.xf_graph_destroy.restype = None
        lib.xf_graph_destroy.argtypes = [C.c_void_p]
        lib.xf_graph_values.restype = C.c_int
        lib.xf_graph_values.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_int64,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]
        lib.xf_limbs.restype = C.c_int
        _lib = lib
    return _lib


class FastGraph:
    """One graph dict (graph_to_dict) held by the C oracle."""

    def __init__(self, d: dict):
        from ksim_amd import _lib as kl
        self.lib = load()
        self.keep: list = []
        self.view = kl.make_view(d, self.keep)
        self.P = int(self.view.P)
        self.h = self.lib.xf_graph_create(C.byref(self.view))

    def values(self, params: np.ndarray) -> list:
        """Canonical exact values (coeffs, exponent) at parameter rows (B, P)."""
        params = np.ascontiguousarray(params, dtype=np.uint8)
        if params.ndim == 1:
            params = params.reshape(1, -1)
        if params.shape[1] != self.P:
            raise ValueError(f"params have {params.shape[1]} columns, the graph has P={self.P}")
        B = params.shape[0]
        nl = self.lib.xf_limbs()
        coeffs = np.zeros((B, 4, nl), np.uint64)
        exps = np.zeros(B, np.int64)
        pp = params.ctypes.data_as(C.POINTER(C.c_uint8)) if params.size else \
            np.zeros(1, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
        rc = self.lib.xf_graph_values(self.h, pp, B, coeffs.ctypes.data_as(C.POINTER(C.c_uint64)),
                                      exps.ctypes.data_as(C.POINTER(C.c_int64)))
        if rc != 0:
            raise OverflowError("exact value beyond the C oracle's 512-bit range")
        out = []
        raw = coeffs.view(np.uint8).reshape(B, 4, 8 * nl)
        for r in range(B):
            c = tuple(int.from_bytes(raw[r, k].tobytes(), "little", signed=True) for k in range(4))
            out.append((c, int(exps[r])))
        return out

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            self.lib.xf_graph_destroy(h)


def graph_values(d: dict, params: np.ndarray) -> list:
    return FastGraph(d).values(params)


def magnitudes(d: dict, rows: np.ndarray) -> np.ndarray:
    """exact.graph_magnitude at every parameter row: exact sums from the C
    code, graphs with approximate floatfactors by exact.graph_float."""
    rows = np.asarray(rows, dtype=np.uint8)
    if d["pf_has_approx"]:
        return np.array([exact.graph_magnitude(d, r) for r in rows], dtype=np.float64)
    return np.array([exact.magnitude(*v) for v in FastGraph(d).values(rows)], dtype=np.float64)


def sample_component(graph_dicts, f_sel: np.ndarray, uniforms: np.ndarray) -> np.ndarray:
    """exact.sample_component, all shots of a graph at a time."""
    B = f_sel.shape[0]
    n_out = len(graph_dicts) - 1
    m = np.zeros((B, n_out), np.uint8)
    f8 = np.asarray(f_sel, dtype=np.uint8)
    prev = magnitudes(graph_dicts[0], f8).reshape(B)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i in range(n_out):
            params = np.hstack([f8, m[:, :i], np.ones((B, 1), np.uint8)])
            p1 = magnitudes(graph_dicts[i + 1], params).reshape(B)
            bits = np.asarray(uniforms[:, i], dtype=np.float64) < p1 / prev
            m[:, i] = bits
            prev = np.where(bits, p1, prev - p1)
    return m


def sample_program(prog, f_params: np.ndarray, uniforms: np.ndarray) -> np.ndarray:
    """exact.sample_program on the C oracle's values."""
    from ksim_amd.program import graph_to_dict
    B = f_params.shape[0]
    outs, col = [], 0
    for comp in prog.components:
        gds = [graph_to_dict(g) for g in comp.graphs]
        n_out = len(gds) - 1
        f_sel = f_params[:, np.asarray(comp.f_selection, dtype=np.int64)]
        outs.append(sample_component(gds, f_sel, uniforms[:, col:col + n_out]))
        col += n_out
    combined = np.concatenate(outs, axis=1) if outs else np.zeros((B, 0), np.uint8)
    return combined[:, np.argsort(prog.output_order)]
