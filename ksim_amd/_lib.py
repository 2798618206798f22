"""ctypes binding of ``libksim_hip.so`` (the C ABI declared in include/ksim_hip.h).

The library is built in-tree by ``ksim_amd.build``.  There is no fallback:
if the library is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libksim_hip.so")

KSIM_OK = 0
KSIM_E_INVALID = -1
KSIM_E_UNSUPPORTED = -2
KSIM_E_DEVICE = -3
KSIM_E_RANGE = -4

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class GraphView(C.Structure):
    """ksim_graph_view (include/ksim_hip.h)."""
    _fields_ = [("G", C.c_int32), ("P", C.c_int32), ("T_np", C.c_int32), ("T_hp", C.c_int32),
                ("T_pp", C.c_int32), ("T_ph", C.c_int32),
                ("np_phases", u8p), ("np_params", u8p), ("np_counts", i32p),
                ("hp_coeffs", u8p), ("hp_params", u8p),
                ("pp_psi_c", u8p), ("pp_psi_p", u8p), ("pp_phi_c", u8p), ("pp_phi_p", u8p),
                ("ph_alpha", u8p), ("ph_alpha_p", u8p), ("ph_beta", u8p), ("ph_beta_p", u8p),
                ("ph_counts", i32p),
                ("pf_phase_idx", u8p), ("pf_floatfactor", i32p), ("pf_power2", i32p),
                ("pf_approx_f32", f32p), ("pf_has_approx", C.c_int32)]


EXPORTS = {
    # name: (restype, argtypes)
    "ksim_component_create": (C.c_int, [C.POINTER(GraphView), C.c_int32, C.c_int32, C.POINTER(vp)]),
    "ksim_component_sample": (C.c_int, [vp, u8p, C.c_int64, C.c_int32, C.c_uint64, u8p]),
    "ksim_component_eval_abs": (C.c_int, [vp, C.c_int32, u8p, C.c_int64, C.c_int32, f64p]),
    "ksim_component_n_out": (C.c_int32, [vp]),
    "ksim_program_create": (C.c_int, [C.POINTER(GraphView), i32p, i32p, i32p, C.c_int32,
                                      C.c_int32, i32p, C.c_int32, C.POINTER(vp)]),
    "ksim_program_sample": (C.c_int, [vp, u8p, C.c_int64, C.c_int32, C.c_uint64, u8p]),
    "ksim_program_sample_uniforms": (C.c_int, [vp, u8p, C.c_int64, C.c_int32, f64p, u8p]),
    "ksim_program_eval_exact": (C.c_int, [vp, C.c_int32, C.c_int32, u8p, C.c_int64, C.c_int32,
                                          i64p, i32p]),
    "ksim_program_sample_device": (C.c_int, [vp, vp, C.c_int64, C.c_int32, C.c_int64, C.c_uint64,
                                             vp, vp, vp]),
    "ksim_program_set_tiles": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "ksim_program_split_hint": (C.c_int, [vp, C.c_int32, C.c_double, i32p,
                                          C.POINTER(C.c_double)]),
    "ksim_program_check": (C.c_int, [vp, vp]),
    "ksim_program_device_bytes": (C.c_int64, [vp]),
    "ksim_program_exact": (C.c_int32, [vp]),
    "ksim_program_stats": (C.c_int, [vp, f64p]),
    "ksim_program_destroy": (None, [vp]),
    "ksim_batch_create": (C.c_int, [C.POINTER(vp), C.c_int32, C.POINTER(vp), i32p, i64p,
                                    C.POINTER(C.c_uint64), C.POINTER(vp), C.c_int32, C.c_double,
                                    C.POINTER(vp)]),
    "ksim_batch_run": (C.c_int, [vp, C.c_int32, C.c_int32, vp]),
    "ksim_batch_check": (C.c_int, [vp]),
    "ksim_batch_destroy": (None, [vp]),
    "ksim_batch_sample_static": (C.c_int, [C.POINTER(vp), C.c_int32, C.POINTER(vp), i32p, i64p,
                                           C.POINTER(C.c_uint64), C.POINTER(vp), C.c_int32,
                                           C.c_double, C.c_int32, vp]),
    "ksim_sched_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(vp)]),
    "ksim_sched_set_slot": (C.c_int, [vp, C.c_int32, vp, vp, C.c_int32, C.c_int64, C.c_uint64, vp]),
    "ksim_sched_clear_slot": (C.c_int, [vp, C.c_int32]),
    "ksim_sched_launch": (C.c_int, [vp, i32p, C.c_int32, C.c_int32, vp]),
    "ksim_sched_check": (C.c_int, [vp, vp]),
    "ksim_sched_destroy": (None, [vp]),
    "ksim_program_pack": (C.c_int, [C.POINTER(GraphView), i32p, i32p, i32p, C.c_int32, C.c_int32,
                                    i32p, C.c_int32, C.POINTER(C.c_uint32), C.c_int64, i64p,
                                    C.POINTER(C.c_double)]),
    "ksim_tables": (C.c_int, [i32p, C.c_int64, i32p, C.c_int64, i32p]),
    "ksim_program_serialize": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8), C.c_int64, i64p]),
    "ksim_program_pack_blob": (C.c_int, [C.POINTER(GraphView), i32p, i32p, i32p, C.c_int32,
                                         C.c_int32, i32p, C.c_int32, C.POINTER(C.c_uint8),
                                         C.c_int64, i64p]),
    "ksim_blob_check": (C.c_int, [C.POINTER(C.c_uint8), C.c_int64, C.POINTER(C.c_double)]),
    "ksim_program_create_from_blob": (C.c_int, [C.POINTER(C.c_uint8), C.c_int64,
                                                C.POINTER(C.c_void_p)]),
    "ksim_program_info": (C.c_int, [C.c_void_p, i64p]),
    "ksim_noise_create": (C.c_int, [C.POINTER(C.c_double), i32p, C.POINTER(C.c_uint8), C.c_int32,
                                    C.c_int32, C.POINTER(C.c_void_p)]),
    "ksim_noise_info": (C.c_int, [C.c_void_p, i64p]),
    "ksim_noise_tables": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint32), i32p]),
    "ksim_noise_sample_device": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_uint64,
                                           C.c_void_p, C.c_int32, C.c_void_p]),
    "ksim_noise_sample": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_uint64,
                                    C.POINTER(C.c_uint8)]),
    "ksim_program_sample_noisy_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                                                   C.c_uint64, C.c_uint64, C.c_void_p,
                                                   C.c_void_p, C.c_void_p]),
    "ksim_program_sample_noisy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64,
                                            C.c_uint64, C.POINTER(C.c_uint8)]),
    "ksim_noise_destroy": (None, [C.c_void_p]),
    "ksim_program_sample_packed": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8), C.c_int64,
                                             C.c_int32, C.c_uint64, C.POINTER(C.c_uint32)]),
    "ksim_program_sample_device_packed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64,
                                                    C.c_int32, C.c_int64, C.c_uint64,
                                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "ksim_last_error": (C.c_char_p, []),
    "ksim_device_info": (C.c_int, [i64p]),
}

_lib = None


def load() -> C.CDLL:
    """Load the library (once); raise ImportError if it was not built.

    PyTorch ships its own HIP and HSA runtimes.  When torch is importable it
    is imported first, so the library binds to the runtime torch already
    loaded (same soname) and the process holds one HIP runtime; loaded the
    other way round, torch would bring a second runtime that cannot open
    the device."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m ksim_amd.build` "
            "(there is no non-HIP fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class KsimError(RuntimeError):
    """A device-side or runtime failure of the native library."""


def check(rc: int) -> None:
    if rc == KSIM_OK:
        return
    msg = load().ksim_last_error().decode(errors="replace")
    if rc == KSIM_E_INVALID:
        raise ValueError(msg)
    if rc == KSIM_E_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise KsimError(msg)


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_view(d: dict, keep: list) -> GraphView:
    """GraphView over the arrays of a graph dict (as made by graph_to_dict);
    the arrays are appended to ``keep`` so they outlive the call."""
    from .program import graph_shape
    G, P, Tnp, Thp, Tpp, Tph = graph_shape(d)
    v = GraphView()
    v.G, v.P, v.T_np, v.T_hp, v.T_pp, v.T_ph = G, P, Tnp, Thp, Tpp, Tph
    for name, ctype in (("np_phases", C.c_uint8), ("np_params", C.c_uint8),
                        ("np_counts", C.c_int32), ("hp_coeffs", C.c_uint8),
                        ("hp_params", C.c_uint8), ("pp_psi_c", C.c_uint8),
                        ("pp_psi_p", C.c_uint8), ("pp_phi_c", C.c_uint8),
                        ("pp_phi_p", C.c_uint8), ("ph_alpha", C.c_uint8),
                        ("ph_alpha_p", C.c_uint8), ("ph_beta", C.c_uint8),
                        ("ph_beta_p", C.c_uint8), ("ph_counts", C.c_int32),
                        ("pf_phase_idx", C.c_uint8), ("pf_floatfactor", C.c_int32),
                        ("pf_power2", C.c_int32), ("pf_approx_f32", C.c_float)):
        dt = {C.c_uint8: np.uint8, C.c_int32: np.int32, C.c_float: np.float32}[ctype]
        a = np.ascontiguousarray(d[name], dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dtype=dt)
        keep.append(a)
        setattr(v, name, ptr(a, ctype))
    v.pf_has_approx = 1 if d["pf_has_approx"] else 0
    return v
