"""Build the CPU checkers under oracle/ (test infrastructure).

* ``oracle/libkokkos_port.so`` — the C++/OpenMP restatement of the
  reference's native sampler (kokkos_port.cpp), the CPU baseline.
* ``oracle/libexact_fast.so`` — the exact oracle's graph values in C
  (exact_fast.c), the checker for programs too large for oracle/exact.py.

The reference's own native sampler (src/cpp/ksim/kokkos_sim.cpp,
zx_eval.hpp) is not built: it needs Kokkos, nanobind and MPI
(src/cpp/CMakeLists.txt:8-11), none of which are available, so no
``oracle/_ref`` build exists.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PORT_LIB = os.path.join(HERE, "libkokkos_port.so")
EXACT_FAST_LIB = os.path.join(HERE, "libexact_fast.so")


def build_exact_fast() -> str:
    """``oracle/libexact_fast.so``: the exact oracle's graph values in C."""
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall",
           os.path.join(HERE, "exact_fast.c"), "-o", EXACT_FAST_LIB + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(EXACT_FAST_LIB + ".tmp", EXACT_FAST_LIB)
    return EXACT_FAST_LIB


def build() -> str:
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           os.path.join(HERE, "kokkos_port.cpp"), "-o", PORT_LIB + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(PORT_LIB + ".tmp", PORT_LIB)
    build_exact_fast()
    return PORT_LIB


def load_port():
    if not os.path.exists(PORT_LIB):
        build()
    from ksim_amd import _lib
    lib = C.CDLL(PORT_LIB)
    lib.kport_create.restype = C.c_void_p
    lib.kport_create.argtypes = [C.POINTER(_lib.GraphView), C.c_int32, C.c_int32]
    lib.kport_sample.restype = C.c_int
    lib.kport_sample.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_int64, C.c_uint64,
                                 C.POINTER(C.c_double), C.POINTER(C.c_uint8), C.c_int32]
    lib.kport_destroy.restype = None
    lib.kport_destroy.argtypes = [C.c_void_p]
    lib.kport_max_threads.restype = C.c_int32
    return lib


class PortComponent:
    """One component on the CPU port (same call shape as kokkos_sim.Component)."""

    def __init__(self, graph_dicts, num_f):
        from ksim_amd import _lib
        self.lib = load_port()
        self.keep = []
        self.views = (_lib.GraphView * len(graph_dicts))()
        for i, d in enumerate(graph_dicts):
            self.views[i] = _lib.make_view(d, self.keep)
        self.h = self.lib.kport_create(self.views, len(graph_dicts), int(num_f))
        self.n_out = len(graph_dicts) - 1
        self.num_f = int(num_f)

    def sample(self, f_sel, seed, uniforms=None, threads=0):
        import numpy as np
        from ksim_amd import _lib
        f_sel = np.ascontiguousarray(f_sel, dtype=np.uint8)
        B = f_sel.shape[0]
        out = np.zeros((B, max(self.n_out, 1)), np.uint8)
        U = None
        if uniforms is not None:
            uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
            U = _lib.ptr(uniforms, C.c_double)
        fp = _lib.ptr(f_sel, C.c_uint8) if f_sel.size else _lib.ptr(np.zeros(1, np.uint8), C.c_uint8)
        self.lib.kport_sample(self.h, fp, B, int(seed), U, _lib.ptr(out, C.c_uint8), int(threads))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.kport_destroy(self.h)
            self.h = None


if __name__ == "__main__":
    print(build())
