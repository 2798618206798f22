"""The C-ABI library loads and exports every function include/ksim_hip.h
declares (no device calls)."""
import ctypes
import os
import re

from ksim_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(REPO, "include", "ksim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(ksim_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ksim_component_create", "ksim_component_sample", "ksim_component_eval_abs",
                 "ksim_component_n_out", "ksim_program_create", "ksim_program_sample",
                 "ksim_program_sample_uniforms", "ksim_program_eval_exact",
                 "ksim_program_destroy", "ksim_last_error", "ksim_batch_create",
                 "ksim_batch_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(_lib.EXPORTS)


def test_last_error_is_a_string():
    lib = _lib.load()
    assert isinstance(lib.ksim_last_error(), bytes)
