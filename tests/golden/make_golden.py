"""Write golden vectors from the reference's own numpy sampler.

Run in a container that has the reference checkout (path below).  The
reference's ``ksim`` package ``__init__`` imports stim/pyzx_param, which are
absent, so the three numpy-only modules it needs are loaded by file path under
a stub package:

    src/python/ksim/program.py
    src/python/ksim/sample/exact_scalar.py
    src/python/ksim/sample/evaluate.py

For a handful of programs built by ``ksim_amd.synth`` it records

* ``evaluate_flat(graph, params)`` (complex128) for every graph on random
  parameter rows — pins the float64 per-graph amplitudes, including the
  reference's summation order;
* ``sample_flat(prog, f, np.random.default_rng(seed))`` — pins the exact
  shot stream for a seed (rng draw order and decision rule).

Output: ``tests/golden/golden_<name>.npz`` holding the program arrays, the
inputs and the reference outputs, plus the numpy version used.
"""
from __future__ import annotations

import importlib.util
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("KSIM_REFERENCE", "/root/reference/src/python/ksim")
sys.path.insert(0, REPO)

from ksim_amd import synth  # noqa: E402
from tests.golden.golden_io import save_program  # noqa: E402


def load_reference():
    pkg = types.ModuleType("ksim")
    pkg.__path__ = [REF]
    sys.modules["ksim"] = pkg
    sub = types.ModuleType("ksim.sample")
    sub.__path__ = [os.path.join(REF, "sample")]
    sys.modules["ksim.sample"] = sub

    def load(name, path):
        spec = importlib.util.spec_from_file_location(name, path)
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
        return mod

    load("ksim.program", os.path.join(REF, "program.py"))
    load("ksim.sample.exact_scalar", os.path.join(REF, "sample", "exact_scalar.py"))
    return load("ksim.sample.evaluate", os.path.join(REF, "sample", "evaluate.py"))


def programs():
    return {
        "htkh3": (synth.htkh_program(3), 0),
        "htkh_noisy": (synth.htkh_program(1, noisy_f=True), 1),
        "ghz5": (synth.single_program(synth.ghz_graphs(5)), 0),
        "split_product": (synth.split_program(
            synth.product_program([synth.htkh_program(1), synth.htkh_program(3, noisy_f=True)]),
            4, seed=7), 1),
        "shape_small": (synth.shape_program(n_q=6, t=4, F=6, n_out=6, seed=11), 6),
        "shape_two_comp": (synth.shape_program(n_q=5, t=6, F=8, n_out=7, seed=12,
                                               n_components=2), 8),
    }


def main():
    ev = load_reference()
    out_dir = HERE
    for name, (prog, num_f) in programs().items():
        rng = np.random.default_rng(2024)
        data = {}
        save_program(data, prog)
        # per-graph amplitudes on random parameter rows
        for ci, comp in enumerate(prog.components):
            for gi, g in enumerate(comp.graphs):
                P = g.np_params.shape[2]
                pv = rng.integers(0, 2, size=(16, P)).astype(np.uint8)
                data[f"eval_params_{ci}_{gi}"] = pv
                data[f"eval_out_{ci}_{gi}"] = ev.evaluate_flat(g, pv)
        # sampled streams for two seeds
        B = 512
        f = (rng.random((B, num_f)) < 0.3).astype(np.uint8)
        data["f_params"] = f
        for seed in (5, 6):
            with np.errstate(divide="ignore", invalid="ignore"):
                data[f"sample_seed_{seed}"] = ev.sample_flat(
                    prog, f, np.random.default_rng(seed)).astype(np.uint8)
        data["numpy_version"] = np.array(np.__version__)
        np.savez_compressed(os.path.join(out_dir, f"golden_{name}.npz"), **data)
        print("wrote", name)


if __name__ == "__main__":
    main()
