"""Save/load a FlatProgram to/from a flat dict of numpy arrays (npz fixtures)."""
from __future__ import annotations

import numpy as np

from ksim_amd.program import FlatComponent, FlatProgram, FlatScalarGraph

_FIELDS = [f for f in FlatScalarGraph.__dataclass_fields__]


def save_program(data: dict, prog: FlatProgram) -> None:
    data["prog_num_outputs"] = np.array(prog.num_outputs)
    data["prog_num_detectors"] = np.array(prog.num_detectors)
    data["prog_output_order"] = np.asarray(prog.output_order, np.int32)
    data["prog_n_comp"] = np.array(len(prog.components))
    for ci, c in enumerate(prog.components):
        data[f"c{ci}_f_selection"] = np.asarray(c.f_selection, np.int32)
        data[f"c{ci}_output_indices"] = np.asarray(c.output_indices, np.int32)
        data[f"c{ci}_n_graphs"] = np.array(len(c.graphs))
        for gi, g in enumerate(c.graphs):
            for f in _FIELDS:
                data[f"c{ci}_g{gi}_{f}"] = np.asarray(getattr(g, f))


def load_program(data) -> FlatProgram:
    comps = []
    for ci in range(int(data["prog_n_comp"])):
        graphs = []
        for gi in range(int(data[f"c{ci}_n_graphs"])):
            kw = {f: data[f"c{ci}_g{gi}_{f}"] for f in _FIELDS}
            kw["pf_has_approx"] = bool(kw["pf_has_approx"])
            graphs.append(FlatScalarGraph(**kw))
        comps.append(FlatComponent(tuple(int(x) for x in data[f"c{ci}_output_indices"]),
                                   data[f"c{ci}_f_selection"], graphs))
    return FlatProgram(comps, data["prog_output_order"], int(data["prog_num_outputs"]),
                       int(data["prog_num_detectors"]))
