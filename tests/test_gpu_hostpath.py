"""GPU: the host-array path of large calls runs as a pipeline of chunks
(pinned staging, upload / kernel / download overlapped on three streams).
Its rows equal one launch over resident buffers bit for bit, and seeded
rows equal the oracle, for byte rows, packed rows and fed uniforms, over a
ragged last chunk, extra f columns and a multi-component program with
permuted output columns.  The reference samples host arrays in one call
(src/cpp/ksim/kokkos_sim.cpp:188-240)."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.sampler import HipProgramSampler, unpack_outputs
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def resident(s, f, seed, U=None):
    """One launch over device buffers (ksim_program_sample_device)."""
    import torch
    from ksim_amd import _lib
    lib = _lib.load()
    df = torch.from_numpy(f).cuda()
    dU = torch.from_numpy(U).cuda() if U is not None else None
    out = torch.zeros((f.shape[0], s.num_outputs), dtype=torch.uint8, device="cuda")
    _lib.check(lib.ksim_program_sample_device(s.handle, df.data_ptr(), f.shape[0], f.shape[1],
                                              0, seed, dU.data_ptr() if dU is not None else None,
                                              out.data_ptr(), None))
    _lib.check(lib.ksim_program_check(s.handle, None))
    return out.cpu().numpy()


@pytest.mark.parametrize("B,F,extra", [(600_003, 40, 3), (400_000, 41, 0), (100_000, 10, 0)])
def test_pipelined_host_path_equals_one_launch(B, F, extra):
    # 600,003 x 43 and 400,000 x 41 bytes of f take the pipeline (chunks of
    # about 8 MiB, the last one ragged); 100,000 x 10 the single launch
    log(f"host path B={B} F={F} extra={extra}")
    prog = synth.shape_program(n_q=6, t=6, F=F, n_out=7, seed=B % 97, n_components=2)
    if len(prog.output_order) > 1:
        prog.output_order = np.random.default_rng(1).permutation(len(prog.output_order))
    s = HipProgramSampler(prog)
    f = synth.random_f(B, s.min_num_f + extra, 0.3, 4)
    want = resident(s, f, 321)
    got = s.sample(f, 321)
    assert np.array_equal(got, want)
    idx = pick(B, 6, 9)
    assert np.array_equal(got[idx], oracle_rows(prog, f, 321, idx))
    assert np.array_equal(unpack_outputs(s.sample_packed(f, 321), s.num_outputs), want)
    U = np.random.default_rng(5).random((B, s.tot_out))
    assert np.array_equal(s.sample_uniforms(f, U), resident(s, f, 0, U))


def test_pipelined_component_sample():
    # Component.sample (one component, f given as its selected columns):
    # 300,000 x 60 bytes take the pipeline; its first 100,000 rows equal a
    # call on those rows alone (one launch), and sampled rows the oracle
    from ksim_amd.program import graph_to_dict
    from ksim_amd.sampler import Component
    from oracle import exact
    prog = synth.shape_program(n_q=6, t=4, F=60, n_out=5, seed=3)
    gds = [graph_to_dict(g) for g in prog.components[0].graphs]
    kc = Component(gds, 60)
    B = 300_000
    f = synth.random_f(B, 60, 0.2, 6)
    got = kc.sample(f, 77)
    assert np.array_equal(got[:100_000], kc.sample(f[:100_000], 77))
    idx = pick(B, 6, 3)
    want = exact.sample_component(gds, f[idx], exact.kokkos_uniforms_at(77, idx, 5))
    assert np.array_equal(got[idx], want)


def test_pipelined_fed_uniforms_without_f_columns():
    # config-1 shape (F = 0): only the fed uniforms (1M x 5 float64, 40 MB)
    # make the call large enough for the pipeline
    prog = synth.config1_program()
    s = HipProgramSampler(prog)
    B = 1_000_000
    f = np.zeros((B, 0), np.uint8)
    U = np.random.default_rng(8).random((B, s.tot_out))
    got = s.sample_uniforms(f, U)
    assert np.array_equal(got, resident(s, f, 0, U))
    idx = pick(B, 8, 4)
    from oracle import exact
    assert np.array_equal(got[idx], exact.sample_program(prog, f[idx], U[idx]))
