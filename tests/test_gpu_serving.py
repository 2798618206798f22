"""GPU: the serving loop.  Every request's output in a continuously batched
mix equals its output alone (a plain launch with the same seed), whatever
the tiling, step size and arrival order; latency figures are produced."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.sampler import HipProgramSampler
from ksim_amd.serve import Request, ServingLoop, poisson_arrivals
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu


def make_requests(n, seed):
    rng = np.random.default_rng(seed)
    progs = [synth.shape_program(n_q=5, t=4, F=3, n_out=5, seed=1),
             synth.shape_program(n_q=8, t=8, F=10, n_out=9, seed=2, n_components=2),
             synth.shape_program(n_q=6, t=0, F=40, n_out=12, seed=3),
             synth.product_program([synth.htkh_program(1), synth.htkh_program(3, noisy_f=True)])]
    samplers = [HipProgramSampler(p) for p in progs]
    reqs = []
    for i in range(n):
        k = int(rng.integers(len(progs)))
        shots = int(rng.choice([0, 1, 63, 64, 65, 300, 1000, 2500]))
        num_f = max(samplers[k].min_num_f, 1)
        f = synth.random_f(shots, num_f, 0.3, 50 + i)
        reqs.append(Request(key=k, prog=samplers[k], f=f, seed=900 + i))
    return reqs, samplers


@pytest.mark.parametrize("tile,step_ms,rate", [(64, 0.05, 2.5e13), (256, 1.0, 2.5e13),
                                               (64, 0.05, 1e6)])
def test_results_same_alone_and_in_mix(tile, step_ms, rate):
    # rate 1e6: a tiny per-step budget, so the heavier requests' tiles have
    # their terms split over 2 or 4 waves
    reqs, samplers = make_requests(24, 3)
    loop = ServingLoop(max_slots=8, step_ms=step_ms, tile_shots=tile, hbm_budget=1 << 30,
                       rate=rate)
    if rate < 1e10:
        assert {loop._split(s, c) for s in samplers for c in range(s.n_comp)} & {2, 4}
    stats = loop.run(reqs, realtime=False)
    assert stats["requests"] == len(reqs)
    for i, r in enumerate(reqs):
        alone = samplers[r.key].sample(r.f, r.seed)
        assert r.out is not None and np.array_equal(r.out, alone), (r.key, r.shots)
        # and a few served shots against the oracle's seeded stream
        idx = pick(r.shots, 4, i)
        if idx.size:
            assert np.array_equal(r.out[idx], oracle_rows(samplers[r.key].prog, r.f, r.seed,
                                                          idx)), (r.key, r.shots)
    loop.close()


def test_poisson_stream_latencies():
    reqs, _ = make_requests(16, 4)
    for r, t in zip(reqs, poisson_arrivals(len(reqs), 400.0, 5)):
        r.arrival = float(t)
    loop = ServingLoop(max_slots=4, step_ms=0.5, tile_shots=128)
    stats = loop.run(reqs, realtime=True)
    assert stats["p50_ms"] is not None and stats["p99_ms"] >= stats["p50_ms"] >= 0
    assert all(r.done >= r.arrival for r in reqs)
    loop.close()


def test_content_keyed_residency():
    # requests without a key: FlatPrograms, blobs and samplers of the same
    # program share one resident copy, and results match a plain launch
    from ksim_amd.sampler import program_blob
    prog_a = synth.shape_program(n_q=5, t=4, F=3, n_out=5, seed=1)
    prog_b = synth.shape_program(n_q=6, t=2, F=4, n_out=3, seed=2)
    sa, sb = HipProgramSampler(prog_a), HipProgramSampler(prog_b)
    forms = [prog_a, program_blob(prog_a), sa, prog_b, program_blob(prog_b)]
    reqs = []
    for i, p in enumerate(forms * 2):
        f = synth.random_f(200 + 37 * i, 4, 0.3, 70 + i)
        reqs.append(Request(key=None, prog=p, f=f, seed=40 + i))
    loop = ServingLoop(max_slots=4, step_ms=0.2, hbm_budget=1 << 30)
    loop.run(reqs, realtime=False)
    assert len(loop.programs) == 2
    for r in reqs:
        ref = sa if r.key == loop.content_key(program_blob(prog_a)) else sb
        assert np.array_equal(r.out, ref.sample(r.f, r.seed))
    loop.close()


def test_admission_counts_program_bytes():
    # a budget that holds one of two programs (with its f and output rows)
    # at a time: the second request is admitted only once the first has
    # finished and its idle program has been evicted
    from ksim_amd.sampler import program_blob, blob_stats
    prog_a = synth.shape_program(n_q=8, t=10, F=6, n_out=12, seed=21)
    prog_b = synth.shape_program(n_q=8, t=10, F=6, n_out=12, seed=22)
    words = [4 * int(blob_stats(program_blob(p))["words"]) for p in (prog_a, prog_b)]
    shots = 500
    rows = shots * (6 + 12)  # f columns + output columns per shot
    budget = max(words) + rows + max(words) // 2
    reqs = [Request(key="a", prog=prog_a, f=synth.random_f(shots, 6, 0.3, 1), seed=5),
            Request(key="b", prog=prog_b, f=synth.random_f(shots, 6, 0.3, 2), seed=6)]
    loop = ServingLoop(max_slots=4, step_ms=0.2, hbm_budget=budget)
    loop.run(reqs, realtime=False)
    assert reqs[1].admitted >= reqs[0].done
    assert list(loop.programs) == ["b"]
    for r, p in zip(reqs, (prog_a, prog_b)):
        assert np.array_equal(r.out, HipProgramSampler(p).sample(r.f, r.seed))
    loop.close()


def test_sched_rejects_requests_beyond_32bit_offsets():
    import ctypes as C
    import torch
    from ksim_amd import _lib
    s = HipProgramSampler(synth.shape_program(n_q=5, t=2, F=3, n_out=4, seed=1))
    lib = _lib.load()
    h = C.c_void_p()
    _lib.check(lib.ksim_sched_create(2, 16, C.byref(h)))
    f = torch.zeros((4, 3), dtype=torch.uint8, device="cuda")
    out = torch.zeros((4, 4), dtype=torch.uint8, device="cuda")
    try:
        with pytest.raises(ValueError, match="2\\^31"):
            _lib.check(lib.ksim_sched_set_slot(h, 0, s.handle, f.data_ptr(), 3, 1 << 31, 1,
                                               out.data_ptr()))
        _lib.check(lib.ksim_sched_set_slot(h, 0, s.handle, f.data_ptr(), 3, 4, 1, out.data_ptr()))
    finally:
        lib.ksim_sched_destroy(h)
