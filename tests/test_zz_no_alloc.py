"""GPU: the batched launch without allocation (ksim_batch_sample_static)
gives exactly what the allocating batch (ksim_batch_create + ksim_batch_run)
gives on the same inputs, on two calls in a row (the second smaller, so a
stale work list or cursor in the shared array would show), and rejects calls
beyond its fixed sizes before launching anything."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.batch import BatchedSampler, sample_static
from ksim_amd.sampler import HipProgramSampler
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def _mix():
    import torch
    progs = [synth.shape_program(n_q=5, t=4, F=3, n_out=5, seed=1),
             synth.shape_program(n_q=8, t=10, F=10, n_out=9, seed=2, n_components=2),
             synth.wide_exponent_program(seed=5),
             synth.shape_program(n_q=6, t=0, F=40, n_out=12, seed=3),
             synth.shape_program(n_q=12, t=14, F=20, n_out=12, seed=4)]
    shots = [700, 1300, 257, 5000, 900]
    samplers, fs, seeds = [], [], []
    for k, (p, B) in enumerate(zip(progs, shots)):
        s = HipProgramSampler(p)
        samplers.append(s)
        fs.append(torch.from_numpy(synth.random_f(B, max(s.min_num_f, 1), 0.3, k)).cuda())
        seeds.append(500 + k)
    return samplers, fs, seeds


def _reference(samplers, fs, seeds, tile_ops):
    import torch
    b = BatchedSampler(samplers, fs, seeds, wave_shots=64, tile_ops=tile_ops)
    b.run(use_graph=False)
    torch.cuda.synchronize()
    b.check()
    return [o.clone() for o in b.out]


def _static(samplers, fs, seeds, tile_ops):
    import torch
    outs = [torch.full((f.shape[0], max(s.num_outputs, 1)), 7, dtype=torch.uint8, device="cuda")
            for s, f in zip(samplers, fs)]
    sample_static(samplers, fs, seeds, outs, wave_shots=64, tile_ops=tile_ops)
    return outs


@pytest.mark.parametrize("tile_ops", [0.0, 1.0])
def test_static_batch_equals_allocating_batch(tile_ops):
    import torch
    samplers, fs, seeds = _mix()
    log(f"no-alloc batch: first call, tile_ops={tile_ops}")
    want = _reference(samplers, fs, seeds, tile_ops)
    got = _static(samplers, fs, seeds, tile_ops)
    for k in range(len(samplers)):
        assert torch.equal(got[k], want[k]), k
    # a few shots of every program against the oracle's seeded stream
    for k, s in enumerate(samplers):
        f = fs[k].cpu().numpy()
        idx = pick(f.shape[0], 4, k)
        assert np.array_equal(got[k].cpu().numpy()[idx, :s.num_outputs],
                              oracle_rows(s.prog, f, seeds[k], idx)), k
    # second call, smaller: three programs, fewer shots each
    log("no-alloc batch: second, smaller call")
    s2 = samplers[:3]
    f2 = [fs[0][:100], fs[1][:65], fs[2][:3]]
    d2 = [11, 12, 13]
    want2 = _reference(s2, f2, d2, tile_ops)
    got2 = _static(s2, f2, d2, tile_ops)
    for k in range(3):
        assert torch.equal(got2[k], want2[k]), k


def test_static_batch_rejects_sizes_beyond_its_array():
    import torch
    s = HipProgramSampler(synth.shape_program(n_q=5, t=2, F=3, n_out=4, seed=1))
    f = torch.zeros((10, 3), dtype=torch.uint8, device="cuda")
    o = torch.zeros((10, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="KSIM_STATIC_MAX_PROGS"):
        sample_static([s] * 257, [f] * 257, [1] * 257, [o] * 257)
    # 65537 items of 256 shots (real buffers: nothing would run past them)
    B = 65537 * 256
    fz = torch.zeros((B, 0), dtype=torch.uint8, device="cuda")
    s0 = HipProgramSampler(synth.htkh_program(1))
    oz = torch.zeros((B, 1), dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="KSIM_STATIC_MAX_ITEMS"):
        sample_static([s0], [fz], [1], [oz], wave_shots=64, tile_ops=1e300)
    # and the array still serves a valid call afterwards
    want = _reference([s], [f], [3], 0.0)
    got = _static([s], [f], [3], 0.0)
    assert torch.equal(got[0], want[0])
