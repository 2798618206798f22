"""GPU: the batched persistent kernel gives every program exactly the result
it gets alone (plain launch, same seed), with and without hipGraph replay."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.batch import BatchedSampler
from ksim_amd.sampler import HipProgramSampler
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu


SPECS = [dict(n_q=5, t=4, F=3, n_out=5, seed=1, B=700),
             dict(n_q=8, t=8, F=10, n_out=9, seed=2, B=130),
             dict(n_q=6, t=0, F=40, n_out=12, seed=3, B=1000),
         dict(n_q=4, t=6, F=0, n_out=3, seed=4, B=65)]


def build(specs):
    import torch
    samplers, fs, seeds, alone = [], [], [], []
    for k, sp in enumerate(specs):
        prog = synth.shape_program(sp["n_q"], sp["t"], sp["F"], sp["n_out"], sp["seed"],
                                   n_components=1 + (k % 2))
        s = HipProgramSampler(prog)
        f = synth.random_f(sp["B"], sp["F"], 0.3, k)
        samplers.append(s)
        fs.append(torch.from_numpy(f).cuda())
        seeds.append(1000 + k)
        alone.append(s.sample(f, 1000 + k))
    return samplers, fs, seeds, alone


def check_same(b, samplers, alone, oracle=True):
    """Every program's batched rows equal its plain-launch rows and, on a
    few shots (first, last and random ones), the oracle's rows under the
    same seeded stream."""
    for k in range(len(samplers)):
        got = b.out[k].cpu().numpy()[:, :samplers[k].num_outputs]
        assert np.array_equal(got, alone[k]), k
        if oracle:
            s = samplers[k]
            f = b.f[k].cpu().numpy()
            idx = pick(f.shape[0], 8, k)
            assert np.array_equal(got[idx], oracle_rows(s.prog, f, b.seeds[k], idx)), k


def test_batch_equals_alone():
    import torch
    specs = SPECS
    samplers, fs, seeds, alone = build(specs)
    for wave_shots in (64, 128, 256):
        b = BatchedSampler(samplers, fs, seeds, wave_shots=wave_shots)
        for use_graph in (False, True, True):
            for o in b.out:
                o.fill_(7)
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            b.check()
            check_same(b, samplers, alone, oracle=use_graph and wave_shots == 64)


def test_autotuned_batch_equals_alone(tmp_path, monkeypatch):
    import torch
    from ksim_amd import autotune
    monkeypatch.setenv("KSIM_TUNE_CACHE", str(tmp_path / "tune.json"))
    samplers, fs, seeds, alone = build(SPECS)
    b = BatchedSampler.tuned(samplers, fs, seeds, reps=1, wave_shots=(64, 128),
                             blocks_per_cu=(2, 5), split_shares=(8, None))
    assert not b.tune["cached"] and len(b.tune["trials"]) == 8
    from ksim_amd.batch import device_info
    cus = device_info()["cus"]
    assert b.wave_shots in (64, 128) and b.blocks in (2 * cus, 5 * cus)
    b.run()
    torch.cuda.synchronize()
    b.check()
    check_same(b, samplers, alone)
    # second call: served from the cache file
    again = autotune.tune_batch(samplers, fs, seeds, reps=1, wave_shots=(64, 128),
                                blocks_per_cu=(2, 5), split_shares=(8, None))
    assert again["cached"] and again["wave_shots"] == b.wave_shots


# terms shared by 1, 2 or 4 waves: heavier programs (many terms per graph,
# several components, a partial last tile) under a tiny tile budget (every
# component split as far as its graphs allow) and with the split off
SPLIT_SPECS = [dict(n_q=10, t=14, F=12, n_out=10, seed=11, B=300),
               dict(n_q=8, t=10, F=6, n_out=7, seed=12, B=129),
               dict(n_q=6, t=0, F=20, n_out=8, seed=13, B=200),
               dict(n_q=12, t=16, F=30, n_out=6, seed=14, B=64)]


@pytest.mark.parametrize("tile_ops", [1.0, 1e300])
def test_term_split_batch_equals_alone(tile_ops):
    import torch
    samplers, fs, seeds, alone = build(SPLIT_SPECS)
    splits = {s.split_hint(c, tile_ops)[0] for s in samplers for c in range(s.n_comp)}
    if tile_ops < 1e100:
        assert {2, 4} & splits, splits  # the split paths are exercised
    else:
        assert splits == {1}
    for wave_shots in (64, 128):
        b = BatchedSampler(samplers, fs, seeds, wave_shots=wave_shots, tile_ops=tile_ops)
        for use_graph in (False, True, True):
            for o in b.out:
                o.fill_(7)
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            b.check()
            check_same(b, samplers, alone)


@pytest.mark.parametrize("tile_ops", [1.0, 1e300])
def test_batch_with_128bit_graphs_equals_alone(tile_ops):
    # a program whose graphs need 128-bit sums in the batch selects the
    # kernel with that path for the whole batch; results stay the same
    import torch
    # (the second reads its f columns in a shuffled order)
    from tests.test_gpu_parity import permute_selections
    progs = [synth.wide_exponent_program(seed=5),
             permute_selections(synth.wide_exponent_program(n_out=2, F=40, seed=7), 8),
             synth.shape_program(n_q=8, t=10, F=6, n_out=7, seed=12)]
    samplers, fs, seeds, alone = [], [], [], []
    for k, prog in enumerate(progs):
        s = HipProgramSampler(prog)
        B = (300, 129, 200)[k]
        f = synth.random_f(B, s.min_num_f, 0.3, 60 + k)
        samplers.append(s)
        fs.append(torch.from_numpy(f).cuda())
        seeds.append(3000 + k)
        alone.append(s.sample(f, 3000 + k))
    assert samplers[0].stats()["wide_terms"] > 0
    b = BatchedSampler(samplers, fs, seeds, wave_shots=64, tile_ops=tile_ops)
    for use_graph in (False, True):
        for o in b.out:
            o.fill_(7)
        b.run(use_graph=use_graph)
        torch.cuda.synchronize()
        b.check()
        check_same(b, samplers, alone)


def test_rejects_bad_device_buffers():
    # caller tensors are checked before their addresses reach a kernel
    import torch
    s = HipProgramSampler(synth.shape_program(n_q=5, t=2, F=3, n_out=4, seed=1))
    f = torch.zeros((10, 3), dtype=torch.uint8, device="cuda")
    with pytest.raises(TypeError):
        BatchedSampler([s], [f.to(torch.int32)], [1])
    with pytest.raises(TypeError):
        BatchedSampler([s], [torch.zeros((3, 10), dtype=torch.uint8, device="cuda").t()], [1])
    with pytest.raises(TypeError):
        BatchedSampler([s], [f.cpu()], [1])
    with pytest.raises(ValueError):
        BatchedSampler([s], [f], [1, 2])
    from ksim_amd.noise import NoiseModel
    from tests.test_gpu_noise import model_b
    probs, T = model_b(num_f=3, n_ch=4, p=0.1, seed=1)
    nm = NoiseModel(probs, T, seed=2)
    with pytest.raises(ValueError):
        s.sample_noisy_device(nm, torch.zeros((10, 7), dtype=torch.uint8, device="cuda"), 5)
    out = torch.zeros((10, 4), dtype=torch.uint8, device="cuda")
    s.sample_noisy_device(nm, out, 5)
    s.check()
    assert np.array_equal(out.cpu().numpy(), s.sample_noisy(nm, 10, 5))
