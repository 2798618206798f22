"""GPU: device-side noise.  The f rows the device draws equal the oracle's
restated walk bit for bit; sampling with the noise drawn inside the kernel
equals sampling the drawn rows; the rows have the exact f distribution of
the reference's ChannelSampler at 5 sigma over a million shots."""
import math

import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.noise import NoiseModel, channel_patterns, sample_circuit_program
from ksim_amd.sampler import HipProgramSampler
from oracle import noise_ref
from tests.test_noise import model_a

pytestmark = pytest.mark.gpu


def log(m):
    print(m, flush=True)


def model_b(num_f=100, n_ch=150, p=2e-3, seed=0):
    """Config-2-like: many 1- and 2-bit channels over ~100 f bits."""
    rng = np.random.default_rng(seed)
    probs, cols = [], []
    for c in range(n_ch):
        if c % 3 == 0:
            probs.append(np.array([1 - 3 * p, p, p, p]))
            k = 2
        else:
            probs.append(np.array([1 - p, p]))
            k = 1
        for _ in range(k):
            col = np.zeros(num_f, np.uint8)
            col[rng.choice(num_f, size=int(rng.integers(1, 4)), replace=False)] = 1
            cols.append(col)
    return probs, np.stack(cols, axis=1)


def test_device_rows_equal_the_oracle():
    for probs, T in (model_a(), model_b(num_f=40, n_ch=30, p=0.02)):
        nm = NoiseModel(probs, T, seed=11)
        t = nm.tables()
        for base in (0, 12345):
            log(f"noise rows C={t['C']} base={base}")
            got = nm.sample(3000, shot_base=base)
            want = noise_ref.walk_rows(t, 11, 3000, shot_base=base)
            assert np.array_equal(got, want)


def test_successive_calls_continue_the_stream():
    probs, T = model_a()
    nm = NoiseModel(probs, T, seed=5)
    a, b = nm.sample(1000), nm.sample(1000)
    assert np.array_equal(np.concatenate([a, b]), nm.sample(2000, shot_base=0))


def test_fused_noise_equals_sampling_the_rows():
    probs, T = model_b(num_f=30, n_ch=40, p=0.05, seed=2)
    nm = NoiseModel(probs, T, seed=21)
    progs = [synth.shape_program(n_q=6, t=4, F=30, n_out=8, seed=4),                  # contiguous fsel
             synth.shape_program(n_q=5, t=2, F=12, n_out=6, seed=5, n_components=3),  # multi-comp
             synth.product_program([synth.htkh_program(1, noisy_f=True),
                                    synth.shape_program(n_q=4, t=2, F=6, n_out=3, seed=6)])]
    for k, prog in enumerate(progs):
        s = HipProgramSampler(prog)
        assert s.min_num_f <= 30
        B = 2048
        rows = nm.sample(B, shot_base=0)
        want = s.sample(rows, 77)
        got = s.sample_noisy(nm, B, 77)
        log(f"fused noise program {k}")
        assert np.array_equal(got, want), k
        # a shot range further along the stream
        rows2 = nm.sample(500, shot_base=9000)
        import torch
        d_out = torch.zeros((500, s.num_outputs), dtype=torch.uint8, device="cuda")
        s.sample_noisy_device(nm, d_out, 77, shot_base=9000)
        s.check()
        # the kernel keys the sampler's stream by the shot index too
        from ksim_amd import _lib
        import ctypes as C
        d_f = torch.from_numpy(rows2).cuda()
        ref = torch.zeros_like(d_out)
        _lib.check(_lib.load().ksim_program_sample_device(
            s.handle, d_f.data_ptr(), 500, rows2.shape[1], 9000, 77, None, ref.data_ptr(),
            torch.cuda.current_stream().cuda_stream))
        s.check()
        assert torch.equal(d_out, ref), k


def test_million_shot_distribution():
    import torch
    probs, T = model_b()
    nm = NoiseModel(probs, T, seed=3)
    B = 1_000_000
    d = torch.zeros((B, nm.num_f), dtype=torch.uint8, device="cuda")
    nm.sample_device(d)
    torch.cuda.synchronize()
    P, counts, prow, nf = channel_patterns(probs, T)
    want = noise_ref.exact_marginals(probs, prow, counts, nf)
    got = d.double().mean(dim=0).cpu().numpy()
    sig = np.sqrt(want * (1 - want) / B)
    log(f"max z = {np.max(np.abs(got - want) / np.maximum(sig, 1e-300)):.2f}")
    assert np.all(np.abs(got - want) <= 5 * sig + 1e-12)
    for i, j in [(0, 1), (3, 50), (10, 99)]:
        w = noise_ref.exact_pair_parity(probs, prow, counts, i, j)
        g = float((d[:, i] ^ d[:, j]).double().mean())
        assert abs(g - w) <= 5 * math.sqrt(w * (1 - w) / B) + 1e-12


def test_sample_circuit_program_splits_det_obs():
    probs, T = model_b(num_f=12, n_ch=10, p=0.01)
    prog = synth.shape_program(n_q=6, t=2, F=12, n_out=7, seed=8)
    prog.num_detectors = 5
    det, obs = sample_circuit_program(prog, probs, T, 4000, seed=9)
    assert det.shape == (4000, 5) and obs.shape == (4000, 2)
    s = HipProgramSampler(prog)
    nm = NoiseModel(probs, T, seed=9)
    full = s.sample(nm.sample(4000), 9)
    assert np.array_equal(np.concatenate([det, obs], axis=1), full)
