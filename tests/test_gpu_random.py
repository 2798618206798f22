"""GPU: seeded random programs around the 32-bit word boundaries of the
parameter vector (F and output counts of 31-33, 63-65, ...; one to three
components; components without outputs; f columns selected in a shuffled
order), through every plain entry point
at term splits 1, 2 and 4, with f bits drawn on the device from random
noise models, and batched in groups of eight, against the oracle and
against each other."""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.sampler import HipProgramSampler, unpack_outputs
from oracle import exact
from tests.seeded import oracle_rows, pick

pytestmark = pytest.mark.gpu

F_CHOICES = [0, 1, 5, 30, 31, 32, 33, 62, 63, 64, 65, 95, 96, 97, 120]
OUT_CHOICES = [1, 2, 7, 30, 31, 32, 33, 40, 64]


def random_programs(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        nc = int(rng.integers(1, 4))
        F = int(rng.choice(F_CHOICES))
        n_out = max(int(rng.choice(OUT_CHOICES)), nc)
        t = int(rng.integers(0, 11))
        prog = synth.shape_program(n_q=int(rng.integers(4, 10)), t=t, F=F, n_out=n_out,
                                   seed=int(rng.integers(1 << 30)), n_components=nc)
        if rng.random() < 0.3:
            for c in prog.components:  # f columns read in a shuffled order
                c.f_selection = rng.permutation(np.asarray(c.f_selection))
        B = int(rng.integers(1, 200))
        yield prog, B, float(rng.choice([1e-3, 0.3, 0.5])), rng


def random_noise(num_f, rng):
    """A noise model of 1- and 2-bit channels on random f columns."""
    from ksim_amd.noise import NoiseModel
    probs, cols = [], []
    for _ in range(int(rng.integers(1, 60))):
        k = int(rng.integers(1, 3))
        p = float(rng.choice([1e-3, 0.02, 0.1]))
        probs.append(np.array([1 - (2 ** k - 1) * p] + [p] * (2 ** k - 1)))
        for _ in range(k):
            col = np.zeros(num_f, np.uint8)
            col[rng.choice(num_f, size=int(rng.integers(1, min(4, num_f) + 1)), replace=False)] = 1
            cols.append(col)
    return NoiseModel(probs, np.stack(cols, axis=1), seed=int(rng.integers(1 << 30)))


@pytest.mark.parametrize("seed", [11, 12])
def test_random_programs(seed):
    import torch
    from ksim_amd.batch import BatchedSampler
    pool = []
    for prog, B, p_f, rng in random_programs(24, seed):
        s = HipProgramSampler(prog)
        f = synth.random_f(B, max(s.min_num_f, 1), p_f, B)
        U = rng.random((B, s.tot_out))
        nb = min(B, 16)
        want = exact.sample_program(prog, f[:nb], U[:nb])
        got = {}
        for L in (1, 2, 4):
            s.set_term_split(L)
            got[L] = (s.sample_uniforms(f, U), s.sample(f, 7), s.sample_packed(f, 7))
        assert np.array_equal(got[1][0][:nb], want)
        # the seeded stream too
        idx = pick(B, 8, B)
        assert np.array_equal(got[1][1][idx], oracle_rows(prog, f, 7, idx))
        for L in (2, 4):
            for j in range(3):
                assert np.array_equal(got[L][j], got[1][j]), (L, j)
        assert np.array_equal(unpack_outputs(got[1][2], s.num_outputs), got[1][1])
        if s.min_num_f > 0 and rng.random() < 0.5:
            # f bits drawn in the kernel from a random noise model equal
            # sampling the rows the same model draws
            nm = random_noise(s.min_num_f + int(rng.integers(0, 3)), rng)
            Bn = int(rng.integers(1, 600))
            rows = nm.sample(Bn, shot_base=0)
            for L in (1, 4):
                s.set_term_split(L)
                assert np.array_equal(s.sample_noisy(nm, Bn, 9), s.sample(rows, 9)), L
            s.set_term_split(1)
        pool.append((s, f, got[1][1]))
        if len(pool) == 8:
            tile_ops = float(rng.choice([1.0, 0.0, 1e300]))
            b = BatchedSampler([q[0] for q in pool], [torch.from_numpy(q[1]).cuda() for q in pool],
                               [7] * 8, wave_shots=64, tile_ops=tile_ops)
            b.run()
            torch.cuda.synchronize()
            b.check()
            for j, q in enumerate(pool):
                got_b = b.out[j].cpu().numpy()[:, :q[0].num_outputs]
                assert np.array_equal(got_b, q[2]), j
                idx = pick(q[1].shape[0], 4, j)
                assert np.array_equal(got_b[idx], oracle_rows(q[0].prog, q[1], 7, idx)), j
            pool = []
