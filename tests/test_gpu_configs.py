"""GPU: the benchmark workloads of configs 1, 4 and 5 against the oracle.

* config 1 (5 q, 4 T, 5 outputs, 100k shots, plain launch): sampled shots
  from the whole run equal the oracle under the seeded per-shot stream, and
  fed uniforms give sample_flat's stream;
* config 4 (the 64-program mix at its full shot counts): the batched
  persistent kernel, with and without a term split and hipGraph replay,
  gives every program its plain-launch result, and shots of every program
  equal the oracle's under the same seeded stream;
* config 5 (serving): the two heaviest programs of the generator forced to
  t = 40, q = 85 (128-bit sums, multi-digit table rows) served in a mix with
  lighter requests: their exact graph values equal the oracle's and their
  shots equal the oracle's seeded stream.

The oracle is the C restatement (oracle/exact_fast.py) for the large
programs, the Python big-integer oracle for config 1; both are pinned on the
CPU (tests/test_oracle.py, tests/test_oracle_fast.py).
"""
import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.program import graph_to_dict
from ksim_amd.sampler import HipProgramSampler
from oracle import exact, exact_fast
from tests.seeded import seeded_columns

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def test_config1_full_run_equals_oracle():
    prog = synth.config1_program()
    s = HipProgramSampler(prog)
    B, seed = 100_000, 31
    f = np.zeros((B, 0), np.uint8)
    log("config1: seeded 100k shots")
    out = s.sample(f, seed)
    idx = np.sort(np.random.default_rng(1).choice(B, 300, replace=False))
    want = exact.sample_program(prog, f[idx], seeded_columns(prog, seed, idx))
    assert np.array_equal(out[idx], want)
    # every launch shape gives the same rows
    for L in (2, 4):
        s.set_term_split(L)
        assert np.array_equal(s.sample(f, seed), out), L
    s.set_term_split(0)
    log("config1: fed uniforms (sample_flat's stream)")
    U = exact.flat_uniforms(prog, 2000, np.random.default_rng(5))
    assert np.array_equal(s.sample_uniforms(f[:2000], U),
                          exact.sample_program(prog, f[:2000], U))


@pytest.fixture(scope="module")
def config4():
    import torch
    specs = synth.config4_specs()
    log(f"config4: building {len(specs)} programs")
    samplers, fs, seeds, progs = [], [], [], []
    for k, sp in enumerate(specs):
        p = synth.program_from_spec(sp)
        progs.append(p)
        samplers.append(HipProgramSampler(p))
        fs.append(torch.from_numpy(synth.random_f(sp["shots"], sp["F"], 0.05, 500 + k)).cuda())
        seeds.append(7000 + k)
    return specs, progs, samplers, fs, seeds


def _plain_outputs(samplers, fs, seeds):
    import torch
    from ksim_amd import _lib
    lib = _lib.load()
    outs = []
    for s, f, sd in zip(samplers, fs, seeds):
        o = torch.zeros((f.shape[0], s.num_outputs), dtype=torch.uint8, device="cuda")
        _lib.check(lib.ksim_program_sample_device(s.handle, f.data_ptr(), f.shape[0], f.shape[1],
                                                  0, sd, None, o.data_ptr(), None))
        _lib.check(lib.ksim_program_check(s.handle, None))
        outs.append(o)
    return outs


def test_config4_batch_equals_plain_and_oracle(config4):
    import torch
    from ksim_amd.autotune import split_budget
    from ksim_amd.batch import BatchedSampler, device_info
    specs, progs, samplers, fs, seeds = config4
    log("config4: plain launches")
    plain = _plain_outputs(samplers, fs, seeds)
    cus = device_info()["cus"]
    shapes = [(64, None, 5), (128, 8.0, 6), (64, 32.0, 5)]
    for ws, share, bpc in shapes:
        log(f"config4: batch wave_shots={ws} split_share={share} blocks_per_cu={bpc}")
        tops = split_budget(samplers, fs, cus, share)
        b = BatchedSampler(samplers, fs, seeds, wave_shots=ws, blocks_per_cu=bpc, tile_ops=tops)
        for use_graph in (False, True, True):
            for o in b.out:
                o.fill_(7)
            b.run(use_graph=use_graph)
            torch.cuda.synchronize()
            b.check()
            for k in range(len(specs)):
                assert torch.equal(b.out[k][:, :samplers[k].num_outputs], plain[k]), \
                    (ws, share, use_graph, k)
    log("config4: oracle on 6 shots of every program")
    rng = np.random.default_rng(9)
    for k, (sp, p) in enumerate(zip(specs, progs)):
        idx = np.sort(rng.choice(sp["shots"], 6, replace=False))
        f_np = fs[k][torch.from_numpy(idx).cuda()].cpu().numpy()
        want = exact_fast.sample_program(p, f_np, seeded_columns(p, seeds[k], idx))
        assert np.array_equal(plain[k][torch.from_numpy(idx).cuda()].cpu().numpy(), want), k


def _config5_heavy(n=2):
    """The n heaviest programs of 24 config-5 generator draws, forced to the
    config's largest shape (t = 40, q = 85)."""
    rng = np.random.default_rng(55)
    specs = [synth.config5_spec(rng) for _ in range(24)]
    specs.sort(key=lambda s: -(s["t"] * 100 + s["n_q"]))
    out = []
    for sp in specs[:n]:
        sp = dict(sp, t=40, n_q=85, F=80, n_out=85)
        out.append(sp)
    return out


def test_config5_heavy_programs_served_equal_oracle():
    from ksim_amd.serve import Request, ServingLoop
    heavy = _config5_heavy()
    log(f"config5: building {len(heavy)} programs at t=40, q=85")
    progs = [synth.program_from_spec(sp) for sp in heavy]
    samplers = [HipProgramSampler(p) for p in progs]
    for s in samplers:
        st = s.stats()
        log(f"config5: program stats {st}")
        assert st["sum_bits"] > 62  # 128-bit sums
    assert any(s.stats()["wide_terms"] > 0 for s in samplers)  # multi-digit rows
    # exact graph values (the 128-bit and multi-digit paths) against the oracle
    rng = np.random.default_rng(3)
    for p, s in zip(progs, samplers):
        comp = p.components[0]
        F = len(comp.f_selection)
        for i in (0, 1, 40, len(comp.graphs) - 1):
            d = graph_to_dict(comp.graphs[i])
            pv = rng.integers(0, 2, size=(3, F + i)).astype(np.uint8)
            coeffs, e = s.eval_exact(0, i, pv)
            want = exact_fast.graph_values(d, pv)
            for r in range(3):
                assert exact.canonical(tuple(int(x) for x in coeffs[r]), int(e[r])) == want[r], \
                    (i, r)
    # a served mix: the heavy requests beside lighter ones of the generator
    light = [synth.program_from_spec(dict(n_q=20, t=8, F=16, n_out=20, shots=0, seed=77 + j))
             for j in range(2)]
    light_s = [HipProgramSampler(p) for p in light]
    reqs = []
    shots = [3000, 1500, 5000, 700]
    pool = [(0, samplers[0], heavy[0]["F"]), (1, samplers[1], heavy[1]["F"]),
            (2, light_s[0], 16), (3, light_s[1], 16)]
    for i, (key, smp, F) in enumerate(pool):
        f = synth.random_f(shots[i], F, 0.3, 60 + i)
        reqs.append(Request(key=key, prog=smp, f=f, seed=9100 + i))
    log("config5: serving the mix")
    loop = ServingLoop(max_slots=8, step_ms=5.0, tile_shots=64, hbm_budget=8 << 30)
    loop.run(reqs, realtime=False, max_wall_s=120)
    loop.close()
    for r, (key, smp, F) in zip(reqs, pool):
        alone = smp.sample(r.f, r.seed)
        assert np.array_equal(r.out, alone), key
    log("config5: oracle on 2 shots of each heavy request")
    for r, p in zip(reqs[:2], progs):
        idx = np.sort(rng.choice(r.shots, 2, replace=False))
        want = exact_fast.sample_program(p, r.f[idx], seeded_columns(p, r.seed, idx))
        assert np.array_equal(r.out[idx], want)
