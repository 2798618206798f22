"""Benchmark of the MI355X FlatProgram sampler.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--quick]

The headline (default) is the batched workload the metric is quoted on:
config 4, 64 programs (q 16-60, t 4-24, 1k-100k shots each, 3.41M shots in
all) sampled by one persistent batched kernel replayed from a hipGraph, with
the f matrices and output buffers already resident in HBM.  A step is one
batched launch over the whole mix; the timed region brackets exactly K steps
with torch.cuda.synchronize() on both sides.  ``value`` is the batch's
aggregate shots/s; beside it on the same JSON line:

* the plain path on the same 64 programs (one launch per program, one lane
  per shot, and with the automatic term split) and the batched/plain ratios;
* ``roofline``: algorithmic int64 ops/s of the batch against the VALU peak
  (kernel time from HIP events on the launch stream), the profiled HBM
  traffic and counter VALU issue fraction of k_batch from profiles/;
* ``cpu_baseline``: the C++/OpenMP restatement of the reference's native
  sampler on this host's cores, on a bounded sample of every program of the
  mix (per-program rates combined with the mix's shot counts);
* unless ``--quick``: plain-launch rows of configs 1-3, the T-count sweep
  (30 q, t = 0..40, 100k shots) and the config-5 serving run (Poisson
  stream, p50/p99 latency and aggregate shots/s).

Other workloads (``--workload``, one line each): config1, config2, config3
(plain launches; config2 adds the device-noise rows), config5 (serving).
Every workload is a synthetic FlatProgram of the stated shape (ksim.compile
is not available to build them from the circuits).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# MI355X peaks.  VALU: 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 lane-ops/s, the
# issue rate of 32-bit integer VALU ops (and of fp64 FMAs: the published
# 78.6 TFLOP/s fp64 vector peak counts 2 flops per FMA).  The roofline
# divides algorithmic ops (the reference formulation's 64-bit integer ops,
# counted by the packer) by this peak: one lane-op per algorithmic op is the
# bound this kernel can reach, since its bit-sliced masks and packed exponent
# adds spend about one 32-bit op per 64-bit op of the reference formulation.
VALU32_PEAK = 256 * 64 * 2.4e9
HBM_PEAK = 8.0e12


# HBM bytes per dispatch of the dominant kernel, measured by two rocprofv3
# --pmc passes (FETCH_SIZE, WRITE_SIZE, each scaled by the same counter on a
# copy kernel of known size) and recorded in these summaries
PROFILE_OF = {"config1": "k_sample_config1_plain_100k.md",
              "config2": "k_sample_config2_plain_1M.md",
              "config3": "k_sample_config3_plain_1M.md",
              "config4": "k_batch_config4_64prog.md",
              "shape:85,40,80,85,200000,7": "k_sample_split_t40q85_200k.md"}


def profiled(workload, key):
    """A figure (``key: value`` line) of the committed counter summary of
    this workload's dominant kernel, with the summary's path, or (None, None)."""
    name = PROFILE_OF.get(workload)
    if not name:
        return None, None
    path = os.path.join(HERE, "profiles", name)
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


def valu_roofline(ops_per_s, bytes_per_s, workload=None):
    """Roofline object of a VALU-bound launch: algorithmic ops/s against the
    VALU lane-op peak, with the HBM rate beside it; ``traffic`` is the
    profiled HBM bytes per dispatch and ``valu_issue_frac`` the profiled
    VALU instructions x 64 lanes / (kernel time x peak), the share of the
    VALU issue rate the kernel's executed instructions use."""
    traffic, src = profiled(workload, "traffic_bytes_per_dispatch")
    issue, _ = profiled(workload, "valu_issue_frac")
    return dict(bound="valu", achieved=ops_per_s, peak=VALU32_PEAK, unit="op/s",
                frac=ops_per_s / VALU32_PEAK, traffic=traffic, traffic_source=src,
                valu_issue_frac=issue,
                hbm_GBps=bytes_per_s / 1e9, hbm_frac=bytes_per_s / HBM_PEAK)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name: str):
    from ksim_amd import synth
    if name == "config1":
        return dict(prog=synth.config1_program(), B=100_000, F=0,
                    desc="5q random Clifford+T, 4 T, 5 outputs, plain launch")
    if name == "config2":
        return dict(prog=synth.config2_program(), B=1_000_000, F=100,
                    desc="d=5 rep-code memory 5 rounds, 9q Clifford, 29 outputs, F=100, plain launch")
    if name == "config3":
        return dict(prog=synth.config3_program(), B=1_000_000, F=24,
                    desc="30q brickwork encoder, 12 T, 30 outputs, F=24, plain launch")
    if name.startswith("shape:"):
        # shape:Q,T,F,NOUT,SHOTS[,SEED] -- one shape-model program, plain launch
        v = [int(x) for x in name.split(":")[1].split(",")]
        q, t, F, n_out, shots = v[:5]
        seed = v[5] if len(v) > 5 else 900 + t
        return dict(prog=synth.shape_program(n_q=q, t=t, F=F, n_out=n_out, seed=seed), B=shots,
                    F=F, desc=f"shape model q={q} t={t} F={F} outputs={n_out}, plain launch")
    raise ValueError(name)


class Plain:
    """One program sampled with ksim_program_sample_device on resident data."""

    def __init__(self, prog, B, F, p_f=1e-3, seed=0):
        import torch
        from ksim_amd import synth
        from ksim_amd.sampler import HipProgramSampler
        self.torch = torch
        self.s = HipProgramSampler(prog)
        self.B, self.F = B, F
        f = synth.random_f(B, F, p_f, 100 + seed) if F else np.zeros((B, 0), np.uint8)
        self.f = torch.from_numpy(f).cuda()
        self.out = torch.zeros((B, max(self.s.num_outputs, 1)), dtype=torch.uint8, device="cuda")
        self.ops = self.s.ops_per_shot
        self.bytes_per_shot = F + self.s.num_outputs

    def step(self):
        from ksim_amd import _lib
        st = self.torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().ksim_program_sample_device(
            self.s.handle, self.f.data_ptr() if self.F else None, self.B, self.F, 0, 12345,
            None, self.out.data_ptr(), st))

    def check(self):
        from ksim_amd import _lib
        _lib.check(_lib.load().ksim_program_check(
            self.s.handle, self.torch.cuda.current_stream().cuda_stream))


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, ev0.elapsed_time(ev1) / 1e3


def cpu_baseline_port(prog, F, shots, seed=0):
    """The C++/OpenMP restatement of the reference's native sampler on this
    host's cores, on a bounded sample of the same program."""
    from ksim_amd import synth
    from ksim_amd.program import graph_to_dict
    from oracle.build_ref import PortComponent
    comps = [PortComponent([graph_to_dict(g) for g in c.graphs], len(c.f_selection))
             for c in prog.components]
    threads = comps[0].lib.kport_max_threads() if comps else 1
    f = synth.random_f(shots, F, 1e-3, 200 + seed) if F else np.zeros((shots, 0), np.uint8)
    t0 = time.perf_counter()
    for k, (pc, c) in enumerate(zip(comps, prog.components)):
        pc.sample(np.ascontiguousarray(f[:, np.asarray(c.f_selection, np.int64)]), 7 + 0x9E37 * k)
    dt = time.perf_counter() - t0
    return dict(value=shots / dt, unit="shots/s", cores=int(threads), kind="port",
                us_per_shot=dt / shots * 1e6,
                sample=f"{shots} shots of the same program, C++/OpenMP restatement of "
                       f"kokkos_sim (u8 masks, thread per shot, float64 term sum)")


def cpu_baseline_mix(progs, Fs, shots_of, budget_s=15.0):
    """The C++/OpenMP port on every program of a mix: each program timed on
    a bounded sample (its share of ``budget_s`` by a probe of 32 shots), the
    mix's rate = total shots / sum of each program's shots x its measured
    time per shot."""
    probes = [cpu_baseline_port(p, F, 32, seed=k) for k, (p, F) in enumerate(zip(progs, Fs))]
    est = sum(pr["us_per_shot"] for pr in probes) * 32e-6  # seconds the probes took
    scale = max(1.0, budget_s / max(est, 1e-6))
    total_t, sampled, threads = 0.0, 0, 1
    for k, (p, F, pr) in enumerate(zip(progs, Fs, probes)):
        n = int(min(shots_of[k], max(32, 32 * scale)))
        r = cpu_baseline_port(p, F, n, seed=k)
        total_t += r["us_per_shot"] * 1e-6 * shots_of[k]
        sampled += n
        threads = r["cores"]
    total = float(sum(shots_of))
    return dict(value=total / total_t, unit="shots/s", cores=int(threads), kind="port",
                us_per_shot=total_t / total * 1e6,
                sample=f"{sampled} shots over all {len(progs)} programs (a bounded sample of "
                       f"each), rates weighted by the mix's shot counts; C++/OpenMP "
                       f"restatement of kokkos_sim (u8 masks, thread per shot, float64 term sum)")


def cpu_baseline_numpy(prog, F, shots):
    """numpy restatement of sample_flat, one process."""
    from ksim_amd import synth
    from oracle import numpy_ref
    f = synth.random_f(shots, F, 1e-3, 300) if F else np.zeros((shots, 0), np.uint8)
    t0 = time.perf_counter()
    numpy_ref.sample_flat(prog, f, np.random.default_rng(0))
    dt = time.perf_counter() - t0
    return dict(value=shots / dt, unit="shots/s", cores=1, kind="port",
                us_per_shot=dt / shots * 1e6, sample=f"{shots} shots, numpy sample_flat restatement")


def scale_cpu_shots(us_guess, budget_s=12.0, lo=200, hi=200_000):
    return int(min(hi, max(lo, budget_s / max(us_guess, 1e-3) * 1e6)))


def host_path(pl, reps=3):
    """End to end through the host-array API, the reference's calling
    convention: f from host memory, H2D, kernel, D2H of the output rows
    (u8 rows, and bit-packed rows), returned as numpy arrays.  Reported
    beside the resident rate, never as the headline value."""
    f = pl.f.cpu().numpy()
    res = {}
    for name, fn in (("u8_rows", pl.s.sample), ("packed_rows", pl.s.sample_packed)):
        fn(f, 12345)  # warm-up: staging buffers
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(f, 12345)
        dt = (time.perf_counter() - t0) / reps
        res[name] = dict(ms=dt * 1e3, shots_per_s=pl.B / dt)
    res["note"] = ("HipProgramSampler.sample / sample_packed with numpy f in and numpy rows "
                   "out (pageable host memory), mean of %d calls" % reps)
    return res


def plain_result(name, steps, warmup, with_cpu=True, term_split=0, with_host=True):
    import torch
    w = make_workload(name)
    log(f"[bench] {name}: building program")
    pl = Plain(w["prog"], w["B"], w["F"])
    pl.s.set_term_split(term_split)
    wall, dev = time_steps(pl.step, steps, warmup)
    pl.check()
    ms = wall / steps * 1e3
    shots_s = w["B"] * steps / wall
    kern = dev / steps
    achieved = pl.ops * w["B"] / kern
    res = dict(workload=name, desc=w["desc"], shots=w["B"], ms_per_step=ms,
               shots_per_s=shots_s, us_per_shot=1e6 / shots_s,
               kernel_ms=kern * 1e3, ops_per_shot=pl.ops, stats=pl.s.stats(),
               roofline=valu_roofline(achieved, pl.bytes_per_shot * w["B"] / kern, name))
    if with_host:
        res["host_path"] = host_path(pl)
    if with_cpu:
        # about 10-20 s of CPU work
        probe = cpu_baseline_port(w["prog"], w["F"], 200)
        n = scale_cpu_shots(probe["us_per_shot"], 12.0)
        log(f"[bench] {name}: CPU port baseline on {n} shots")
        res["cpu_baseline"] = cpu_baseline_port(w["prog"], w["F"], n)
        # the reference's other CPU path, numpy sample_flat (one core), on
        # about 6 s of work
        probe = cpu_baseline_numpy(w["prog"], w["F"], 100)
        n = scale_cpu_shots(probe["us_per_shot"], 6.0, lo=100, hi=100_000)
        log(f"[bench] {name}: numpy sample_flat baseline on {n} shots")
        res["cpu_baseline_numpy"] = cpu_baseline_numpy(w["prog"], w["F"], n)
    return res


def noise_result(prog, B, steps, warmup, with_cpu=True):
    """config2 with its noise: f bits drawn on the device, (a) fused into the
    sampling kernel, (b) as f rows then a plain launch; beside it the host
    path of the reference (numpy channel sampler restated, timed on a
    bounded sample, plus the H2D copy of the f rows)."""
    import torch
    from ksim_amd import _lib, synth
    from ksim_amd.noise import NoiseModel
    from ksim_amd.sampler import HipProgramSampler
    probs, T = synth.config2_noise()
    nm = NoiseModel(probs, T, seed=102)
    s = HipProgramSampler(prog)
    lib = _lib.load()
    out = torch.zeros((B, s.num_outputs), dtype=torch.uint8, device="cuda")
    f = torch.zeros((B, nm.num_f), dtype=torch.uint8, device="cuda")

    def fused():
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.ksim_program_sample_noisy_device(s.handle, nm.handle, B, 0, 12345, 102,
                                                        None, out.data_ptr(), st))

    def rows_then_sample():
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.ksim_noise_sample_device(nm.handle, B, 0, 102, f.data_ptr(), nm.num_f, st))
        _lib.check(lib.ksim_program_sample_device(s.handle, f.data_ptr(), B, nm.num_f, 0, 12345,
                                                  None, out.data_ptr(), st))

    def rows_only():
        st = torch.cuda.current_stream().cuda_stream
        _lib.check(lib.ksim_noise_sample_device(nm.handle, B, 0, 102, f.data_ptr(), nm.num_f, st))

    _, dev_f = time_steps(fused, steps, warmup)
    s.check()
    fused_out = out.clone()
    _, dev_r = time_steps(rows_then_sample, steps, warmup)
    s.check()
    same = bool(torch.equal(fused_out, out))
    _, dev_n = time_steps(rows_only, steps, warmup)
    res = dict(channels=len(probs), e_bits=int(T.shape[1]), f_bits=int(T.shape[0]), shots=B,
               fused_kernel_ms=dev_f / steps * 1e3, fused_shots_per_s=B * steps / dev_f,
               rows_then_sample_ms=dev_r / steps * 1e3, noise_rows_ms=dev_n / steps * 1e3,
               fused_equals_rows_then_sample=same)
    if with_cpu:
        from oracle import numpy_ref
        n = 200_000
        t0 = time.perf_counter()
        rows = numpy_ref.channel_sample(probs, T, n, np.random.default_rng(0))
        dt = time.perf_counter() - t0
        host = torch.from_numpy(np.zeros((B, nm.num_f), np.uint8)).pin_memory()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        f.copy_(host, non_blocking=True)
        torch.cuda.synchronize()
        h2d = time.perf_counter() - t1
        res["host_path"] = dict(
            kind="port", cores=1, sample=f"{n} shots, numpy restatement of ChannelSampler.sample",
            noise_us_per_shot=dt / n * 1e6, h2d_ms_for_all_shots=h2d * 1e3,
            noise_plus_h2d_ms_for_all_shots=dt / n * B * 1e3 + h2d * 1e3,
            mean_f_weight=float(rows.sum(axis=1).mean()))
    return res


def batched_result(steps, warmup, n_programs=64, tune=True, shape=None):
    """config4: the same 64 programs sampled (a) by one plain launch each,
    back to back on one stream, and (b) by one persistent batched launch
    replayed from a hipGraph; both with resident inputs."""
    import torch
    from ksim_amd import synth
    from ksim_amd.batch import BatchedSampler, device_info
    from ksim_amd.sampler import HipProgramSampler
    specs = synth.config4_specs(n_programs=n_programs)
    log(f"[bench] config4: building {len(specs)} programs")
    samplers, fs, seeds, progs_list = [], [], [], []
    for k, sp in enumerate(specs):
        progs_list.append(synth.program_from_spec(sp))
        samplers.append(HipProgramSampler(progs_list[-1]))
        fs.append(torch.from_numpy(synth.random_f(sp["shots"], sp["F"], 1e-3, 500 + k)).cuda())
        seeds.append(7000 + k)
    outs = [torch.zeros((sp["shots"], s.num_outputs), dtype=torch.uint8, device="cuda")
            for sp, s in zip(specs, samplers)]
    from ksim_amd import _lib
    lib = _lib.load()

    def plain_step():
        st = torch.cuda.current_stream().cuda_stream
        for s, f, o, sd in zip(samplers, fs, outs, seeds):
            _lib.check(lib.ksim_program_sample_device(s.handle, f.data_ptr() if f.numel() else None,
                                                      f.shape[0], f.shape[1], 0, sd, None,
                                                      o.data_ptr(), st))
    # launch shape from the autotuner (timed on this mix, outside the timed region)
    if shape is not None:
        # a given launch shape (wave_shots, split share or None, blocks per CU)
        from ksim_amd.autotune import split_budget
        ws, share, bpc = shape
        tops = split_budget(samplers, fs, device_info()["cus"], share)
        batch = BatchedSampler(samplers, fs, seeds, wave_shots=ws, blocks_per_cu=bpc,
                               tile_ops=tops)
        batch.tune = dict(wave_shots=ws, split_share=share, blocks_per_cu=bpc, tune_s=0.0)
    elif tune:
        batch = BatchedSampler.tuned(samplers, fs, seeds, use_cache=False, log=log)
    else:
        batch = BatchedSampler(samplers, fs, seeds)
        batch.tune = dict(wave_shots=batch.wave_shots,
                          blocks_per_cu=batch.blocks // device_info()["cus"], tune_s=0.0)
    total = batch.total_shots
    # the plain path: one launch per program, one lane per shot (the
    # reference's launch shape), then the same launches with the automatic
    # term split of small launches
    for s in samplers:
        s.set_term_split(1)
    wall_p, dev_p = time_steps(plain_step, steps, warmup)
    for s in samplers:
        _lib.check(lib.ksim_program_check(s.handle, torch.cuda.current_stream().cuda_stream))
        s.set_term_split(0)
    wall_a, dev_a = time_steps(plain_step, steps, warmup)
    for s in samplers:
        _lib.check(lib.ksim_program_check(s.handle, torch.cuda.current_stream().cuda_stream))
    wall_b, dev_b = time_steps(lambda: batch.run(use_graph=True), steps, warmup)
    batch.check()
    # same bits either way
    same = all(torch.equal(o, bo[:, :o.shape[1]]) for o, bo in zip(outs, batch.out))
    plain_sps = total * steps / wall_p
    auto_sps = total * steps / wall_a
    batch_sps = total * steps / wall_b
    kern_b = dev_b / steps
    achieved = batch.ops / kern_b
    res_progs = (progs_list, [sp["F"] for sp in specs], [sp["shots"] for sp in specs])
    return dict(programs=len(specs), shots=total, plain_shots_per_s=plain_sps,
                _mix=res_progs, batched_kernel_ms=kern_b * 1e3,
                batched_shots_per_s=batch_sps, speedup=batch_sps / plain_sps,
                plain_ms=wall_p / steps * 1e3, batched_ms=wall_b / steps * 1e3,
                plain_split_ms=wall_a / steps * 1e3, plain_split_shots_per_s=auto_sps,
                speedup_vs_plain_split=batch_sps / auto_sps,
                batched_equals_plain=bool(same), ops_total=batch.ops,
                tuned=dict(wave_shots=batch.tune["wave_shots"],
                           split_share=batch.tune.get("split_share", 8),
                           blocks_per_cu=batch.tune["blocks_per_cu"],
                           tune_s=batch.tune["tune_s"]),
                roofline=valu_roofline(achieved, batch.bytes / kern_b, "config4"))


def tcount_sweep(ts=(0, 4, 8, 12, 16, 20, 24, 32, 40), shots=100_000, n_q=30):
    """plain launch, n_q = 30, F = 2t, 30 outputs, 100k shots per point."""
    import torch
    from ksim_amd import synth
    rows = []
    for t in ts:
        prog = synth.shape_program(n_q=n_q, t=t, F=2 * t, n_out=n_q, seed=900 + t)
        try:
            pl = Plain(prog, shots, 2 * t, seed=t)
            wall, dev = time_steps(pl.step, 2, 1)
            pl.check()
        except Exception as e:  # report the point, keep the sweep going
            rows.append(dict(t=t, error=str(e)))
            log(f"[bench] T sweep t={t}: {e}")
            continue
        kern = dev / 2
        rows.append(dict(t=t, terms_mean=float(np.mean([g.pf_phase_idx.shape[0]
                                                         for g in prog.components[0].graphs])),
                         shots_per_s=shots * 2 / wall, us_per_shot=wall / 2 / shots * 1e6,
                         kernel_ms=kern * 1e3, valu_frac=pl.ops * shots / kern / VALU32_PEAK))
        log(f"[bench] T sweep t={t}: {rows[-1]['us_per_shot']:.4f} us/shot")
        del pl
        torch.cuda.empty_cache()
    return rows


def sparse_f(B, F, p, seed):
    """A (B, F) uint8 matrix of independent Bernoulli(p) bits for small p,
    drawn by placing Binomial(B F, p) ones at uniform positions (a position
    drawn twice holds one 1, a relative deficit of about B F p / 2 ones in
    B F), so a 1M x 80 request does not build a float64 matrix."""
    rng = np.random.default_rng(seed)
    f = np.zeros((B, F), np.uint8)
    if B * F:
        f.flat[rng.integers(0, B * F, size=int(rng.binomial(B * F, p)))] = 1
    return f


def serving_result(n_requests=200, n_pool=8, load=0.7, t_max=40, q_max=85, seed=5,
                   rate=None, fill_long=True):
    """config5: Poisson request stream; programs from the config-5 generator
    (t <= 40, q <= 85) drawn from a pool of resident programs whose first
    program is the generator's largest shape (t = 40, q = 85, F = 80, 85
    outputs), shots log-uniform in [100, 1M]; arrival rate = ``load`` x the
    saturation rate measured on the same stream (or ``rate`` requests/s).
    The latency of the
    heaviest requests is bounded below by their work at the saturation rate
    (``heaviest``)."""
    import torch
    from ksim_amd import synth
    from ksim_amd.sampler import HipProgramSampler
    from ksim_amd.serve import Request, ServingLoop, poisson_arrivals
    rng = np.random.default_rng(seed)
    specs = [synth.config5_spec(rng, t_max=t_max, q_max=q_max) for _ in range(n_pool)]
    specs[0].update(n_q=q_max, t=t_max, F=2 * t_max, n_out=q_max)
    log(f"[bench] config5: building {n_pool} programs (t = {[s['t'] for s in specs]}, "
        f"q = {[s['n_q'] for s in specs]})")
    pool = [HipProgramSampler(synth.program_from_spec(sp)) for sp in specs]
    reqs = []
    for i in range(n_requests):
        k = int(rng.integers(n_pool))
        shots = int(round(math.exp(rng.uniform(math.log(100), math.log(1_000_000)))))
        f = sparse_f(shots, specs[k]["F"], 1e-3, 10_000 + i)
        reqs.append(Request(key=k, prog=pool[k], f=f, seed=20_000 + i))
    work = np.array([pool[r.key].ops_per_shot * r.shots for r in reqs])
    total_ops = float(work.sum())
    # saturation: the whole stream at once (no arrival gaps)
    loop = ServingLoop(hbm_budget=96 << 30, max_slots=64, step_ms=30.0, fill_long=fill_long)
    warm = [Request(key=r.key, prog=r.prog, f=r.f[:64], seed=r.seed) for r in reqs[:8]]
    loop.run(warm, realtime=False, progress=True, max_wall_s=60)
    log(f"[bench] config5: saturation run over {n_requests} requests, {total_ops:.3g} ops")
    sat = loop.run([Request(key=r.key, prog=r.prog, f=r.f, seed=r.seed) for r in reqs],
                   realtime=False, progress=True, max_wall_s=240)
    rate_ops = total_ops / sat["wall_s"]
    lam = rate if rate else load * rate_ops / (total_ops / n_requests)
    for r, t in zip(reqs, poisson_arrivals(n_requests, lam, seed)):
        r.arrival = float(t)
    log(f"[bench] config5: saturation {sat['wall_s']:.2f}s for {n_requests} requests; "
        f"Poisson rate {lam:.2f}/s")
    st = loop.run(reqs, realtime=True, progress=True, max_wall_s=300)
    loop.close()
    lat = np.array([r.latency for r in reqs]) * 1e3
    top = np.argsort(work)[::-1][:3]
    return dict(requests=n_requests, pool=n_pool, load=load, arrival_rate_per_s=lam,
                pool_t=[sp["t"] for sp in specs], pool_q=[sp["n_q"] for sp in specs],
                pool_ops_per_shot=[p.ops_per_shot for p in pool],
                shots=st["shots"], p50_ms=st["p50_ms"], p90_ms=float(np.percentile(lat, 90)),
                p99_ms=st["p99_ms"], max_ms=float(lat.max()), mean_ms=st["mean_ms"],
                shots_per_s=st["shots_per_s"], wall_s=st["wall_s"],
                saturation_shots_per_s=sat["shots_per_s"], saturation_wall_s=sat["wall_s"],
                steps=st["steps"], algorithmic_ops_per_s=rate_ops,
                heaviest=[dict(program=int(reqs[i].key), t=specs[reqs[i].key]["t"],
                               q=specs[reqs[i].key]["n_q"], shots=reqs[i].shots,
                               latency_ms=float(lat[i]),
                               work_at_saturation_ms=float(work[i] / rate_ops * 1e3))
                          for i in top])


METRIC = "shots/s & \u00b5s/shot vs T-count, plain vs batched (exact); p50/p99 latency"


def plain_row(r):
    return {k: r[k] for k in ("shots", "shots_per_s", "us_per_shot", "kernel_ms", "ops_per_shot",
                              "roofline", "cpu_baseline", "cpu_baseline_numpy", "host_path",
                              "desc") if k in r}


def headline(args):
    """config 4 batched: the line the metric is quoted on."""
    r4 = batched_result(args.steps, args.warmup, tune=not args.no_tune,
                        shape=_shape(args.batch_shape))
    progs, Fs, shots_of = r4.pop("_mix")
    line = {
        "metric": METRIC,
        "value": r4["batched_shots_per_s"],
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r4["batched_ms"],
        "us_per_shot": 1e6 / r4["batched_shots_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "config4",
                   "desc": "64 programs (q 16-60, t 4-24, F = 2t, q outputs, 1k-100k shots "
                           "each), batched persistent kernel, hipGraph replay, resident f",
                   "programs": r4["programs"], "shots": r4["shots"], "tuned": r4["tuned"]},
        "plain": {"shots_per_s": r4["plain_shots_per_s"], "ms_per_step": r4["plain_ms"],
                  "launch": "one launch per program, one lane per shot"},
        "plain_split": {"shots_per_s": r4["plain_split_shots_per_s"],
                        "ms_per_step": r4["plain_split_ms"],
                        "launch": "one launch per program, automatic term split"},
        "speedup_vs_plain": r4["speedup"],
        "speedup_vs_plain_split": r4["speedup_vs_plain_split"],
        "batched_equals_plain": r4["batched_equals_plain"],
        "batched_kernel_ms": r4["batched_kernel_ms"],
        "ops_total": r4["ops_total"],
        "roofline": r4["roofline"],
    }
    if not args.no_cpu:
        log("[bench] config4: CPU port baseline on every program of the mix")
        line["cpu_baseline"] = cpu_baseline_mix(progs, Fs, shots_of)
    return line


def _shape(text):
    if not text:
        return None
    ws, share, bpc = text.split(",")
    return (int(ws), None if share.lower() == "none" else float(share), int(bpc))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="config4",
                    help="config4 (batched, the headline) | config1 | config2 | config3 | "
                         "config5 (serving) | shape:Q,T,F,NOUT,SHOTS[,SEED] (one plain launch)")
    ap.add_argument("--term-split", type=int, default=0,
                    help="plain workloads: waves per tile (0 = automatic, 1, 2, 4)")
    ap.add_argument("--quick", action="store_true",
                    help="config4: the headline line only (no config 1-3 rows, T-count "
                         "sweep or serving run)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="plain workloads: skip the host-array (PCIe) path timing")
    ap.add_argument("--no-tune", action="store_true",
                    help="config4: default launch shape instead of the autotuner's")
    ap.add_argument("--batch-shape", default=None,
                    help="config4: launch shape WAVE_SHOTS,SPLIT_SHARE|none,BLOCKS_PER_CU "
                         "instead of the autotuner's (e.g. 64,none,5)")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="config5: Poisson arrival rate (requests/s) instead of 0.7 x the "
                         "measured saturation rate")
    ap.add_argument("--serve-no-fill", action="store_true",
                    help="config5: steps sized by step_ms only (ServingLoop fill_long=False)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="launch one elementwise int32 copy kernel of this size first, so "
                         "counter runs can scale FETCH_SIZE / WRITE_SIZE against known bytes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py runs on one GPU only: this sampler build has no multi-GPU path "
                 "(shots shard with no exchange; see the design notes)")
    import torch
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    if args.calib_copy_mib:
        n = args.calib_copy_mib << 18  # int32 elements
        src = torch.ones(n, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)  # reads and writes n * 4 bytes
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {n * 4} bytes read, {n * 4} bytes written")
        del src, dst
    if args.workload == "config4":
        line = headline(args)
        if not args.quick:
            line["configs"] = {}
            for name in ("config1", "config2", "config3"):
                r = plain_result(name, max(3, args.steps // 2), 1, with_cpu=not args.no_cpu)
                if name == "config2":
                    w2 = make_workload("config2")
                    r["noise"] = noise_result(w2["prog"], w2["B"], max(3, args.steps // 2), 1,
                                              with_cpu=not args.no_cpu)
                line["configs"][name] = plain_row(r)
                if "noise" in r:
                    line["configs"][name]["noise"] = r["noise"]
            line["tcount_sweep"] = tcount_sweep()
            line["serving"] = serving_result()
        print(json.dumps(line), flush=True)
        return
    if args.workload == "config5":
        print(json.dumps({"metric": "config5 serving",
                          **serving_result(rate=args.serve_rate,
                                           fill_long=not args.serve_no_fill)}), flush=True)
        return
    res = plain_result(args.workload, args.steps, args.warmup, with_cpu=not args.no_cpu,
                       term_split=args.term_split, with_host=not args.no_host)
    if args.workload == "config2":
        w2 = make_workload("config2")
        res["noise"] = noise_result(w2["prog"], w2["B"], args.steps, args.warmup,
                                    with_cpu=not args.no_cpu)
    line = {
        "metric": f"{args.workload} plain launch",
        "value": res["shots_per_s"],
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "us_per_shot": res["us_per_shot"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": res["workload"], "desc": res["desc"], "shots": res["shots"],
                   "ops_per_shot": res["ops_per_shot"], "kernel_ms": res["kernel_ms"],
                   "program": res["stats"]},
        "roofline": res["roofline"],
        "cpu_baseline": res.get("cpu_baseline"),
        "cpu_baseline_numpy": res.get("cpu_baseline_numpy"),
    }
    if "noise" in res:
        line["noise"] = res["noise"]
    line["host_path"] = res.get("host_path")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
