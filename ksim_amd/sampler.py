"""Drop-in sampler objects over the HIP library.

* :class:`Component` — the native component object, same constructor and
  methods as the reference's ``kokkos_sim.Component``
  (``src/cpp/ksim/kokkos_sim.cpp:244-252``): ``Component(graphs, num_f)``,
  ``.sample(f_sel, seed)``, ``.eval_abs(graph, params)``, ``.n_out``.
* :class:`HipProgramSampler` — same constructor and ``sample(f_params, seed)``
  as ``KokkosProgramSampler`` (``src/python/ksim/sample/sampler.py:46-67``),
  available under that name too.  All components of a program run in one
  kernel launch; the per-component seed offset ``seed + 0x9E37*k`` and the
  ``argsort(output_order)`` column permutation of the reference are kept, so
  for a given seed it draws the same per-shot uniforms as the reference's GPU
  sampler.
* :func:`sample_flat` — the numpy reference sampler's call
  (``src/python/ksim/sample/evaluate.py:69-87``) run on the GPU: the
  uniforms are drawn from ``rng`` in ``sample_flat``'s order and fed to the
  kernel, so the shot stream is the one ``sample_flat`` produces.

Amplitudes are exact: every term and the sum over terms are computed in
Z[w] with integer arithmetic; only the final |amplitude| and the Bernoulli
test are float64 (one fixed sequence of correctly rounded operations).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .program import FlatProgram, graph_shape, graph_to_dict, validate_graph

_M64 = 1 << 64
SEED_STRIDE = 0x9E37  # per-component seed offset, src/python/ksim/sample/sampler.py:63


def _as_u8_2d(a, name: str) -> np.ndarray:
    """The reference's binding accepts only uint8, 2-D, C-contiguous arrays
    (nanobind ``ndarray<const uint8_t, ndim<2>, c_contig>``) and raises
    TypeError otherwise."""
    if not isinstance(a, np.ndarray) or a.dtype != np.uint8 or a.ndim != 2 \
            or not a.flags.c_contiguous:
        raise TypeError(f"{name}: expected a C-contiguous 2-D uint8 ndarray")
    return a


def device_u8_2d(t, name: str, cols: tuple | None = None):
    """Check a caller's device buffer before its address goes to a kernel:
    a 2-D, contiguous uint8 CUDA tensor (TypeError otherwise), with a column
    count among ``cols`` when given (ValueError otherwise): kernels take
    the row stride from the program, not from the tensor."""
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.uint8 or t.dim() != 2 \
            or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name}: expected a contiguous 2-D uint8 CUDA tensor")
    if cols is not None and int(t.shape[1]) not in cols:
        raise ValueError(f"{name}: {int(t.shape[1])} columns, expected one of {sorted(set(cols))}")
    return t


def _check_seed(seed: int, n_comp: int) -> int:
    if isinstance(seed, (bool, np.bool_)) or not isinstance(seed, (int, np.integer)):
        raise TypeError("seed must be an int")
    seed = int(seed)
    if seed < 0 or seed + SEED_STRIDE * max(n_comp - 1, 0) >= _M64:
        raise TypeError("seed out of the uint64 range")
    return seed


class _Native:
    """Owner of one native program handle."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.load().ksim_program_destroy(h)
            except Exception:
                pass


class Component:
    """One component on the device; mirrors ``kokkos_sim.Component``.

    ``graphs`` is the list of graph dicts (``_graph_to_dict`` /
    :func:`ksim_amd.program.graph_to_dict`), graph 0 the normaliser and graph
    i+1 the step for output i; ``num_f`` the f-parameter count
    (kokkos_sim.cpp:67-151).
    """

    def __init__(self, graphs: list, num_f: int):
        lib = _lib.load()
        graphs = list(graphs)
        if not graphs:
            raise ValueError("a component needs at least one graph")
        keep: list = []
        views = (_lib.GraphView * len(graphs))()
        for i, d in enumerate(graphs):
            validate_graph(d)
            views[i] = _lib.make_view(d, keep)
        h = C.c_void_p()
        _lib.check(lib.ksim_component_create(views, len(graphs), int(num_f), C.byref(h)))
        self._native = _Native(h)
        self._num_f = int(num_f)
        self._P = [graph_shape(d)[1] for d in graphs]
        self._n_out = len(graphs) - 1

    @property
    def n_out(self) -> int:
        return self._n_out

    def sample(self, f_sel: np.ndarray, seed: int) -> np.ndarray:
        """f_sel (B, F) uint8 -> (B, max(n_out, 1)) uint8 (kokkos_sim.cpp:188-240)."""
        f_sel = _as_u8_2d(f_sel, "f_sel")
        seed = _check_seed(seed, 1)
        B = f_sel.shape[0]
        # every byte of an n_out > 0 row is written by the call, so the rows
        # need no zero fill (a fresh 1M x 29 array filled first cost
        # 1.5-3 ms of a 6 ms call)
        out = (np.empty if self._n_out else np.zeros)((B, max(self._n_out, 1)), dtype=np.uint8)
        lib = _lib.load()
        _lib.check(lib.ksim_component_sample(self._native._h, _lib.ptr(f_sel, C.c_uint8), B,
                                             f_sel.shape[1], seed, _lib.ptr(out, C.c_uint8)))
        return out

    def eval_abs(self, graph: int, params: np.ndarray) -> np.ndarray:
        """|amplitude| of graph ``graph`` at rows params (B, P) -> float64 (B,)
        (kokkos_sim.cpp:157-186)."""
        params = _as_u8_2d(params, "params")
        B, P = params.shape
        out = np.zeros(B, dtype=np.float64)
        lib = _lib.load()
        _lib.check(lib.ksim_component_eval_abs(self._native._h, int(graph),
                                               _lib.ptr(params, C.c_uint8), B, P,
                                               _lib.ptr(out, C.c_double)))
        return out


def output_dest(order: np.ndarray) -> np.ndarray:
    """Result column of each component-major output: the inverse of
    argsort(output_order) (src/python/ksim/sample/sampler.py:65-67)."""
    order = np.asarray(order)
    n = order.shape[0]
    dest = np.empty(n, dtype=np.int32)
    dest[np.argsort(order)] = np.arange(n, dtype=np.int32)
    return dest


def algorithmic_ops(prog: FlatProgram) -> float:
    """Algorithmic 64-bit integer operations per shot of the reference
    formulation (one term = parity of every mask word, the family products
    and the prefactor multiply), used as the roofline work figure:

        ops(i,g) = (2 W_i + 1) M + 4 n_np + 12 n_ph + 12 + 28 + 4,
        M = n_np + T_hp + 2 T_pp + 2 n_ph,  W_i = ceil(P_i / 64)
        ops/shot = sum_c sum_i sum_g ops(i,g) + 32 * sum_c n_out_c
    """
    total = 0.0
    for comp in prog.components:
        for g in comp.graphs:
            G = g.pf_phase_idx.shape[0]
            if G == 0:
                continue
            P = g.np_params.shape[2] if g.np_params.ndim == 3 else 0
            W = (P + 63) // 64
            Thp = g.hp_coeffs.shape[1] if g.hp_coeffs.ndim == 2 else 0
            Tpp = g.pp_psi_c.shape[1] if g.pp_psi_c.ndim == 2 else 0
            n_np = np.asarray(g.np_counts, dtype=np.float64)
            n_ph = np.asarray(g.ph_counts, dtype=np.float64)
            M = n_np + Thp + 2 * Tpp + 2 * n_ph
            total += float(np.sum((2 * W + 1) * M + 4 * n_np + 12 * n_ph + 48))
        total += 32.0 * (len(comp.graphs) - 1)
    return total


class ProgramArgs:
    """The arrays ksim_program_create / ksim_program_pack take for a
    FlatProgram (graph views kept alive by this object)."""

    def __init__(self, prog: FlatProgram):
        comps = list(prog.components)
        self.n_comp = len(comps)
        self.num_outputs = int(prog.num_outputs)
        self.keep: list = []
        graphs = []
        gpc, fsel, Fpc = [], [], []
        num_f = 0
        for comp in comps:
            sel = np.asarray(comp.f_selection, dtype=np.int64).reshape(-1)
            if sel.size and sel.min() < 0:
                raise ValueError("negative f_selection index")
            num_f = max(num_f, int(sel.max()) + 1 if sel.size else 0)
            gds = [graph_to_dict(g) for g in comp.graphs]
            for d in gds:
                validate_graph(d)
            graphs.extend(gds)
            gpc.append(len(gds))
            fsel.extend(int(x) for x in sel)
            Fpc.append(int(sel.size))
        self.num_f = num_f
        self.tot_out = int(sum(n - 1 for n in gpc))
        order = np.asarray(prog.output_order)
        if order.shape[0] != self.tot_out:
            raise ValueError("output_order length != total component outputs")
        dest = output_dest(order) if self.tot_out else np.zeros(0, np.int32)
        if self.tot_out and int(dest.max()) >= self.num_outputs:
            raise ValueError("output_order has more entries than num_outputs")
        self.views = (_lib.GraphView * max(len(graphs), 1))()
        for i, d in enumerate(graphs):
            self.views[i] = _lib.make_view(d, self.keep)
        self.gpc = np.asarray(gpc, np.int32)
        self.fsel = np.asarray(fsel if fsel else [0], np.int32)
        self.Fpc = np.asarray(Fpc if Fpc else [0], np.int32)
        self.dest = np.asarray(dest if dest.size else [0], np.int32)

    def args(self):
        return (self.views, _lib.ptr(self.gpc, C.c_int32), _lib.ptr(self.fsel, C.c_int32),
                _lib.ptr(self.Fpc, C.c_int32), self.n_comp, self.num_f,
                _lib.ptr(self.dest, C.c_int32), self.num_outputs)


_STAT_KEYS = ("graphs", "terms", "records", "words", "ops_per_shot", "n2_bound", "wide_terms",
              "sum_bits")


def _stats_dict(st: np.ndarray) -> dict:
    d = {k: (float(v) if k == "ops_per_shot" else int(v)) for k, v in zip(_STAT_KEYS, st)}
    d["device_bytes"] = d["words"] * 4
    return d


def pack_words(prog: FlatProgram, stats: dict | None = None) -> np.ndarray:
    """The device program words of ``prog`` (host-only packing); ``stats``,
    when given, receives the packer's figures (as HipProgramSampler.stats)."""
    lib = _lib.load()
    pa = ProgramArgs(prog)
    n = C.c_int64()
    _lib.check(lib.ksim_program_pack(*pa.args(), None, 0, C.byref(n), None))
    w = np.zeros(max(n.value, 1), dtype=np.uint32)
    st = np.zeros(8, np.float64)
    _lib.check(lib.ksim_program_pack(*pa.args(), _lib.ptr(w, C.c_uint32), w.size, C.byref(n),
                                     _lib.ptr(st, C.c_double)))
    if stats is not None:
        stats.update(_stats_dict(st))
    return w[:n.value]


def unpack_outputs(packed: np.ndarray, num_outputs: int) -> np.ndarray:
    """(B, W) uint32 packed rows -> (B, num_outputs) uint8 (the layout of
    ``HipProgramSampler.sample``)."""
    packed = np.ascontiguousarray(packed, dtype="<u4")
    B = packed.shape[0]
    return np.unpackbits(packed.view(np.uint8).reshape(B, -1), axis=1, count=num_outputs,
                         bitorder="little")


def split_detectors(out: np.ndarray, num_detectors: int):
    """(detector bits, observable bits) of a result, as the reference's
    ``run`` returns them (src/python/ksim/run.py:75-76)."""
    nd = int(num_detectors)
    return out[:, :nd].astype(np.uint8), out[:, nd:].astype(np.uint8)


def program_blob(prog: FlatProgram) -> bytes:
    """Pack ``prog`` into the wire format on the host (no device needed)."""
    lib = _lib.load()
    pa = ProgramArgs(prog)
    n = C.c_int64()
    _lib.check(lib.ksim_program_pack_blob(*pa.args()[:8], None, 0, C.byref(n)))
    out = np.zeros(n.value, np.uint8)
    _lib.check(lib.ksim_program_pack_blob(*pa.args()[:8], _lib.ptr(out, C.c_uint8), out.size,
                                          C.byref(n)))
    return out.tobytes()


def blob_stats(blob: bytes) -> dict:
    """Validate a program blob and return its stats (ValueError if rejected),
    with the program header fields n_comp, num_f, num_outputs, tot_out and
    lane_words (the u32 fields at bytes 96-116 of the validated header)."""
    buf = np.frombuffer(blob, dtype=np.uint8)
    st = np.zeros(8, np.float64)
    _lib.check(_lib.load().ksim_blob_check(_lib.ptr(buf, C.c_uint8) if buf.size else None,
                                           buf.size, _lib.ptr(st, C.c_double)))
    d = _stats_dict(st)
    hdr = np.frombuffer(bytes(blob[96:116]), dtype="<u4")
    d.update(n_comp=int(hdr[0]), num_f=int(hdr[1]), num_outputs=int(hdr[2]),
             tot_out=int(hdr[3]), lane_words=int(hdr[4]))
    return d


def lookup_tables():
    """Materialisation tables of the kernels.

    Returns (tab (D, (2W+1)*32, 4) int32 digit planes, pitab (PI_MAX*PI_MAX, 4)
    int32, W, PI_MAX, D, G): entry ((n+W)*4+r)*8+s of plane d is digit d, in
    balanced base 2^G, of the coefficients of w^s (1+w)^r (1+sqrt2)^n.
    """
    lib = _lib.load()
    sizes = np.zeros(4, np.int32)
    _lib.check(lib.ksim_tables(None, 0, None, 0, _lib.ptr(sizes, C.c_int32)))
    W, PM, D, G = (int(x) for x in sizes)
    tab = np.zeros((D, (2 * W + 1) * 32, 4), np.int32)
    pit = np.zeros((PM * PM, 4), np.int32)
    _lib.check(lib.ksim_tables(_lib.ptr(tab, C.c_int32), tab.size, _lib.ptr(pit, C.c_int32),
                               pit.size, None))
    return tab, pit, W, PM, D, G


class HipProgramSampler:
    """GPU sampler for a whole FlatProgram; drop-in for KokkosProgramSampler.

    ``HipProgramSampler(prog).sample(f_params, seed)`` returns the
    (B, num_outputs) uint8 array in ``output_order`` (sampler.py:46-67).
    """

    def __init__(self, prog: FlatProgram):
        lib = _lib.load()
        pa = ProgramArgs(prog)
        h = C.c_void_p()
        _lib.check(lib.ksim_program_create(*pa.args()[:8], C.byref(h)))
        self._adopt(h, prog)

    def _adopt(self, h, prog):
        self.prog = prog
        self._native = _Native(h)
        info = np.zeros(8, np.int64)
        _lib.check(_lib.load().ksim_program_info(h, _lib.ptr(info, C.c_int64)))
        self.n_comp, self.min_num_f, self.num_outputs, self.tot_out = (int(x) for x in info[:4])
        self.out_words = int(info[7])
        self.ops_per_shot = self.stats()["ops_per_shot"]

    @classmethod
    def from_bytes(cls, blob: bytes) -> "HipProgramSampler":
        """Upload a program blob (``to_bytes`` / :func:`program_blob`)
        without re-packing; the blob is validated first.  ``prog`` is None."""
        buf = np.frombuffer(blob, dtype=np.uint8)
        h = C.c_void_p()
        _lib.check(_lib.load().ksim_program_create_from_blob(
            _lib.ptr(buf, C.c_uint8) if buf.size else None, buf.size, C.byref(h)))
        self = cls.__new__(cls)
        self._adopt(h, None)
        return self

    def to_bytes(self) -> bytes:
        """The program as a versioned, checksummed blob."""
        lib = _lib.load()
        n = C.c_int64()
        _lib.check(lib.ksim_program_serialize(self._native._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint8)
        _lib.check(lib.ksim_program_serialize(self._native._h, _lib.ptr(out, C.c_uint8),
                                              out.size, C.byref(n)))
        return out.tobytes()

    def _prep_f(self, f_params) -> np.ndarray:
        f = np.ascontiguousarray(f_params, dtype=np.uint8)
        if f.ndim != 2:
            raise ValueError("f_params must be 2-D (shots, num_f)")
        if f.shape[1] < self.min_num_f:
            raise IndexError("f_selection index out of range for f_params")
        return f

    def sample(self, f_params: np.ndarray, seed: int) -> np.ndarray:
        f = self._prep_f(f_params)
        seed = _check_seed(seed, self.n_comp)
        B = f.shape[0]
        out = np.empty((B, self.num_outputs), dtype=np.uint8)  # every byte is written
        _lib.check(_lib.load().ksim_program_sample(self._native._h, _lib.ptr(f, C.c_uint8), B,
                                                    f.shape[1], seed, _lib.ptr(out, C.c_uint8)))
        return out

    def sample_noisy(self, noise, shots: int, seed: int, noise_seed: int | None = None,
                     shot_base: int = 0) -> np.ndarray:
        """Sample ``shots`` shots with the f bits drawn inside the kernel from
        ``noise`` (a :class:`ksim_amd.noise.NoiseModel`, streams keyed by
        ``noise_seed``, default the model's seed): the result of ``sample`` on
        ``noise.sample(shots, shot_base)`` with no f matrix anywhere."""
        seed = _check_seed(seed, self.n_comp)
        B = int(shots)
        out = np.zeros((B, self.num_outputs), dtype=np.uint8)
        ns = noise.seed if noise_seed is None else int(noise_seed)
        if B and shot_base == 0:
            _lib.check(_lib.load().ksim_program_sample_noisy(
                self._native._h, noise.handle, B, seed, ns, _lib.ptr(out, C.c_uint8)))
        elif B:
            import torch
            d_out = torch.zeros((B, max(self.num_outputs, 1)), dtype=torch.uint8, device="cuda")
            self.sample_noisy_device(noise, d_out, seed, ns, shot_base)
            self.check()
            out = d_out.cpu().numpy()[:, :self.num_outputs]
        return out

    def sample_noisy_device(self, noise, d_out, seed: int, noise_seed: int | None = None,
                            shot_base: int = 0, stream=None):
        """Resident variant: d_out a (B, num_outputs) uint8 CUDA tensor."""
        import torch
        device_u8_2d(d_out, "d_out", (self.num_outputs, max(self.num_outputs, 1)))
        seed = _check_seed(seed, self.n_comp)
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        ns = noise.seed if noise_seed is None else int(noise_seed)
        _lib.check(_lib.load().ksim_program_sample_noisy_device(
            self._native._h, noise.handle, int(d_out.shape[0]), int(shot_base), int(seed), ns,
            None, d_out.data_ptr() if d_out.numel() else None, st))

    def check(self, stream=None):
        """Wait for the program's launches on ``stream`` and raise a device
        range error, if any."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().ksim_program_check(self._native._h, st))

    def sample_packed(self, f_params: np.ndarray, seed: int) -> np.ndarray:
        """``sample`` with bit-packed rows: (B, ceil(num_outputs/32)) uint32,
        bit c of a row = output column c; :func:`unpack_outputs` restores
        the u8 array."""
        f = self._prep_f(f_params)
        seed = _check_seed(seed, self.n_comp)
        B = f.shape[0]
        out = np.empty((B, self.out_words), dtype=np.uint32)  # every word is written
        if self.out_words:
            _lib.check(_lib.load().ksim_program_sample_packed(
                self._native._h, _lib.ptr(f, C.c_uint8), B, f.shape[1], seed,
                _lib.ptr(out, C.c_uint32)))
        return out

    def sample_uniforms(self, f_params: np.ndarray, uniforms: np.ndarray) -> np.ndarray:
        """Sample with fed uniforms (B, total outputs), columns
        component-major then output-major (sample_flat's draw order)."""
        f = self._prep_f(f_params)
        B = f.shape[0]
        U = np.ascontiguousarray(uniforms, dtype=np.float64)
        if U.shape != (B, self.tot_out):
            raise ValueError(f"uniforms must have shape {(B, self.tot_out)}")
        out = np.empty((B, self.num_outputs), dtype=np.uint8)  # every byte is written
        Up = _lib.ptr(U, C.c_double) if U.size else None
        _lib.check(_lib.load().ksim_program_sample_uniforms(
            self._native._h, _lib.ptr(f, C.c_uint8), B, f.shape[1], Up,
            _lib.ptr(out, C.c_uint8)))
        return out

    def eval_exact(self, comp: int, graph: int, params: np.ndarray):
        """Exact value of one graph at param rows (B, P): returns
        (coeffs, exp2) with coeffs a (B, 4) object array of Python ints and
        exp2 (B,) int, value = coeffs * 2**exp2 (not canonicalised)."""
        params = np.ascontiguousarray(params, dtype=np.uint8)
        if params.ndim != 2:
            raise ValueError("params must be 2-D")
        B, P = params.shape
        raw = np.zeros((B, 4, 2), dtype=np.int64)
        e = np.zeros(B, dtype=np.int32)
        _lib.check(_lib.load().ksim_program_eval_exact(
            self._native._h, int(comp), int(graph), _lib.ptr(params, C.c_uint8), B, P,
            _lib.ptr(raw, C.c_int64), _lib.ptr(e, C.c_int32)))
        coeffs = np.empty((B, 4), dtype=object)
        for r in range(B):
            for k in range(4):
                lo = int(raw[r, k, 0]) & ((1 << 64) - 1)
                hi = int(raw[r, k, 1])
                coeffs[r, k] = (hi << 64) | lo
        return coeffs, e.astype(np.int64)

    @property
    def handle(self):
        return self._native._h

    @property
    def exact(self) -> bool:
        """False when some graph carries approximate (complex64)
        floatfactors: its terms are then summed in float64 like the
        reference's (zx_eval.hpp:168-178) instead of exactly, and
        eval_exact refuses it."""
        return _lib.load().ksim_program_exact(self.handle) == 1

    def set_term_split(self, term_split: int) -> None:
        """Waves that share one 64-shot tile's terms in plain launches: 1, 2
        or 4, or 0 (default) to choose per launch (split only launches too
        small to fill the GPU).  Results do not depend on it."""
        _lib.check(_lib.load().ksim_program_set_tiles(self.handle, 256, int(term_split)))

    def split_hint(self, comp: int, tile_ops: float) -> tuple[int, float]:
        """(term split, per-wave tile cost in algorithmic ops) the persistent
        kernels use for component ``comp`` under a per-wave tile budget."""
        L = C.c_int32(0)
        cost = C.c_double(0.0)
        _lib.check(_lib.load().ksim_program_split_hint(self.handle, int(comp), float(tile_ops),
                                                       C.byref(L), C.byref(cost)))
        return int(L.value), float(cost.value)

    def stats(self) -> dict:
        st = np.zeros(8, dtype=np.float64)
        _lib.check(_lib.load().ksim_program_stats(self._native._h, _lib.ptr(st, C.c_double)))
        return _stats_dict(st)


KokkosProgramSampler = HipProgramSampler


def flat_uniforms(prog: FlatProgram, B: int, rng: np.random.Generator) -> np.ndarray:
    """One rng.random(B) per output, component-major: the draws of
    sample_flat (src/python/ksim/sample/evaluate.py:73-82)."""
    cols = [rng.random(B) for comp in prog.components for _ in range(len(comp.graphs) - 1)]
    return np.stack(cols, axis=1) if cols else np.zeros((B, 0))


def sample_flat(prog: FlatProgram, f_params: np.ndarray, rng: np.random.Generator,
                sampler: HipProgramSampler | None = None) -> np.ndarray:
    """``sample_flat(prog, f_params, rng)`` evaluated on the GPU."""
    s = sampler if sampler is not None else HipProgramSampler(prog)
    f = np.ascontiguousarray(f_params, dtype=np.uint8)
    U = flat_uniforms(prog, f.shape[0], rng)
    return s.sample_uniforms(f, U)


def run_sampler(flat: FlatProgram, f: np.ndarray, seed: int, backend: str = "gpu") -> np.ndarray:
    """The reference's backend switch (src/python/ksim/run.py:79-91), every
    branch on the GPU:

    * "gpu": ``KokkosProgramSampler(flat).sample(f, seed=seed)`` — the
      per-shot xorshift64 streams of the native sampler;
    * "cpu": ``sample_flat(flat, f, np.random.default_rng(seed))`` — the
      numpy sampler's PCG64 stream, drawn on the host in its order and fed to
      the kernel, so the result is the reference's CPU result;
    * "auto": the reference picks "gpu" when its native module imports; this
      library is the native module, so "auto" is "gpu".
    """
    if backend == "auto":
        backend = "gpu"
    if backend == "gpu":
        return np.asarray(HipProgramSampler(flat).sample(f, seed=seed))
    if backend == "cpu":
        return np.asarray(sample_flat(flat, f, np.random.default_rng(seed)))
    raise ValueError(f"unknown backend {backend!r} (expected gpu/cpu/auto)")
