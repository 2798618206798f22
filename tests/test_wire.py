"""Host-only: the program wire format.  A blob holds exactly the packed
words, round-trips through validation, and every structural corruption a
kernel could follow out of bounds is rejected before upload."""
import os
import struct

import numpy as np
import pytest

from ksim_amd import synth
from ksim_amd.sampler import blob_stats, pack_words, program_blob

HDR = 128
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fnv1a64(b: bytes) -> int:
    h = 0xcbf29ce484222325
    for x in b:
        h ^= x
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def words_of(blob: bytes) -> np.ndarray:
    return np.frombuffer(blob[HDR:], dtype=np.uint32).copy()


def with_words(blob: bytes, w: np.ndarray) -> bytes:
    """Replace the words and re-seal the checksum, so only the structural
    validation can catch what was changed."""
    body = w.astype(np.uint32).tobytes()
    head = bytearray(blob[:HDR])
    head[16:24] = struct.pack("<Q", len(w))
    head[24:32] = struct.pack("<Q", fnv1a64(body))
    return bytes(head) + body


# blobs written by the version-4 packer (this repository before sign-record
# groups) for these programs: program_blob(make()) at that version
LEGACY = {
    "two_comp": (lambda: synth.shape_program(n_q=6, t=6, F=8, n_out=4, seed=3, n_components=2),
                 open(os.path.join(GOLDEN, "blob_v4_two_comp.bin"), "rb").read()),
    "wide": (synth.wide_exponent_program,
             open(os.path.join(GOLDEN, "blob_v4_wide.bin"), "rb").read()),
}


@pytest.fixture(scope="module")
def prog():
    return synth.shape_program(n_q=6, t=6, F=8, n_out=4, seed=3, n_components=2)


@pytest.fixture(scope="module")
def blob(prog):
    return program_blob(prog)


def test_blob_holds_the_packed_words(prog, blob):
    st = {}
    w = pack_words(prog, st)
    assert blob[:8] == b"KSIMPRG\0"
    assert struct.unpack("<II", blob[8:16]) == (6, HDR)
    assert np.array_equal(words_of(blob), w)
    assert struct.unpack("<Q", blob[24:32])[0] == fnv1a64(w.tobytes())
    got = blob_stats(blob)
    for k in ("graphs", "terms", "records", "words", "ops_per_shot", "sum_bits"):
        assert got[k] == st[k], k


def test_version_3_blobs_still_load(prog, blob):
    # version 3 differs from version 4 only in what it could not hold
    # (approximate factors, more than 8 words): its number is accepted
    b = bytearray(LEGACY["two_comp"][1])
    b[8:12] = struct.pack("<I", 3)
    assert blob_stats(bytes(b))["words"] == blob_stats(LEGACY["two_comp"][1])["words"]


def test_version_5_blobs_still_load():
    # version 5 is version 6 without parity tables: a mask-form blob under
    # its number loads and keeps its words
    b = bytearray(program_blob(synth.shape_program(n_q=20, t=12, F=70, n_out=8, seed=12)))
    from tests.devfmt import CH, GF_TABLE, GH_DWORDS, GH_FLAGS, PH_DWORDS
    w = words_of(bytes(b))
    c0 = int(w[PH_DWORDS])
    g0, ng = int(w[c0 + CH["GRAPH"]]), int(w[c0 + CH["NGRAPHS"]])
    assert not any(int(w[g0 + GH_DWORDS * i + GH_FLAGS]) & GF_TABLE for i in range(ng))
    b[8:12] = struct.pack("<I", 5)
    assert blob_stats(bytes(b))["words"] == len(w)


@pytest.mark.parametrize("name", list(LEGACY))
def test_version_4_blobs_evaluate_as_the_oracle(name):
    # blobs written by the version-4 packer: every LIN record with its own
    # delta, no sign-record groups.  They validate, and their words evaluate
    # (CPU interpreter of the device format) to the oracle's exact values
    from oracle import exact
    from tests.devfmt import Packed
    from ksim_amd.program import graph_to_dict
    from ksim_amd.sampler import lookup_tables
    make, b = LEGACY[name]
    assert struct.unpack("<I", b[8:12])[0] == 4
    words = words_of(b)
    blob_stats(b)
    assert not np.array_equal(words, pack_words(make())), "the current packer writes sign records"
    pk = Packed(words, lookup_tables())
    prog = make()
    rng = np.random.default_rng(5)
    for c, comp in enumerate(prog.components):
        for i, g in enumerate(comp.graphs):
            d = graph_to_dict(g)
            P = g.np_params.shape[2] if g.np_params.ndim == 3 else 0
            for _ in range(4):
                bits = rng.integers(0, 2, size=P)
                coeffs, e0 = pk.eval_graph(c, i, bits)
                assert exact.canonical(coeffs, e0) == exact.graph_value(d, bits), (name, c, i)


def test_sign_records_written_and_counted():
    # node phases 1 + w^phi (phi odd or 2 mod 4) become sign records: the
    # packed words of a program with such factors are smaller than one
    # (mask, delta) pair per LIN record would make them
    from tests.devfmt import CH, GF_TABLE, GH_DWORDS, GH_FLAGS, PH_DWORDS, TH, TH_DWORDS, chunk_walk
    w = [int(x) for x in pack_words(synth.shape_program(n_q=12, t=8, F=16, n_out=4, seed=2))]
    off = w[PH_DWORDS]
    n_pm = 0
    for i in range(w[off + CH["NGRAPHS"]]):
        gh = w[w[off + CH["GRAPH"]] + i * GH_DWORDS:][:GH_DWORDS]
        W32, th = gh[1], gh[3]
        for _ in range(gh[0]):
            h = w[th:th + TH_DWORDS]
            rp = th + TH_DWORDS + h[TH["NPROJ"]] * (W32 + 1) + h[TH["NPPX"]] * (2 * W32 + 2)
            for _ in range(h[TH["NCHUNK"]]):
                npm, nl, npr, _, _, rp = chunk_walk(w, rp, W32, bool(gh[GH_FLAGS] & GF_TABLE))
                assert nl + npr + npm <= 63
                n_pm += npm
            th += TH_DWORDS + h[TH["SIZE"]]
    assert n_pm > 0


def test_blob_of_every_golden_program(golden_files):
    from tests.golden.golden_io import load_program
    for fn in golden_files:
        b = program_blob(load_program(np.load(fn)))
        assert blob_stats(b)["words"] == (len(b) - HDR) // 4


@pytest.mark.parametrize("case", ["magic", "version", "truncated", "extra", "checksum", "header"])
def test_rejects_damaged_envelope(blob, case):
    b = bytearray(blob)
    if case == "magic":
        b[0] ^= 1
    elif case == "version":
        b[8] ^= 0x40  # a version number never written
    elif case == "truncated":
        b = b[:-4]
    elif case == "extra":
        b += b"\0\0\0\0"
    elif case == "checksum":
        b[HDR + 40] ^= 0x10
    elif case == "header":
        b[96] ^= 1  # n_comp field disagrees with the words
    with pytest.raises(ValueError):
        blob_stats(bytes(b))


def structural_cases(w):
    """(name, mutated words) pairs, each a corruption the validator must catch."""
    PH, CH, GH, TH = 8, 16, 16, 16
    c0 = int(w[PH])                       # component 0 header offset
    fsel, dest, graphs = int(w[c0 + 5]), int(w[c0 + 6]), int(w[c0 + 7])
    g1 = graphs + GH                      # graph 1 header
    t0 = int(w[g1 + 3])                   # its first term
    out = []

    def mut(name, idx, val):
        v = w.copy()
        v[idx] = val
        out.append((name, v))

    mut("lane words", 4, 129)
    mut("component offset", PH, len(w) + 5)
    mut("graph count", c0 + 0, 0)
    mut("f_selection", fsel, int(w[1]) + 3)
    mut("output column", dest, int(w[2]))
    mut("graph mask words", g1 + 1, 9)
    mut("graph width", g1 + 1, 2 if int(w[g1 + 1]) == 1 else 1)
    mut("term offset", g1 + 3, len(w) - 4)
    mut("term count", g1 + 0, int(w[g1 + 0]) + 1000)
    mut("term size", t0 + 14, int(w[t0 + 14]) + 1)
    mut("term flags", t0 + 3, 8)
    mut("chunk count", t0 + 2, int(w[t0 + 2]) + 1000)
    mut("graph flags", g1 + 5, 16)
    mut("term part order", g1 + 6, int(w[g1 + 0]) + 1)
    mut("term part offset", g1 + 9, int(w[g1 + 9]) + 1)
    ng = int(w[c0 + 0])
    multi = [graphs + GH * i for i in range(ng) if int(w[graphs + GH * i]) >= 2]
    mut("single-term flag", multi[0] + 5, int(w[multi[0] + 5]) | 1)
    return out


def test_rejects_structural_corruption(blob):
    w = words_of(blob)
    cases = structural_cases(w)
    assert len(cases) == 16
    for name, v in cases:
        with pytest.raises(ValueError):
            blob_stats(with_words(blob, v))
        # the unchanged words still pass (the re-sealing itself is sound)
    blob_stats(with_words(blob, w))


def test_rejects_approximate_graph_corruption():
    # approximate factors beyond the words, an exact-sum flag on an
    # approximate graph, and an approximate graph split into term parts
    prog = synth.with_approx_factors(synth.shape_program(n_q=6, t=6, F=5, n_out=4, seed=8), 1,
                                     frac=1.0)
    blob = program_blob(prog)
    w = words_of(blob)
    PH, GH = 8, 16
    c0 = int(w[PH])
    g1 = int(w[c0 + 7]) + GH
    assert int(w[g1 + 5]) & 4  # GF_APPROX
    assert int(w[g1 + 0]) >= 2
    cases = []
    for idx, val in ((g1 + 12, len(w) - 1), (g1 + 5, 4 | 2), (g1 + 6, 1)):
        v = w.copy()
        v[idx] = val
        cases.append(v)
    for v in cases:
        with pytest.raises(ValueError):
            blob_stats(with_words(blob, v))
    blob_stats(with_words(blob, w))


def first_table_chunk(w):
    """(graph header offset, offset of the first record chunk with parity
    groups) in the words of a program with parity tables."""
    PH, GH = 8, 16
    c0 = int(w[PH])
    g0, ng = int(w[c0 + 7]), int(w[c0])
    for i in range(ng):
        gh = g0 + GH * i
        if not int(w[gh + 5]) & 16:  # GF_TABLE
            continue
        W32, th = int(w[gh + 1]), int(w[gh + 3])
        for _ in range(int(w[gh])):
            rp = th + 16 + int(w[th]) * (W32 + 1) + int(w[th + 1]) * (2 * W32 + 2)
            if int(w[th + 2]) and int(w[rp + 3]) & 0xFF:
                return gh, rp
            th += 16 + int(w[th + 14])
    raise AssertionError("no parity groups")


def test_parity_group_blobs():
    # a program that runs the 128-bit kernels carries parity tables: its blob
    # round-trips; tables off a 16-byte boundary, tables on a graph wider
    # than 4 mask words and tables in a program whose kernels hold no buffer
    # for them are each rejected
    prog = synth.shape_program(n_q=22, t=6, F=12, n_out=22, seed=4062)
    blob = program_blob(prog)
    w = words_of(blob)
    assert struct.unpack("<I", blob[8:12])[0] == 6
    gh, rp = first_table_chunk(w)
    assert (rp + 4 + (int(w[rp + 3]) >> 8) + 8) % 4 == 0
    blob_stats(with_words(blob, w))
    v = w.copy()
    v[rp + 3] ^= np.uint32(1 << 8)  # one dword more or less of padding
    with pytest.raises(ValueError):
        blob_stats(with_words(blob, v))
    v = w.copy()
    v[gh + 1] = 5  # W32 beyond the tables' limit (also the graph's own width)
    with pytest.raises(ValueError):
        blob_stats(with_words(blob, v))
    # an all-64-bit program carries tables only while its kernels hold a
    # buffer: at most 4 lane words (the same words claiming 5 are rejected)
    narrow = program_blob(synth.shape_program(n_q=8, t=6, F=6, n_out=7, seed=3))
    u = words_of(narrow)
    first_table_chunk(u)

    def lane_words(b, k):
        v = u.copy()
        v[4] = k  # PH_WMAX, and the header's copy of it
        out = bytearray(with_words(b, v))
        out[112:116] = struct.pack("<I", k)
        return bytes(out)
    blob_stats(lane_words(narrow, 4))
    with pytest.raises(ValueError, match="cannot read"):
        blob_stats(lane_words(narrow, 5))


def test_stream_flag_needs_its_padding():
    # the packer flags every multi-term graph GF_STREAM (record read-ahead)
    # and ends the words in 256 zero dwords; a flagged graph without that
    # padding past its last record is rejected
    prog = synth.shape_program(n_q=6, t=6, F=5, n_out=4, seed=8)
    blob = program_blob(prog)
    w = words_of(blob)
    PH, GH = 8, 16
    c0 = int(w[PH])
    g0, ng = int(w[c0 + 7]), int(w[c0])
    for i in range(ng):
        assert bool(int(w[g0 + GH * i + 5]) & 8) == (int(w[g0 + GH * i]) > 1), i
    assert not w[-256:].any()
    last = g0 + GH * (ng - 1)
    assert int(w[last]) >= 2 and int(w[last + 5]) & 8
    with pytest.raises(ValueError, match="read-ahead"):
        blob_stats(with_words(blob, w[:-1]))
    # without the flags the padding is not needed
    v = w[:-256].copy()
    for i in range(ng):
        v[g0 + GH * i + 5] &= ~np.uint32(8)
    blob_stats(with_words(blob, v))


def test_unpack_outputs_and_split():
    from ksim_amd.sampler import split_detectors, unpack_outputs
    rng = np.random.default_rng(0)
    out = rng.integers(0, 2, size=(9, 45)).astype(np.uint8)
    # pack as the kernel does: bit c of the row = column c, little-endian u32 words
    words = np.zeros((9, 2), np.uint32)
    for c in range(45):
        words[:, c // 32] |= out[:, c].astype(np.uint32) << np.uint32(c % 32)
    assert np.array_equal(unpack_outputs(words, 45), out)
    det, obs = split_detectors(out, 40)
    assert det.shape == (9, 40) and obs.shape == (9, 5)
    assert np.array_equal(np.concatenate([det, obs], axis=1), out)
