"""Python interpreter of the packed device program (test helper).

Evaluates a graph of ``ksim_amd.sampler.pack_words(prog)`` with the same
record semantics and integer arithmetic as the gfx950 kernel
(ksim_amd/csrc/device.hpp), using the lookup tables the library exports.
Lets the host packer's factorisation be checked against the exact oracle on
a machine without a GPU.
"""
from __future__ import annotations

import numpy as np

# layout.hpp
PK_OV, PK_WV, PK_ON, PK_WN, PK_OA, PK_WA, PK_OB, PK_WB, PK_OS = 0, 9, 9, 8, 17, 6, 23, 6, 29
TH = dict(NPROJ=0, NPPX=1, NCHUNK=2, FLAGS=3, S=4, V=5, N=6, A=7, B=8, E=9, R0=10, SIZE=14)
TH_DWORDS = 16
CHUNK_HDR = 4
PM_GROUP, PM_HDR = 32, 4
PM_COUNT_MASK, PM_PAD_SHIFT = 0xFF, 8
GH_DWORDS = 16
CH = dict(NGRAPHS=0, F=1, NOUT=2, WMAX=3, COLBASE=4, FSEL=5, DEST=6, GRAPH=7)
PH_DWORDS = 8
TF_RESIDUAL, TF_PI = 1, 2
GF_APPROX = 4
GF_TABLE = 16
GH_FLAGS = 5


PM_TAB_HDR = 8


def chunk_walk(w, rp, W32, table):
    """(count, n_lin, n_pair, pad, first group offset, end offset) of the
    record chunk at rp; count is the number of sign records (mask form) or
    of parity groups (GF_TABLE graphs)."""
    nl, npr, nw = w[rp], w[rp + 1], w[rp + 3]
    npm, pad = nw & PM_COUNT_MASK, nw >> PM_PAD_SHIFT
    g = rp + CHUNK_HDR + pad
    if table:
        end = g + npm * (PM_TAB_HDR + 128 * W32)
    else:
        end = g + (npm + PM_GROUP - 1) // PM_GROUP * PM_HDR + npm * W32
    end += nl * (W32 + 1) + npr * (2 * W32 + 3)
    return npm, nl, npr, pad, g, end


def s32(x: int) -> int:
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x & 0x80000000 else x


def zmul(x, y):
    a1, b1, c1, d1 = x
    a2, b2, c2, d2 = y
    return (a1 * a2 + b1 * d2 - c1 * c2 + d1 * b2,
            a1 * b2 + b1 * a2 + c1 * d2 + d1 * c2,
            a1 * c2 + b1 * b2 + c1 * a2 - d1 * d2,
            a1 * d2 - b1 * c2 - c1 * b2 + d1 * a2)


def param_words(bits, W32):
    pv = [0] * W32
    for j, b in enumerate(bits):
        if int(b) & 1:
            pv[j >> 5] |= 1 << (j & 31)
    return pv


def popc_and(pv, words):
    return sum(bin(p & m).count("1") for p, m in zip(pv, words))


class Packed:
    def __init__(self, words: np.ndarray, tables):
        self.w = [int(x) for x in words]
        self.tab, self.pitab, self.W, self.PM, self.D, self.G = tables
        self.max_n2 = 0  # largest |n'| of a live term seen by eval_graph

    def comp(self, c):
        off = self.w[PH_DWORDS + c]
        return {k: self.w[off + v] for k, v in CH.items()}

    def graph_hdr(self, c, i):
        ch = self.comp(c)
        o = ch["GRAPH"] + i * GH_DWORDS
        return self.w[o:o + GH_DWORDS]

    def eval_graph(self, c, i, bits):
        """(coeffs, exp2) of graph i of component c at parameter bits."""
        gh = self.graph_hdr(c, i)
        e0 = s32(gh[2])
        acc = [0, 0, 0, 0]
        for X, sh in self.terms(c, i, bits):
            assert sh >= 0
            acc = [a + (v << sh) for a, v in zip(acc, X)]
        return tuple(acc), e0

    def magnitude(self, c, i, bits):
        """The float64 |amplitude| the kernels decide on: magnitude() of the
        exact sum, or for a GF_APPROX graph the float64 sum in term order of
        each term's cast times its complex64 factor (GH_APX_OFF table)."""
        import math
        import struct
        from oracle import exact
        gh = self.graph_hdr(c, i)
        e0 = s32(gh[2])
        if not gh[5] & GF_APPROX:
            return exact.magnitude(*self.eval_graph(c, i, bits))
        apx = gh[12]
        fr = fi = 0.0
        for t, (X, sh) in enumerate(self.terms(c, i, bits, keep_dead=True)):
            if X is None:
                continue
            re, im = exact.term_float(X, e0 + sh)
            ar = struct.unpack("<f", struct.pack("<I", self.w[apx + 2 * t]))[0]
            ai = struct.unpack("<f", struct.pack("<I", self.w[apx + 2 * t + 1]))[0]
            fr += re * ar - im * ai
            fi += re * ai + im * ar
        return math.sqrt(fr * fr + fi * fi)

    def terms(self, c, i, bits, keep_dead=False):
        """(value, shift) of every term of graph i at parameter bits: value
        * 2^(e0 + shift); a vanished term gives zero (None with keep_dead)."""
        w = self.w
        gh = self.graph_hdr(c, i)
        nterms, W32, off = gh[0], gh[1], gh[3]
        table = bool(gh[GH_FLAGS] & GF_TABLE)
        pv = param_words(bits, W32)
        th = off
        for _ in range(nterms):
            h = w[th:th + TH_DWORDS]
            S, V, N, A, B, E = (s32(h[TH[k]]) for k in ("S", "V", "N", "A", "B", "E"))
            rp = th + TH_DWORDS
            dead = 0
            for _ in range(h[TH["NPROJ"]]):
                c0 = w[rp + W32] + popc_and(pv, w[rp:rp + W32])
                dead |= c0
                rp += W32 + 1
            ex = 0
            for _ in range(h[TH["NPPX"]]):
                cs = w[rp + 2 * W32] + popc_and(pv, w[rp:rp + W32])
                cf = w[rp + 2 * W32 + 1] + popc_and(pv, w[rp + W32:rp + 2 * W32])
                ex ^= cs & cf
                rp += 2 * W32 + 2
            jd = 0
            for _ in range(h[TH["NCHUNK"]]):
                x = w[rp + 2]
                npm, nl, npr, _, rp, _ = chunk_walk(w, rp, W32, table)
                groups = []
                if table:
                    for _ in range(npm):
                        # parity word: XOR of the nibble tables' entries (16-byte aligned)
                        hdr = w[rp:rp + PM_TAB_HDR]
                        rp += PM_TAB_HDR
                        assert rp % 4 == 0
                        y = 0
                        for j in range(8 * W32):
                            y ^= w[rp + 16 * j + ((pv[j // 8] >> (4 * (j % 8))) & 15)]
                        rp += 128 * W32
                        groups.append((y, hdr))
                else:
                    for g0 in range(0, npm, PM_GROUP):
                        n = min(npm - g0, PM_GROUP)
                        hdr = w[rp:rp + PM_HDR] + [0, 0, 0, 0]
                        rp += PM_HDR
                        y = 0
                        for j in range(n):
                            if popc_and(pv, w[rp:rp + W32]) & 1:
                                y |= 1 << (PM_GROUP - n + j)
                            rp += W32
                        groups.append((y, hdr))
                for y, (sp, sm, tp, tm, xm, xc, jm, jc) in groups:
                    ny = ~y & 0xFFFFFFFF
                    cs = bin((y & sp) | (ny & sm)).count("1")
                    cn = bin((y & tp) | (ny & tm)).count("1")
                    x = (x + (cs << 30) + (cn << 9)) & 0xFFFFFFFF
                    z = y ^ xc
                    ex += bin(z & (z >> 1) & xm).count("1")
                    jd |= (y ^ jc) & jm
                for _ in range(nl):
                    if popc_and(pv, w[rp:rp + W32]) & 1:
                        x = (x + w[rp + W32]) & 0xFFFFFFFF
                    rp += W32 + 1
                for _ in range(npr):
                    pa = popc_and(pv, w[rp:rp + W32]) & 1
                    pb = popc_and(pv, w[rp + W32:rp + 2 * W32]) & 1
                    if pa:
                        x += w[rp + 2 * W32]
                    if pb:
                        x += w[rp + 2 * W32 + 1]
                    if pa and pb:
                        x += w[rp + 2 * W32 + 2]
                    x &= 0xFFFFFFFF
                    rp += 2 * W32 + 3
                V += (x >> PK_OV) & ((1 << PK_WV) - 1)
                N += (x >> PK_ON) & ((1 << PK_WN) - 1)
                A += (x >> PK_OA) & ((1 << PK_WA) - 1)
                B += (x >> PK_OB) & ((1 << PK_WB) - 1)
                S += x >> PK_OS
            S += (ex & 1) << 2
            dead |= 1 if jd else 0
            q, r = V >> 2, V & 3
            n2, s2 = N + 2 * q, (S + 2 * q) & 7
            if not -self.W <= n2 <= self.W:
                raise OverflowError("table range")
            if not dead & 1:
                self.max_n2 = max(self.max_n2, abs(n2))
            gi = ((n2 + self.W) * 4 + r) * 8 + s2
            X = tuple(sum(int(self.tab[d, gi, k]) << (self.G * d) for d in range(self.D))
                      for k in range(4))
            if h[TH["FLAGS"]] & TF_PI and (A or B):
                X = zmul(X, tuple(int(v) for v in self.pitab[A * self.PM + B]))
            if h[TH["FLAGS"]] & TF_RESIDUAL:
                X = zmul(X, tuple(s32(h[TH["R0"] + k]) for k in range(4)))
            if dead & 1:
                X = None if keep_dead else (0, 0, 0, 0)
            yield X, q + E
            th = th + TH_DWORDS + h[TH["SIZE"]]
