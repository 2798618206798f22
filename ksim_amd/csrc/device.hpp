// device.hpp — gfx950 device code of the FlatProgram sampler.
//
// One lane evaluates one shot.  All program data a wave touches is the same
// for its 64 lanes (one component, one graph, one term, one record at a
// time), so records are read through scalar loads into SGPRs and each record
// costs a few VALU ops per lane: AND + popcount per 32-bit mask word, then a
// masked add of a packed exponent delta (see layout.hpp).  The exact term
// value is built from an LDS table, shifted to a per-graph exponent and added
// into four 128-bit accumulators, so the sum over terms is exact (the
// reference sums in float64, src/cpp/ksim/zx_eval.hpp:168-178).
#pragma once
#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace ksim {

typedef __int128 i128_t;
typedef unsigned __int128 u128_t;

struct Acc4 {
  i128_t a, b, c, d;
};

// ---------------------------------------------------------------- exact->f64
// Correctly rounded signed 128-bit -> double.  The top 64 significant bits
// are converted with a sticky bit OR-ed into bit 0, so the one rounding of the
// u64 -> f64 conversion sees every discarded bit.
__device__ __host__ inline double i128_to_double(i128_t x) {
  const bool neg = x < 0;
  const u128_t m = neg ? (u128_t)0 - (u128_t)x : (u128_t)x;
  const uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
  double r;
  if (hi == 0) {
    r = (double)lo;
  } else {
    const int s = 64 - __builtin_clzll(hi);  // 1..64
    uint64_t top, sticky;
    if (s == 64) {
      top = hi;
      sticky = lo != 0;
    } else {
      top = (hi << (64 - s)) | (lo >> s);
      sticky = (lo & ((1ull << s) - 1)) != 0;
    }
    r = __builtin_ldexp((double)(top | sticky), s);
  }
  return neg ? -r : r;
}

// |value * 2^e| by the fixed IEEE sequence
//   re = dbl(a) + dbl(b+d)*SQ, im = dbl(c) + dbl(b-d)*SQ, ldexp(sqrt(re^2+im^2), e)
// (the cast of zx_eval.hpp:168-169 applied once to the exact sum).
__device__ inline double magnitude(const Acc4 &v, int e) {
#pragma clang fp contract(off)
  const double SQ = 0.70710678118654752440;
  const double re = __dadd_rn(i128_to_double(v.a), __dmul_rn(i128_to_double(v.b + v.d), SQ));
  const double im = __dadd_rn(i128_to_double(v.c), __dmul_rn(i128_to_double(v.b - v.d), SQ));
  const double n2 = __dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im));
  return __builtin_ldexp(__dsqrt_rn(n2), e);
}

// ------------------------------------------------------------------- RNG
// Per-shot xorshift64 stream of the reference: seeding at
// src/cpp/ksim/kokkos_sim.cpp:215-217, step and 53-bit uniform at :43-46.
__device__ inline uint64_t rng_seed(uint64_t seed, uint64_t s) {
  uint64_t r = seed ^ (s * 0x9e3779b97f4a7c15ull);
  r ^= r >> 30;
  r *= 0xbf58476d1ce4e5b9ull;
  r ^= r >> 27;
  return r ? r : 1ull;
}
__device__ inline double rng_uniform(uint64_t &s) {
  s ^= s << 13;
  s ^= s >> 7;
  s ^= s << 17;
  return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}

// --------------------------------------------------------------- Z[w] x Z[w]
__device__ inline void zmul64(int64_t &a, int64_t &b, int64_t &c, int64_t &d, int64_t ya,
                              int64_t yb, int64_t yc, int64_t yd) {
  const int64_t na = a * ya + b * yd - c * yc + d * yb;
  const int64_t nb = a * yb + b * ya + c * yd + d * yc;
  const int64_t nc = a * yc + b * yb + c * ya - d * yd;
  const int64_t nd = a * yd - b * yc - c * yb + d * ya;
  a = na; b = nb; c = nc; d = nd;
}

// acc += x * 2^(32 L + r) for wave-uniform r < 32, L < 4 (mod 2^128)
__device__ __forceinline__ void add_shifted_uniform(i128_t &acc, int64_t x, int r, int L) {
  __builtin_assume(r >= 0 && r < 32);
  const u128_t y = (u128_t)(i128_t)x << r;  // 96 significant bits
  switch (L) {
    case 0: acc = (i128_t)((u128_t)acc + y); break;
    case 1: acc = (i128_t)((u128_t)acc + (y << 32)); break;
    case 2: acc = (i128_t)((u128_t)acc + (y << 64)); break;
    default: acc = (i128_t)((u128_t)acc + (y << 96)); break;
  }
}

// -------------------------------------------------------------- one graph
// Exact amplitude of one graph for this lane's parameter words pv[0..W32):
// acc = sum_terms value * 2^{-e0}; the caller scales by 2^{GH_E0}.
template <int W32, int WL>
__device__ __forceinline__ void eval_graph(const uint32_t *__restrict__ P,
                                           const uint32_t *__restrict__ gh,
                                           const uint32_t (&pv)[WL], const int4 *tab,
                                           const int4 *__restrict__ pitab, Acc4 &acc,
                                           uint32_t &err) {
  static_assert(W32 <= WL, "graph wider than the lane's parameter words");
  acc.a = acc.b = acc.c = acc.d = 0;
  const uint32_t nterms = gh[GH_NTERMS];
  const uint32_t *__restrict__ th = P + gh[GH_TERM_OFF];
  for (uint32_t t = 0; t < nterms; ++t) {
    const uint32_t nproj = th[TH_NPROJ], nppx = th[TH_NPPX], nchunk = th[TH_NCHUNK];
    const uint32_t flags = th[TH_FLAGS];
    int S = (int)th[TH_S], V = (int)th[TH_V], N = (int)th[TH_N];
    int A = (int)th[TH_A], Bq = (int)th[TH_B];
    const int E = (int)th[TH_E];
    const uint32_t *__restrict__ rp = th + TH_DWORDS;
    // PROJ: the term vanishes unless every projector parity matches
    uint32_t dead = 0;
    for (uint32_t k = 0; k < nproj; ++k, rp += W32 + 1) {
      uint32_t c = rp[W32];
#pragma unroll
      for (int w = 0; w < W32; ++w) c += __popc(pv[w] & rp[w]);
      dead |= c;
    }
    // PPX: (-1)^{psi phi}
    uint32_t ex = 0;
    for (uint32_t k = 0; k < nppx; ++k, rp += 2 * W32 + 2) {
      uint32_t cs = rp[2 * W32], cf = rp[2 * W32 + 1];
#pragma unroll
      for (int w = 0; w < W32; ++w) {
        cs += __popc(pv[w] & rp[w]);
        cf += __popc(pv[w] & rp[W32 + w]);
      }
      ex ^= cs & cf;
    }
    S += (int)((ex & 1u) << 2);
    // LIN / PAIR chunks: packed exponent adds
    for (uint32_t ch = 0; ch < nchunk; ++ch) {
      const uint32_t nl = rp[0], npr = rp[1];
      uint32_t x = rp[2];
      rp += CHUNK_HDR;
#pragma unroll 2
      for (uint32_t k = 0; k < nl; ++k, rp += W32 + 1) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < W32; ++w) c += __popc(pv[w] & rp[w]);
        x += (0u - (c & 1u)) & rp[W32];
      }
      for (uint32_t k = 0; k < npr; ++k, rp += 2 * W32 + 3) {
        uint32_t ca = 0, cb = 0;
#pragma unroll
        for (int w = 0; w < W32; ++w) {
          ca += __popc(pv[w] & rp[w]);
          cb += __popc(pv[w] & rp[W32 + w]);
        }
        const uint32_t ma = 0u - (ca & 1u), mb = 0u - (cb & 1u);
        x += (ma & rp[2 * W32]) + (mb & rp[2 * W32 + 1]) + (ma & mb & rp[2 * W32 + 2]);
      }
      V += (int)((x >> PK_OV) & ((1u << PK_WV) - 1));
      N += (int)((x >> PK_ON) & ((1u << PK_WN) - 1));
      A += (int)((x >> PK_OA) & ((1u << PK_WA) - 1));
      Bq += (int)((x >> PK_OB) & ((1u << PK_WB) - 1));
      S += (int)(x >> PK_OS);
    }
    // materialise w^S (1+w)^V (1+sqrt2)^N pi^A pibar^B R:
    // (1+w)^4 = 2 w^2 (1+sqrt2)^2, so with q = V>>2, r = V&3 the value is
    // 2^q w^{S+2q} (1+w)^r (1+sqrt2)^{N+2q} pi^A pibar^B R.
    const int q = V >> 2, r = V & 3;
    const int n2 = N + 2 * q;
    const int s2 = (S + 2 * q) & 7;
    int64_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;  // < 2^30 from the table
    if (n2 < -TAB_R || n2 > TAB_R) {
      err |= ERR_TAB_RANGE;
    } else {
      const int4 e = tab[((n2 + TAB_R) * 4 + r) * 8 + s2];
      x0 = e.x; x1 = e.y; x2 = e.z; x3 = e.w;
    }
    if ((flags & TF_PI) && (A | Bq)) {
      if (A >= PI_MAX || Bq >= PI_MAX) {
        err |= ERR_PI_RANGE;
        x0 = x1 = x2 = x3 = 0;
      } else {
        // (< 2^30) x (< 2^24) products, four per coefficient: exact in int64
        const int4 p = pitab[A * PI_MAX + Bq];
        zmul64(x0, x1, x2, x3, p.x, p.y, p.z, p.w);
      }
    }
    if (dead & 1u) x0 = x1 = x2 = x3 = 0;
    int sh = q + E;
    if (sh < 0 || sh > SHIFT_MAX) {
      if (x0 | x1 | x2 | x3) err |= ERR_SHIFT;
      x0 = x1 = x2 = x3 = 0;
      sh = 0;
    }
    if (flags & TF_RESIDUAL) {
      // a floatfactor that does not factor: one Z[w] multiply in 128 bits
      const i128_t R0 = (int32_t)th[TH_R0], R1 = (int32_t)th[TH_R1];
      const i128_t R2 = (int32_t)th[TH_R2], R3 = (int32_t)th[TH_R3];
      const i128_t y0 = x0 * R0 + x1 * R3 - x2 * R2 + x3 * R1;
      const i128_t y1 = x0 * R1 + x1 * R0 + x2 * R3 + x3 * R2;
      const i128_t y2 = x0 * R2 + x1 * R1 + x2 * R0 - x3 * R3;
      const i128_t y3 = x0 * R3 - x1 * R2 - x2 * R1 + x3 * R0;
      acc.a += (i128_t)((u128_t)y0 << sh);
      acc.b += (i128_t)((u128_t)y1 << sh);
      acc.c += (i128_t)((u128_t)y2 << sh);
      acc.d += (i128_t)((u128_t)y3 << sh);
    } else {
      // add X * 2^sh into the four 128-bit sums.  The shift is almost always
      // the same on every lane (it varies only through phase-pair records),
      // so the common path adds 32-bit limbs at a wave-uniform offset.
      const int sh0 = __builtin_amdgcn_readfirstlane(sh);
      if (dead & 1u) sh = sh0;  // a vanished term adds zero at any shift
      if (__builtin_amdgcn_ballot_w64(sh != sh0) == 0) {
        const int rr = sh0 & 31, L = sh0 >> 5;
        add_shifted_uniform(acc.a, x0, rr, L);
        add_shifted_uniform(acc.b, x1, rr, L);
        add_shifted_uniform(acc.c, x2, rr, L);
        add_shifted_uniform(acc.d, x3, rr, L);
      } else {
        acc.a += (i128_t)((u128_t)(i128_t)x0 << sh);
        acc.b += (i128_t)((u128_t)(i128_t)x1 << sh);
        acc.c += (i128_t)((u128_t)(i128_t)x2 << sh);
        acc.d += (i128_t)((u128_t)(i128_t)x3 << sh);
      }
    }
    th = th + TH_DWORDS + th[TH_SIZE];
  }
}

template <int WL>
__device__ __forceinline__ void eval_graph_any(const uint32_t *__restrict__ P,
                                               const uint32_t *__restrict__ gh,
                                               const uint32_t (&pv)[WL], const int4 *tab,
                                               const int4 *__restrict__ pitab, Acc4 &acc,
                                               uint32_t &err) {
  switch (gh[GH_W32]) {
    case 1: eval_graph<1, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 2: if constexpr (WL >= 2) eval_graph<2, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 3: if constexpr (WL >= 3) eval_graph<3, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 4: if constexpr (WL >= 4) eval_graph<4, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 5: if constexpr (WL >= 5) eval_graph<5, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 6: if constexpr (WL >= 6) eval_graph<6, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 7: if constexpr (WL >= 7) eval_graph<7, WL>(P, gh, pv, tab, pitab, acc, err); break;
    case 8: if constexpr (WL >= 8) eval_graph<8, WL>(P, gh, pv, tab, pitab, acc, err); break;
    default: break;
  }
}

// set parameter bit `pos` of the lane's words to `bit` (pos is wave-uniform)
template <int WL>
__device__ __forceinline__ void set_param_bit(uint32_t (&pv)[WL], int pos, uint32_t bit) {
  // every word is rewritten with a word-selected mask (no indexed store, so
  // pv stays in registers)
  const int word = pos >> 5;
  const uint32_t m = 1u << (pos & 31);
  const uint32_t b = 0u - (bit & 1u);
#pragma unroll
  for (int w = 0; w < WL; ++w) {
    const uint32_t mw = m & (0u - (uint32_t)(w == word));
    pv[w] = (pv[w] & ~mw) | (mw & b);
  }
}

struct ShotArgs {
  const uint32_t *prog;
  const uint8_t *f;   // (B, num_f)
  const double *U;    // (B, tot_out) or null
  uint8_t *out;       // (B, num_outputs)
  uint32_t *err;
  int64_t B, shot_base;
  uint64_t seed;
  int32_t num_f, num_outputs, tot_out, pad;
};

// The whole autoregressive loop of one shot of component `comp`
// (src/cpp/ksim/kokkos_sim.cpp:214-230, src/python/ksim/sample/evaluate.py:76-85).
template <int WL>
__device__ __forceinline__ void run_shot(const ShotArgs &a, int comp, int64_t s,
                                         const int4 *tab, const int4 *__restrict__ pitab) {
  const uint32_t *__restrict__ P = a.prog;
  const uint32_t *__restrict__ ch = P + P[PH_DWORDS + comp];
  const int ng = (int)ch[CH_NGRAPHS], F = (int)ch[CH_F];
  const int colbase = (int)ch[CH_COLBASE];
  const uint32_t *__restrict__ fsel = P + ch[CH_FSEL_OFF];
  const uint32_t *__restrict__ dest = P + ch[CH_DEST_OFF];
  const uint32_t *__restrict__ ghs = P + ch[CH_GRAPH_OFF];
  const uint8_t *frow = a.f + s * (int64_t)a.num_f;
  uint32_t pv[WL];
  const uint32_t fsel0 = ch[CH_FSEL0];
  if (fsel0 != 0xffffffffu && (((uint32_t)a.num_f | fsel0 | (uint32_t)(uintptr_t)a.f) & 3u) == 0) {
    // contiguous selection on a dword-aligned row: 4 f bytes per load,
    // bit 0 of each byte gathered by one multiply (bits 24..27 of the product)
    const uint32_t *row4 = (const uint32_t *)(frow + fsel0);
#pragma unroll
    for (int w = 0; w < WL; ++w) {
      uint32_t word = 0;
      const int j0 = 32 * w;
      const int jn = min(32, F - j0);
      int jj = 0;
      for (; jj + 4 <= jn; jj += 4) {
        const uint32_t x = row4[(j0 + jj) >> 2] & 0x01010101u;
        word |= ((x * 0x01020408u) >> 24) << jj;
      }
      for (; jj < jn; ++jj) word |= (uint32_t)(frow[fsel0 + j0 + jj] & 1u) << jj;
      pv[w] = word;
    }
  } else {
#pragma unroll
    for (int w = 0; w < WL; ++w) {
      uint32_t word = 0;
      const int j0 = 32 * w;
      const int jn = min(32, F - j0);
      for (int jj = 0; jj < jn; ++jj) word |= (uint32_t)(frow[fsel[j0 + jj]] & 1u) << jj;
      pv[w] = word;
    }
  }
  // per-component stream: seed + 0x9E37*k (src/python/ksim/sample/sampler.py:63)
  uint64_t rs = rng_seed(a.seed + 0x9E37ull * (uint64_t)comp, (uint64_t)(a.shot_base + s));
  uint32_t err = 0;
  Acc4 acc;
  eval_graph_any<WL>(P, ghs, pv, tab, pitab, acc, err);
  double prev = magnitude(acc, (int)ghs[GH_E0]);
  const double *urow = a.U ? a.U + s * (int64_t)a.tot_out + colbase : nullptr;
  uint8_t *orow = a.out + s * (int64_t)a.num_outputs;
  for (int i = 0; i + 1 < ng; ++i) {
    const uint32_t *gh = ghs + (size_t)(i + 1) * GH_DWORDS;
    set_param_bit<WL>(pv, F + i, 1u);  // trying bit
    eval_graph_any<WL>(P, gh, pv, tab, pitab, acc, err);
    const double p1 = magnitude(acc, (int)gh[GH_E0]);
    const double u = urow ? urow[i] : rng_uniform(rs);
    const uint32_t bit = (u < __ddiv_rn(p1, prev)) ? 1u : 0u;
    set_param_bit<WL>(pv, F + i, bit);
    orow[dest[i]] = (uint8_t)bit;
    prev = bit ? p1 : __dsub_rn(prev, p1);
  }
  if (err) atomicOr(a.err, err);
}

__device__ inline void load_table(int4 *s_tab, const int4 *tab) {
  for (int i = threadIdx.x; i < TAB_N; i += blockDim.x) s_tab[i] = tab[i];
  __syncthreads();
}

}  // namespace ksim
