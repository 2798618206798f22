// device.hpp — gfx950 device code of the FlatProgram sampler.
//
// One lane evaluates one shot.  All program data a wave touches is the same
// for its 64 lanes (one component, one graph, one term, one record at a
// time), so records are read through scalar loads into SGPRs and each record
// costs a few VALU ops per lane: AND + popcount per 32-bit mask word, then a
// masked add of a packed exponent delta (see layout.hpp).  The exact term
// value is built from a table (first digit plane read from global memory and
// held in the vector L1; pi powers and float64 magnitudes staged in LDS),
// shifted to a per-graph exponent and added into four integer accumulators:
// int64 when the packer proved the graph's sums fit, 128-bit otherwise, so
// the sum over terms is exact (the reference sums in float64,
// src/cpp/ksim/zx_eval.hpp:168-178).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "layout.hpp"
#include "noise.hpp"

namespace ksim {

typedef __int128 i128_t;
typedef unsigned __int128 u128_t;

// Program words are read through the constant address space: the kernels
// never write them, and loads from this space are invariant, so the compiler
// keeps every header and record load on the scalar path (SMEM into SGPRs)
// even after the kernel's own output stores.
typedef __attribute__((address_space(4))) const uint32_t cword;
__device__ __forceinline__ const cword *as_const(const uint32_t *p) { return (const cword *)p; }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const u32x4 cword4;
typedef __attribute__((address_space(3))) const uint32_t lword;
typedef __attribute__((address_space(3))) u32x4 lword4;

// LDS byte address of a pointer into shared memory
__device__ __forceinline__ uint32_t lds_addr(const void *p) {
  return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void *)p;
}

struct Acc4 {
  i128_t a, b, c, d;
};

// ---------------------------------------------------------------- exact->f64
// |x| < 2^62 for a signed 128-bit x
__device__ __forceinline__ bool fits62(i128_t x) {
  const int64_t hi = (int64_t)(x >> 64);
  const uint64_t lo = (uint64_t)(u128_t)x;
  return (hi == 0 && lo < (1ull << 62)) || (hi == -1 && lo > ~((1ull << 62)));
}

// bit length of |x|
__device__ __forceinline__ int bitlen128(i128_t x) {
  const u128_t m = x < 0 ? (u128_t)0 - (u128_t)x : (u128_t)x;
  const uint64_t hi = (uint64_t)(m >> 64), lo = (uint64_t)m;
  return hi ? 128 - __builtin_clzll(hi) : (lo ? 64 - __builtin_clzll(lo) : 0);
}

// |(a + b w + c i + d conj(w)) * 2^e| in float64, by one fixed sequence:
//   s = b + d, t = b - d (exact);  k = max(0, L - 62) with L the largest bit
//   length of |a|, |c|, |s|, |t|;  a' = a >> k etc. (floor);
//   re = dbl(a') + dbl(s') * SQ,  im = dbl(c') + dbl(t') * SQ,
//   |z| = ldexp(sqrt(re*re + im*im), e + k)
// dbl() is the correctly rounded int64 -> double conversion and every float
// step one IEEE operation (no FMA).  The value does not depend on how the
// exact sum is represented (scaling by 2 commutes with every step), the
// truncation costs < 2^-61 relative, and the oracle runs the same sequence.
// The re/im formula is the cast of zx_eval.hpp:168-169, applied once to the
// exact sum instead of to each term.
__device__ inline double magnitude(const Acc4 &v, int e) {
#pragma clang fp contract(off)
  const double SQ = 0.70710678118654752440;
  const i128_t s = v.b + v.d, t = v.b - v.d;
  int64_t a_, c_, s_, t_;
  int k = 0;
  if (fits62(v.a) && fits62(v.c) && fits62(s) && fits62(t)) {
    a_ = (int64_t)v.a; c_ = (int64_t)v.c; s_ = (int64_t)s; t_ = (int64_t)t;
  } else {
    const int L = max(max(bitlen128(v.a), bitlen128(v.c)), max(bitlen128(s), bitlen128(t)));
    k = L - 62;
    a_ = (int64_t)(v.a >> k); c_ = (int64_t)(v.c >> k);
    s_ = (int64_t)(s >> k); t_ = (int64_t)(t >> k);
  }
  const double re = __dadd_rn((double)a_, __dmul_rn((double)s_, SQ));
  const double im = __dadd_rn((double)c_, __dmul_rn((double)t_, SQ));
  const double n2 = __dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im));
  return __builtin_ldexp(__dsqrt_rn(n2), e + k);
}

// float64 (re, im) of an exact value (a + b w + c i + d conj(w)) * 2^e by the
// first steps of magnitude(): the same 62-bit window, re = dbl(a') +
// dbl(s') * SQ, im = dbl(c') + dbl(t') * SQ, then scaled by 2^(e + k).  For
// coefficients below 2^53 this is the reference's cast of one term
// (zx_eval.hpp:168-171: dbl(b) + dbl(d) is dbl(b + d) there) whatever the
// representation, since scaling by 2 commutes with every step.
__device__ inline void term_float(const Acc4 &v, int e, double &re, double &im) {
#pragma clang fp contract(off)
  const double SQ = 0.70710678118654752440;
  const i128_t s = v.b + v.d, t = v.b - v.d;
  int64_t a_, c_, s_, t_;
  int k = 0;
  if (fits62(v.a) && fits62(v.c) && fits62(s) && fits62(t)) {
    a_ = (int64_t)v.a; c_ = (int64_t)v.c; s_ = (int64_t)s; t_ = (int64_t)t;
  } else {
    const int L = max(max(bitlen128(v.a), bitlen128(v.c)), max(bitlen128(s), bitlen128(t)));
    k = L - 62;
    a_ = (int64_t)(v.a >> k); c_ = (int64_t)(v.c >> k);
    s_ = (int64_t)(s >> k); t_ = (int64_t)(t >> k);
  }
  re = __builtin_ldexp(__dadd_rn((double)a_, __dmul_rn((double)s_, SQ)), e + k);
  im = __builtin_ldexp(__dadd_rn((double)c_, __dmul_rn((double)t_, SQ)), e + k);
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

// x * r exactly, x < 2^62 in magnitude, |r| < 2^31: x = xh 2^32 + xl
__device__ __forceinline__ i128_t mul_64x32(int64_t x, int32_t r) {
  const int64_t hi = (int64_t)(int32_t)(x >> 32) * (int64_t)r;
  const int64_t lo = (int64_t)(uint64_t)(uint32_t)x * (int64_t)r;
  return (i128_t)((u128_t)(i128_t)hi << 32) + (i128_t)lo;
}

// acc + (x * R) * 2^sh for a residual floatfactor R
__device__ __forceinline__ Acc4 add_residual_term(Acc4 acc, int64_t x0, int64_t x1,
                                                            int64_t x2, int64_t x3, int32_t R0,
                                                            int32_t R1, int32_t R2, int32_t R3,
                                                            int sh) {
  const i128_t y0 = mul_64x32(x0, R0) + mul_64x32(x1, R3) - mul_64x32(x2, R2) + mul_64x32(x3, R1);
  const i128_t y1 = mul_64x32(x0, R1) + mul_64x32(x1, R0) + mul_64x32(x2, R3) + mul_64x32(x3, R2);
  const i128_t y2 = mul_64x32(x0, R2) + mul_64x32(x1, R1) + mul_64x32(x2, R0) - mul_64x32(x3, R3);
  const i128_t y3 = mul_64x32(x0, R3) - mul_64x32(x1, R2) - mul_64x32(x2, R1) + mul_64x32(x3, R0);
  acc.a += (i128_t)((u128_t)y0 << sh);
  acc.b += (i128_t)((u128_t)y1 << sh);
  acc.c += (i128_t)((u128_t)y2 << sh);
  acc.d += (i128_t)((u128_t)y3 << sh);
  return acc;
}

// LDS-resident lookup tables of one block
struct LdsTables {
  int4 pi[PI_MAX * PI_MAX];     // pi^a pibar^b
  double mag[MAG_LDS_N];        // magnitude() of the entries with |n| <= MAG_LDS_R
};

// Per-device read-only tables the kernels take.
struct Tables {
  const int4 *tab;     // TAB_DIGITS planes of TAB_N entries
  const int4 *pi;      // PI_MAX x PI_MAX
  const double *mag;   // TAB_N: magnitude() of each (single-plane-summed) entry
};

// XOR over the lane's parameter words of (word & mask word): its popcount
// has the parity of <params, mask>.  One AND, then one v_bitop3 per further
// word computing (mask & word) ^ t (truth table 0x6c for operands
// (mask, t, word)), so a W32-word mask costs W32 VALU ops.
template <int W32, int WL>
__device__ __forceinline__ uint32_t masked_xor(const uint32_t (&pv)[WL],
                                               const cword *__restrict__ m) {
  uint32_t t = pv[0] & m[0];
#pragma unroll
  for (int w = 1; w < W32; ++w) t = __builtin_amdgcn_bitop3_b32(m[w], t, pv[w], 0x6c);
  return t;
}

// The lane's parameter words as the records see them: mxor(m) is the XOR
// over the words of (word & mask word) for a mask of width() words.
// RegWords: the words in registers, the width fixed at compile time (the
// programs of at most 256 parameters).  LdsWords: the words in LDS, word w
// of the lane at pv[64 w] (a wave's words interleaved by lane, so the 64
// lanes of one read hit 64 consecutive dwords), any width at run time.
//
// Parity tables (GF_TABLE graphs, at most PM_TAB_W32_MAX words): a parity
// group's 8 W32 nibble tables are copied into the wave's LDS buffer (2 KiB,
// 1 KiB in the plain kernel without the 128-bit path; one 16-byte
// global-to-LDS load per lane and KiB) and the group's parity word is the
// XOR of one LDS read per nibble, T_j[nibble j of p], instead of W32 + 2
// VALU ops per record.  Each read address is the buffer base | 4 v for
// nibble value v (one v_perm), table j at the read's immediate offset 64 j.
constexpr uint32_t NO_TBUF = 0xffffffffu;  // no table buffer: the groups hold masks
struct TabBase {
  uint32_t base;  // the wave's table buffer (LDS byte address), or NO_TBUF
};
struct NoTab {
  static constexpr uint32_t base = NO_TBUF;
};

// wave-local order between LDS writes and the reads of other lanes' data
// (a wave's LDS operations complete in issue order, the global-to-LDS
// loads once vmcnt says so; the fences keep the compiler from moving
// memory operations across)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Parity word of the group whose tables start at `tabs` (16-byte aligned,
// pm_tab_dwords(W32) dwords) for the lane's W32 parameter words pw, through
// the wave's buffer at LDS byte address `tbase`.
template <int W32, class Word>
__device__ __forceinline__ uint32_t tab_parity_words(const Word &pw, const cword *__restrict__ tabs,
                                                     uint32_t tbase) {
  constexpr int N4 = (int)pm_tab_dwords(W32) / 4;
  const int lane = threadIdx.x & 63;
  // straight from global memory into the buffer (global_load_lds_dwordx4:
  // lane l of a load fills 16 bytes at the base + 16 l), then one wait
  constexpr int NL = (N4 + 63) / 64;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int e = i * 64 + lane;
    if (N4 % 64 == 0 || e < N4)
      __builtin_amdgcn_global_load_lds(
          (const __attribute__((address_space(1))) void *)((uintptr_t)tabs + 16 * (uintptr_t)e),
          (__attribute__((address_space(3))) void *)(uintptr_t)(tbase + 1024u * (uint32_t)i), 16,
          0, 0);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  wave_lds_sync();
  // read addresses: v_perm takes byte k of (word << 2) & 0x3c3c3c3c (the
  // even nibbles times 4) or of (word >> 2) & 0x3c3c3c3c (the odd ones)
  // into byte 0, under bytes 1 and 2 of the base (256-byte aligned, below
  // 2^24).  (The opaque copies keep the compiler from building all 8 W32
  // addresses once per graph, which spilled them.)
  uint32_t base = tbase;
  asm volatile("" : "+s"(base));
  uint32_t y = 0;
#pragma unroll
  for (int w = 0; w < W32; ++w) {
    uint32_t x = pw(w);
    asm volatile("" : "+v"(x));
    const uint32_t pe = (x << 2) & 0x3c3c3c3cu, po = (x >> 2) & 0x3c3c3c3cu;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t sel = 0x0c060500u | (uint32_t)k;
      const int j = 8 * w + 2 * k;
      const uint32_t ae = __builtin_amdgcn_perm(base, pe, sel);
      const uint32_t ao = __builtin_amdgcn_perm(base, po, sel);
      const uint32_t u = ((const lword *)(uintptr_t)ae)[16 * j];
      const uint32_t v = ((const lword *)(uintptr_t)ao)[16 * (j + 1)];
      y = __builtin_amdgcn_bitop3_b32(u, v, y, 0x96);
    }
  }
  wave_lds_sync();  // the next group's copy after these reads
  return y;
}

// Parity words of two groups of at most 2 mask words (tables of 1 KiB or
// less each, the second at +1 KiB of a 2 KiB buffer): one set of read
// addresses serves both (the second group's reads take 1 KiB more in their
// immediate offsets).
template <int W32, class Word>
__device__ __forceinline__ void tab_parity2_words(const Word &pw, const cword *__restrict__ t1,
                                                  const cword *__restrict__ t2, uint32_t tbase,
                                                  uint32_t &y1, uint32_t &y2) {
  static_assert(W32 <= 2, "two groups' tables fill the buffer at 2 mask words");
  constexpr int N4 = (int)pm_tab_dwords(W32) / 4;
  const int lane = threadIdx.x & 63;
  if (N4 % 64 == 0 || lane < N4) {
    __builtin_amdgcn_global_load_lds(
        (const __attribute__((address_space(1))) void *)((uintptr_t)t1 + 16 * (uintptr_t)lane),
        (__attribute__((address_space(3))) void *)(uintptr_t)tbase, 16, 0, 0);
    __builtin_amdgcn_global_load_lds(
        (const __attribute__((address_space(1))) void *)((uintptr_t)t2 + 16 * (uintptr_t)lane),
        (__attribute__((address_space(3))) void *)(uintptr_t)(tbase + 1024u), 16, 0, 0);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  wave_lds_sync();
  uint32_t base = tbase;
  asm volatile("" : "+s"(base));
  uint32_t a = 0, b = 0;
#pragma unroll
  for (int w = 0; w < W32; ++w) {
    uint32_t x = pw(w);
    asm volatile("" : "+v"(x));
    const uint32_t pe = (x << 2) & 0x3c3c3c3cu, po = (x >> 2) & 0x3c3c3c3cu;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t sel = 0x0c060500u | (uint32_t)k;
      const int j = 8 * w + 2 * k;
      const lword *ae = (const lword *)(uintptr_t)__builtin_amdgcn_perm(base, pe, sel);
      const lword *ao = (const lword *)(uintptr_t)__builtin_amdgcn_perm(base, po, sel);
      a = __builtin_amdgcn_bitop3_b32(ae[16 * j], ao[16 * (j + 1)], a, 0x96);
      b = __builtin_amdgcn_bitop3_b32(ae[16 * j + 256], ao[16 * (j + 1) + 256], b, 0x96);
    }
  }
  wave_lds_sync();  // the next groups' copies after these reads
  y1 = a;
  y2 = b;
}

template <int W32, int WL>
struct RegWords {
  static_assert(W32 <= WL, "graph wider than the lane's parameter words");
  static constexpr bool kTab = W32 <= PM_TAB_W32_MAX;
  static constexpr bool kTabNarrow = W32 <= PM_TAB_W32_MAX_NARROW;
  static constexpr bool kTab2 = W32 <= 2;  // two groups per 2 KiB buffer
  using Tab = typename std::conditional<kTab, TabBase, NoTab>::type;
  const uint32_t (&pv)[WL];
  __device__ __forceinline__ int width() const { return W32; }
  __device__ __forceinline__ uint32_t mxor(const cword *__restrict__ m) const {
    return masked_xor<W32, WL>(pv, m);
  }
  __device__ __forceinline__ uint32_t operator()(int w) const { return pv[w]; }
  __device__ __forceinline__ void tab_init(Tab &t, uint32_t base) const {
    if constexpr (kTab) t.base = base;
  }
  __device__ __forceinline__ uint32_t tab_parity(const cword *__restrict__ tabs,
                                                 const Tab &t) const {
    if constexpr (kTab)
      return tab_parity_words<W32>(*this, tabs, t.base);
    else
      return 0;
  }
  __device__ __forceinline__ void tab_parity2(const cword *__restrict__ t1,
                                              const cword *__restrict__ t2, const Tab &t,
                                              uint32_t &y1, uint32_t &y2) const {
    if constexpr (kTab2) tab_parity2_words<W32>(*this, t1, t2, t.base, y1, y2);
  }
};
// The LDS-word kernels read tables too (a batch puts every program in the
// kernel its widest member needs): their exchange area holds the buffer,
// and a graph with tables has at most PM_TAB_W32_MAX words (a switch picks
// the unrolled form).
struct LdsWords {
  static constexpr bool kTab = true;
  static constexpr bool kTabNarrow = true;
  static constexpr bool kTab2 = false;
  using Tab = TabBase;
  const uint32_t *pv;
  int W32;
  __device__ __forceinline__ int width() const { return W32; }
  __device__ __forceinline__ uint32_t mxor(const cword *__restrict__ m) const {
    uint32_t t = pv[0] & m[0];
    for (int w = 1; w < W32; ++w) t = __builtin_amdgcn_bitop3_b32(m[w], t, pv[64 * w], 0x6c);
    return t;
  }
  __device__ __forceinline__ uint32_t operator()(int w) const { return pv[64 * w]; }
  __device__ __forceinline__ void tab_init(Tab &t, uint32_t base) const { t.base = base; }
  __device__ __forceinline__ uint32_t tab_parity(const cword *__restrict__ tabs,
                                                 const Tab &t) const {
    switch (W32) {
      case 1: return tab_parity_words<1>(*this, tabs, t.base);
      case 2: return tab_parity_words<2>(*this, tabs, t.base);
      case 3: return tab_parity_words<3>(*this, tabs, t.base);
      case 4: return tab_parity_words<4>(*this, tabs, t.base);
      default: return 0;  // the validator admits no wider graph with tables
    }
  }
  __device__ __forceinline__ void tab_parity2(const cword *__restrict__, const cword *__restrict__,
                                              const Tab &, uint32_t &, uint32_t &) const {}
};

// Graphs with approximate floatfactors (GF_APPROX) are evaluated only by
// the kernels that keep parameter words in LDS (the host sends every
// program holding one there), so the register kernels carry no float path.
template <class PW>
constexpr bool kApprox = false;
template <>
constexpr bool kApprox<LdsWords> = true;

// ------------------------------------------------------ record read-ahead
// The records of a GF_STREAM graph are pulled into L2 ahead of the scalar
// loads that walk them: before each exponent chunk the wave touches the
// 128-byte lines from the furthest line it touched so far (or the chunk,
// if that is further) up to STREAM_AHEAD dwords past the chunk's end, one
// 4-byte vector load per line, at most 64 lines per touch.  The scalar
// loads then find the lines in L2 instead of waiting on HBM one round trip
// at a time.  The program ends STREAM_PAD >= STREAM_AHEAD dwords past its
// last record, so every touched line is inside the program words.
struct Stream {
  const cword *pf;  // touched up to here; nullptr: the graph is not streamed
};

// Touch the lines of [max(ps.pf, from), want); the returned value is
// discarded by stream_done once the chunk is walked, so the load is waited
// for only then.
__device__ __forceinline__ uint32_t stream_touch(Stream &ps, const cword *from,
                                                 const cword *want) {
  if (ps.pf == nullptr) return 0;
  const cword *lo = ps.pf > from ? ps.pf : from;
  const cword *base = (const cword *)((uintptr_t)lo & ~(uintptr_t)127);
  if (base >= want) return 0;
  const cword *a = base + 32 * (threadIdx.x & 63);
  uint32_t v = 0;
  if (a < want) v = *a;
  const cword *end = base + 64 * 32;
  ps.pf = end < want ? end : want;
  return v;
}

__device__ __forceinline__ void stream_done(uint32_t v) { asm volatile("" ::"v"(v)); }

// -------------------------------------------------------------- one term
// Exponents of one term for this lane's parameter words pw: the term is
// w^S (1+w)^V (1+sqrt2)^N pi^A pibar^B [R] 2^E, or zero when bit 0 of
// `dead` is set.  th points at the term header.  LONG kernels read the
// records of streamed graphs ahead (ps).
template <class PW, bool LONG = false>
__device__ __forceinline__ void term_exps(const cword *__restrict__ th, const PW &pw, int &S,
                                          int &V, int &N, int &A, int &Bq, uint32_t &dead,
                                          Stream &ps, const typename PW::Tab &tab) {
  const int W32 = pw.width();
  const uint32_t nproj = th[TH_NPROJ], nppx = th[TH_NPPX], nchunk = th[TH_NCHUNK];
  S = (int)th[TH_S]; V = (int)th[TH_V]; N = (int)th[TH_N];
  A = (int)th[TH_A]; Bq = (int)th[TH_B];
  const cword *__restrict__ rp = th + TH_DWORDS;
  // PROJ: the term vanishes unless every projector parity matches
  dead = 0;
  for (uint32_t k = 0; k < nproj; ++k, rp += W32 + 1) {
    dead |= rp[W32] + __popc(pw.mxor(rp));
  }
  // PPX: (-1)^{psi phi}.  LONG (the kernels with 128-bit sums, which serve
  // the programs with long terms): unrolled by 4, which amortises the loop's
  // scalar work over many records (measured 4% faster on the 64-program
  // batch, 4% slower on programs of 3-record loops)
  uint32_t ex = 0;
  if constexpr (LONG) {
#pragma unroll 4
    for (uint32_t k = 0; k < nppx; ++k, rp += 2 * W32 + 2) {
      const uint32_t cs = rp[2 * W32] + __popc(pw.mxor(rp));
      const uint32_t cf = rp[2 * W32 + 1] + __popc(pw.mxor(rp + W32));
      ex ^= cs & cf;
    }
  } else {
    for (uint32_t k = 0; k < nppx; ++k, rp += 2 * W32 + 2) {
      const uint32_t cs = rp[2 * W32] + __popc(pw.mxor(rp));
      const uint32_t cf = rp[2 * W32 + 1] + __popc(pw.mxor(rp + W32));
      ex ^= cs & cf;
    }
  }
  // parity-table graphs carry their PROJ and PPX parities in the parity
  // groups of the first chunk: jd collects the failed projectors there
  uint32_t jd = 0;
  // LIN / PAIR chunks: packed exponent adds
  for (uint32_t ch = 0; ch < nchunk; ++ch) {
    const uint32_t nl = rp[0], npr = rp[1], nw = rp[3];
    const uint32_t npm = nw & PM_COUNT_MASK, pad = nw >> PM_PAD_SHIFT;
    uint32_t x = rp[2];
    uint32_t touched = 0;
    if constexpr (LONG) {
      const uint32_t pmw = tab.base != NO_TBUF
                               ? npm * (uint32_t)pm_tab_group_dwords(W32)
                               : (npm + PM_GROUP - 1) / PM_GROUP * PM_HDR + npm * W32;
      touched = stream_touch(ps, rp, rp + CHUNK_HDR + pad + pmw + nl * (W32 + 1) +
                                         npr * (2 * W32 + 3) + STREAM_AHEAD);
    }
    rp += CHUNK_HDR + pad;
    if (tab.base != NO_TBUF) {
      // npm parity groups, each a header [S+, S-, N+, N-, X, Xc, J, Jc] and
      // its nibble tables: the parity word y holds one bit per form; sign
      // records step S and N as below, PPX pairs (psi at a bit b marked in
      // X, phi at b + 1) add the parity of popcount(z & (z >> 1) & X) with
      // z = y ^ Xc to the PPX sign, and a PROJ at a bit of J kills the term
      // unless its parity is Jc's bit
      // (each header is read before the group's tables, so its scalar
      // loads overlap the table copy)
      struct Hdr {
        uint32_t sp, sm, tp, tm, xm, xc, jm, jc;
      };
      auto header = [](const cword *__restrict__ h) {
        return Hdr{h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7]};
      };
      auto apply = [&](const Hdr &h, uint32_t y) {
        const uint32_t cs = __popc((y & h.sp) | (~y & h.sm));
        const uint32_t cn = __popc((y & h.tp) | (~y & h.tm));
        uint32_t sn;
        asm("v_lshl_add_u32 %0, %1, 21, %2" : "=v"(sn) : "v"(cs), "v"(cn));
        asm("v_lshl_add_u32 %0, %1, 9, %0" : "+v"(x) : "v"(sn));
        if (h.xm | h.jm) {  // wave-uniform
          const uint32_t z = y ^ h.xc;
          ex += __popc(z & (z >> 1) & h.xm);
          jd |= (y ^ h.jc) & h.jm;
        }
      };
      const uint32_t gd = pm_tab_group_dwords(W32);
      uint32_t g = 0;
      if constexpr (LONG && PW::kTab2) {
        // the 128-bit kernels' 2 KiB buffers take two groups of up to 2
        // mask words at a time
        for (; g + 1 < npm; g += 2, rp += 2 * gd) {
          const Hdr h1 = header(rp), h2 = header(rp + gd);
          uint32_t y1, y2;
          pw.tab_parity2(rp + PM_TAB_HDR, rp + gd + PM_TAB_HDR, tab, y1, y2);
          apply(h1, y1);
          apply(h2, y2);
        }
      }
      for (; g < npm; ++g, rp += gd) {
        const Hdr h = header(rp);
        apply(h, pw.tab_parity(rp + PM_TAB_HDR, tab));
      }
    } else {
      // sign records (S += 2s, N += t, s, t in {-1, 0, 1}): the parities of
      // a group go into one word, record j of n at bit 32 - n + j (each
      // funnel shift takes bit 0 of the popcount in at the top), then two
      // popcounts over the group's sign masks give the S and N steps; the
      // chunk base holds the -2^30 |S-| - 2^9 |N-| of the parity-0 records
      // counted there
      for (uint32_t g0 = 0; g0 < npm; g0 += PM_GROUP) {
        const uint32_t rem = npm - g0;  // wave-uniform: kept on the scalar unit
        const uint32_t n = rem < (uint32_t)PM_GROUP ? rem : (uint32_t)PM_GROUP;
        const uint32_t sp = rp[0], sm = rp[1], tp = rp[2], tm = rp[3];
        rp += PM_HDR;
        uint32_t pk = 0;
        if constexpr (LONG) {
#pragma unroll 8
          for (uint32_t k = 0; k < n; ++k, rp += W32)
            pk = __builtin_amdgcn_alignbit(__popc(pw.mxor(rp)), pk, 1);
        } else {
#pragma unroll 4
          for (uint32_t k = 0; k < n; ++k, rp += W32)
            pk = __builtin_amdgcn_alignbit(__popc(pw.mxor(rp)), pk, 1);
        }
        const uint32_t cs = __popc((pk & sp) | (~pk & sm));
        const uint32_t cn = __popc((pk & tp) | (~pk & tm));
        // x += 2^30 cs + 2^9 cn as two shift-adds (the compiler's own form
        // is two shifts and an add)
        uint32_t sn;
        asm("v_lshl_add_u32 %0, %1, 21, %2" : "=v"(sn) : "v"(cs), "v"(cn));
        asm("v_lshl_add_u32 %0, %1, 9, %0" : "+v"(x) : "v"(sn));
      }
    }
    // LIN records, unrolled so each scalar load fetches several records
    if constexpr (LONG) {
#pragma unroll 8
      for (uint32_t k = 0; k < nl; ++k, rp += W32 + 1) {
        const uint32_t c = __popc(pw.mxor(rp));
        x += (0u - (c & 1u)) & rp[W32];
      }
    } else {
#pragma unroll 4
      for (uint32_t k = 0; k < nl; ++k, rp += W32 + 1) {
        const uint32_t c = __popc(pw.mxor(rp));
        x += (0u - (c & 1u)) & rp[W32];
      }
    }
    if constexpr (LONG) {
#pragma unroll 2
      for (uint32_t k = 0; k < npr; ++k, rp += 2 * W32 + 3) {
        const uint32_t ca = __popc(pw.mxor(rp));
        const uint32_t cb = __popc(pw.mxor(rp + W32));
        const uint32_t ma = 0u - (ca & 1u), mb = 0u - (cb & 1u);
        x += (ma & rp[2 * W32]) + (mb & rp[2 * W32 + 1]) + (ma & mb & rp[2 * W32 + 2]);
      }
    } else {
      for (uint32_t k = 0; k < npr; ++k, rp += 2 * W32 + 3) {
        const uint32_t ca = __popc(pw.mxor(rp));
        const uint32_t cb = __popc(pw.mxor(rp + W32));
        const uint32_t ma = 0u - (ca & 1u), mb = 0u - (cb & 1u);
        x += (ma & rp[2 * W32]) + (mb & rp[2 * W32 + 1]) + (ma & mb & rp[2 * W32 + 2]);
      }
    }
    if constexpr (LONG) stream_done(touched);
    V += (int)((x >> PK_OV) & ((1u << PK_WV) - 1));
    N += (int)((x >> PK_ON) & ((1u << PK_WN) - 1));
    A += (int)((x >> PK_OA) & ((1u << PK_WA) - 1));
    Bq += (int)((x >> PK_OB) & ((1u << PK_WB) - 1));
    S += (int)(x >> PK_OS);
  }
  S += (int)((ex & 1u) << 2);
  if constexpr (PW::kTab) dead |= jd ? 1u : 0u;
}

// materialise w^S (1+w)^V (1+sqrt2)^N pi^A pibar^B R:
// (1+w)^4 = 2 w^2 (1+sqrt2)^2, so with q = V>>2, r = V&3 the value is
// 2^q w^{S+2q} (1+w)^r (1+sqrt2)^{N+2q} pi^A pibar^B R.  Table row n':
// the LDS copy for the central rows, else the global table, whose rows
// beyond TAB_R spread over TAB_DIGITS planes of base-2^31 digits (each
// < 2^30); digit d adds at shift sh + 31 d.
struct TermIdx {
  int q, r, n2, s2, gi;
  bool in_tab;
};
__device__ __forceinline__ TermIdx term_index(int S, int V, int N) {
  TermIdx t;
  t.q = V >> 2;
  t.r = V & 3;
  t.n2 = N + 2 * t.q;
  t.s2 = (S + 2 * t.q) & 7;
  t.gi = ((t.n2 + TAB_W) * 4 + t.r) * 8 + t.s2;
  t.in_tab = (unsigned)(t.n2 + TAB_W) <= 2u * TAB_W;
  return t;
}

// the first digit plane is read from global memory: its central rows stay
// resident in the CU's vector L1 cache (an LDS copy measured no faster).
// A vanished term, or a row beyond the table (an error), reads the zero
// entry, so every later step may add the lane's value unconditionally.
template <bool ZeroRow>
__device__ __forceinline__ int4 table_digit0(const TermIdx &ti, bool vanished, const LdsTables *lt,
                                             const Tables &tb, uint32_t &err) {
  if constexpr (ZeroRow) {
    if (!ti.in_tab && !vanished) err |= ERR_TAB_RANGE;
    return tb.tab[ti.in_tab && !vanished ? ti.gi : TAB_ZERO];
  } else {
    // the 64-bit path masks vanished lanes itself (add_term_narrow)
    if (ti.in_tab) return tb.tab[ti.gi];
    if (!vanished) err |= ERR_TAB_RANGE;
    return make_int4(0, 0, 0, 0);
  }
}

// e * pi^A pibar^B for a table digit e (|e| < 2^30; pi powers < 2^24), exact
// in int64.
// Always (the 128-bit sums, which serve the programs with many terms and
// phase pairs): a term that may carry pi powers (TF_PI, the same for the
// whole wave) takes the product, pi^0 pibar^0 = 1 included, since most waves
// hold some lane with a power and a branch per lane costs more than the
// products it skips (config-4 batch 4 % faster).  Otherwise the lanes whose
// exponents are both zero skip it (the 64-bit sums of short Clifford-like
// graphs, where powers are rare: config 2 7 % faster that way).
// Exponents of PI_MAX or more break the packer's bound: they set the error
// bit (the host then fails the call) and read a wrapped table row.
template <bool Always>
__device__ __forceinline__ void times_pi(int4 e, uint32_t flags, int A, int Bq,
                                         const LdsTables *lt, int64_t &x0, int64_t &x1,
                                         int64_t &x2, int64_t &x3, uint32_t &err) {
  static_assert((PI_MAX & (PI_MAX - 1)) == 0, "pi table side must be a power of two");
  if ((flags & TF_PI) && (Always || (A | Bq))) {
    int pi_row;
    if constexpr (Always) {
      if ((uint32_t)(A | Bq) >= (uint32_t)PI_MAX) err |= ERR_PI_RANGE;
      pi_row = (A & (PI_MAX - 1)) * PI_MAX + (Bq & (PI_MAX - 1));
    } else {
      if (A >= PI_MAX || Bq >= PI_MAX) {
        err |= ERR_PI_RANGE;
        e = make_int4(0, 0, 0, 0);
      }
      pi_row = min(A, PI_MAX - 1) * PI_MAX + min(Bq, PI_MAX - 1);
    }
    // pi^A pibar^B = P0 + P1 i sqrt2 = (P0, P1, 0, -P1) (pi = 1 + w - conj w,
    // checked when the table is built), so e * pi^A pibar^B = P0 e + P1 (w -
    // conj w) e: (w - conj w) e is a signed sum of e's coefficients (each below
    // 2^30, so the sums fit 32 bits) and two 32 x 32 -> 64-bit products per
    // coefficient remain
    const int4 p = lt->pi[pi_row];
    const int32_t f0 = e.w - e.y, f1 = e.x - e.z, f2 = e.y + e.w, f3 = -(e.x + e.z);
    x0 = (int64_t)e.x * p.x + (int64_t)f0 * p.y;
    x1 = (int64_t)e.y * p.x + (int64_t)f1 * p.y;
    x2 = (int64_t)e.z * p.x + (int64_t)f2 * p.y;
    x3 = (int64_t)e.w * p.x + (int64_t)f3 * p.y;
  } else {
    x0 = e.x; x1 = e.y; x2 = e.z; x3 = e.w;
  }
}

// acc + e * pi^{A,B} * R * 2^sh for one table digit e (|e| < 2^30).
// zero: the lane adds nothing (a vanished term, or no digit at this place).
// Uniform: try the wave-uniform shift path first.
// The low 64 bits of a 128-bit sum as a 64-bit sum: a GF_NARROW graph's exact
// sum (below 2^62 by the packer's bound) kept in the registers of the
// 128-bit sums, so the kernels with both paths hold one set of sums.
__device__ __forceinline__ void add_low64(i128_t &acc, int64_t v) {
  const u128_t u = (u128_t)acc;
  acc = (i128_t)((u & ~(u128_t)~0ull) | (u128_t)((uint64_t)u + (uint64_t)v));
}
__device__ __forceinline__ int64_t low64(i128_t acc) { return (int64_t)(uint64_t)(u128_t)acc; }

template <bool Uniform>
__device__ __forceinline__ Acc4 add_term_digit(Acc4 acc, int4 e, bool zero, uint32_t flags, int A,
                                               int Bq, const LdsTables *lt,
                                               const cword *__restrict__ th, int sh,
                                               uint32_t &err) {
  // e is zero on the lanes that add nothing (table_digit0, or the caller's
  // select of a further digit)
  const uint64_t live = __builtin_amdgcn_ballot_w64(!zero);
  if (live == 0) return acc;
  int64_t x0, x1, x2, x3;
  times_pi<true>(e, flags, A, Bq, lt, x0, x1, x2, x3, err);
  // add X * 2^sh into the four 128-bit sums.  The shift is almost always the
  // same on every lane that adds something (it varies only through
  // phase-pair records), so the common path adds 32-bit limbs at a
  // wave-uniform offset, its range checked once for the wave.
  const int sh0 = __builtin_amdgcn_readlane(sh, __builtin_ctzll(live));
  if (Uniform && !(flags & TF_RESIDUAL) && (uint32_t)sh0 <= (uint32_t)SHIFT_MAX &&
      __builtin_amdgcn_ballot_w64(!zero && sh != sh0) == 0) {
    const int rr = sh0 & 31, L = sh0 >> 5;
    add_shifted_uniform(acc.a, x0, rr, L);
    add_shifted_uniform(acc.b, x1, rr, L);
    add_shifted_uniform(acc.c, x2, rr, L);
    add_shifted_uniform(acc.d, x3, rr, L);
    return acc;
  }
  if (sh < 0 || sh > SHIFT_MAX) {
    if (x0 | x1 | x2 | x3) err |= ERR_SHIFT;
    x0 = x1 = x2 = x3 = 0;
    sh = 0;
  }
  if (flags & TF_RESIDUAL) {
    // a floatfactor that does not factor: one Z[w] multiply in 128 bits
    acc = add_residual_term(acc, x0, x1, x2, x3, (int32_t)th[TH_R0], (int32_t)th[TH_R1],
                            (int32_t)th[TH_R2], (int32_t)th[TH_R3], sh);
  } else {
    {
      acc.a += (i128_t)((u128_t)(i128_t)x0 << sh);
      acc.b += (i128_t)((u128_t)(i128_t)x1 << sh);
      acc.c += (i128_t)((u128_t)(i128_t)x2 << sh);
      acc.d += (i128_t)((u128_t)(i128_t)x3 << sh);
    }
  }
  return acc;
}

// Four signed 64-bit sums: the exact term sum of a GF_NARROW graph, whose
// packer bound keeps every partial sum, shifted term and product below
// 2^62 in magnitude.
struct Acc64 {
  int64_t a, b, c, d;
};

// n: the 64-bit sums, or the low halves of the 128-bit sums (Acc4)
template <class AccT>
__device__ __forceinline__ void add_term_narrow(AccT &n, int4 e, bool zero, uint32_t flags,
                                                int A, int Bq, const LdsTables *lt,
                                                const cword *__restrict__ th, int sh,
                                                uint32_t &err) {
  const int32_t keep = (int32_t)zero - 1;
  e.x &= keep; e.y &= keep; e.z &= keep; e.w &= keep;
  int64_t x0, x1, x2, x3;
  times_pi<false>(e, flags, A, Bq, lt, x0, x1, x2, x3, err);
  if (flags & TF_RESIDUAL)
    zmul64(x0, x1, x2, x3, (int32_t)th[TH_R0], (int32_t)th[TH_R1], (int32_t)th[TH_R2],
           (int32_t)th[TH_R3]);
  if (sh < 0 || sh > 62) {
    if (x0 | x1 | x2 | x3) err |= ERR_SHIFT;
    x0 = x1 = x2 = x3 = 0;
    sh = 0;
  }
  if constexpr (sizeof(n.a) == 8) {
    n.a += (int64_t)((uint64_t)x0 << sh);
    n.b += (int64_t)((uint64_t)x1 << sh);
    n.c += (int64_t)((uint64_t)x2 << sh);
    n.d += (int64_t)((uint64_t)x3 << sh);
  } else {
    add_low64(n.a, (int64_t)((uint64_t)x0 << sh));
    add_low64(n.b, (int64_t)((uint64_t)x1 << sh));
    add_low64(n.c, (int64_t)((uint64_t)x2 << sh));
    add_low64(n.d, (int64_t)((uint64_t)x3 << sh));
  }
}


// Running state of a graph's term sum.
struct GraphSum {
  Acc4 acc;     // 128-bit sums (graphs without GF_NARROW)
  Acc64 n;      // 64-bit sums (GF_NARROW)
  double pm;    // the float64 magnitude (GF_MONO)
  double fr, fi;  // the float64 sum (GF_APPROX)
};

// clear the sums the graph's flags select (the others stay unset: the
// evaluation never reads them).  Kernels with the 128-bit path sum a
// GF_NARROW graph in the low halves of the 128-bit sums (add_low64), so
// only one set of sums is live across the term loop.
template <bool WIDE>
__device__ __forceinline__ void graph_sum_clear(GraphSum &g, uint32_t gflags) {
  if (gflags & GF_MONO) {
    g.pm = 0.0;
  } else if constexpr (WIDE) {
    g.acc.a = g.acc.b = g.acc.c = g.acc.d = 0;
  } else {
    g.n.a = g.n.b = g.n.c = g.n.d = 0;
  }
}

// Add terms [0, nt) starting at program dword `th` into g, by the graph's
// flags: a single table magnitude (GF_MONO), 64-bit sums (GF_NARROW) or
// 128-bit sums.
// GF_TABLE graphs read their parity tables through the wave's LDS buffer at
// `tbuf` (an LDS byte address, 256-byte aligned); the caller has checked
// that the kernel reads tables (tab_ok).
template <class PW, bool WIDE = true>
__device__ __forceinline__ void sum_terms(const cword *__restrict__ th, uint32_t nt,
                                          uint32_t gflags, int e0, const PW &pw,
                                          const LdsTables *lt, const Tables &tb, GraphSum &g,
                                          uint32_t &err, const cword *__restrict__ apx = nullptr,
                                          uint32_t tbuf = NO_TBUF) {
  Stream ps{(gflags & GF_STREAM) ? th : nullptr};
  typename PW::Tab tab;
  if constexpr (PW::kTab) {
    tab.base = NO_TBUF;
    if constexpr (WIDE || PW::kTabNarrow)
      if (gflags & GF_TABLE) pw.tab_init(tab, tbuf);
  }
  for (uint32_t t = 0; t < nt; ++t) {
    const uint32_t flags = th[TH_FLAGS];
    const int E = (int)th[TH_E];
    int S, V, N, A, Bq;
    uint32_t dead;
    term_exps<PW, WIDE>(th, pw, S, V, N, A, Bq, dead, ps, tab);
    const TermIdx ti = term_index(S, V, N);
    const bool vanished = dead & 1u;
    const int sh = ti.q + E;
    if (gflags & GF_MONO) {
      // |value| = magnitude(table entry) * 2^(sh + e0) exactly: scaling by
      // a power of two commutes with every step of magnitude(), and the
      // entry is below 2^31, so the 62-bit window never truncates it
      double m;
      if ((unsigned)(ti.n2 + MAG_LDS_R) <= 2u * MAG_LDS_R) {
        m = lt->mag[((ti.n2 + MAG_LDS_R) * 4 + ti.r) * 8 + ti.s2];
      } else if (ti.in_tab) {
        m = tb.mag[ti.gi];
      } else {
        if (!vanished) err |= ERR_TAB_RANGE;
        m = 0.0;
      }
      g.pm = vanished ? 0.0 : __builtin_ldexp(m, sh + e0);
    } else {
      const int4 e = table_digit0<WIDE>(ti, vanished, lt, tb, err);
      if (!WIDE || (gflags & GF_NARROW)) {
        if (!WIDE && !(gflags & GF_NARROW)) err |= ERR_WIDE;
        if constexpr (WIDE)
          add_term_narrow(g.acc, e, vanished, flags, A, Bq, lt, th, sh, err);
        else
          add_term_narrow(g.n, e, vanished, flags, A, Bq, lt, th, sh, err);
      } else if (kApprox<PW> && WIDE && (gflags & GF_APPROX)) {
        // the term's exact value at its own exponent (digit planes at their
        // own shifts), cast to float64, times its complex64 factor, added
        // to the float64 sum in term order (zx_eval.hpp:168-178)
        Acc4 v;
        v.a = v.b = v.c = v.d = 0;
        v = add_term_digit<true>(v, e, vanished, flags, A, Bq, lt, th, 0, err);
        const bool wide = ti.in_tab && !vanished && (unsigned)(ti.n2 + TAB_R) > 2u * TAB_R;
        if (__builtin_amdgcn_ballot_w64(wide)) {
#pragma nounroll
          for (int d = 1; d < TAB_DIGITS; ++d) {
            const int4 ed = wide ? tb.tab[d * TAB_N + ti.gi] : make_int4(0, 0, 0, 0);
            v = add_term_digit<false>(v, ed, !wide, flags, A, Bq, lt, th, TAB_DIGIT_BITS * d, err);
          }
        }
        if (!vanished) {
          double re, im;
          term_float(v, e0 + sh, re, im);
          const double ar = (double)__uint_as_float(apx[2 * t]);
          const double ai = (double)__uint_as_float(apx[2 * t + 1]);
          g.fr = __dadd_rn(g.fr, __dsub_rn(__dmul_rn(re, ar), __dmul_rn(im, ai)));
          g.fi = __dadd_rn(g.fi, __dadd_rn(__dmul_rn(re, ai), __dmul_rn(im, ar)));
        }
      } else if constexpr (WIDE) {
        g.acc = add_term_digit<true>(g.acc, e, vanished, flags, A, Bq, lt, th, sh, err);
        const bool wide = ti.in_tab && !vanished && (unsigned)(ti.n2 + TAB_R) > 2u * TAB_R;
        if (__builtin_amdgcn_ballot_w64(wide)) {
          // a rare tail: some lane's |n'| exceeds the single-digit rows
#pragma nounroll
          for (int d = 1; d < TAB_DIGITS; ++d) {
            const int4 ed = wide ? tb.tab[d * TAB_N + ti.gi] : make_int4(0, 0, 0, 0);
            g.acc = add_term_digit<false>(g.acc, ed, !wide, flags, A, Bq, lt, th,
                                          sh + TAB_DIGIT_BITS * d, err);
          }
        }
      }
    }
    th = th + TH_DWORDS + th[TH_SIZE];
  }
}

__device__ __forceinline__ Acc4 widen(const Acc64 &n) {
  Acc4 a;
  a.a = n.a; a.b = n.b; a.c = n.c; a.d = n.d;
  return a;
}

// Waves that share the term sums of one 64-shot tile: L waves (1, 2 or 4)
// of a block form a group; wave `part` of the group sums term parts
// [part * GH_PARTS / L, (part + 1) * GH_PARTS / L) of each graph and the
// partial sums meet in `xch`, the block's exchange area in LDS:
// [4 waves][4 int4][64 lanes] (16 KiB).  Every wave of the group then
// holds the same exact sum, so all of them take the same decisions.
struct Split {
  int L;          // waves per group
  int part;       // this wave's index in its group
  int first;      // the group's first wave in the block
  int4 *xch;      // the block's exchange area (nullptr when L == 1)
  uint32_t tbuf = NO_TBUF;  // kernels without an exchange area: the wave's
                            // parity-table buffer (LDS byte address, 256-byte
                            // aligned), or NO_TBUF
};
constexpr int XCH_INT4 = 4 * 4 * 64;

// The wave's 2 KiB parity-table buffer for the graph about to be summed
// (NO_TBUF: the kernel has none).  With an exchange area it is the half of the
// wave's area that the next exchange writes (xs bit 0): the other half may
// still be read by the group's other waves after the last 64-bit exchange,
// and this one was last read before that exchange's barrier.
__device__ __forceinline__ uint32_t table_buffer(const Split &sp, uint32_t xs) {
  if (sp.xch == nullptr) return sp.tbuf;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  return lds_addr(sp.xch + wave * 256 + (int)(xs & 1u) * 128);
}

// Sum the partial sums of the group's waves (exact integer sums, so every
// wave ends with the same value).  Uniform across the block: every wave of
// the block calls it for the same graph, so every barrier is reached by
// all.  64-bit sums use half of a wave's area at a time, alternating
// (xs bit 0), with one barrier: a half is written again only two
// exchanges later, after a barrier that every wave reached when done
// reading it.  128-bit sums use the whole area between two barriers (one
// more first if the previous exchange left a half being read, xs bit 1).
__device__ __forceinline__ void exchange_sums(GraphSum &g, bool narrow, const Split &sp,
                                              uint32_t &xs) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  int4 *buf = sp.xch;
  if (narrow) {
    const int half = (int)(xs & 1u) * 128;
    xs = (xs ^ 1u) | 2u;
    int4 *mine = buf + wave * 256 + half + lane;
    mine[0] = make_int4((int)(uint32_t)g.n.a, (int)(g.n.a >> 32), (int)(uint32_t)g.n.b, (int)(g.n.b >> 32));
    mine[64] = make_int4((int)(uint32_t)g.n.c, (int)(g.n.c >> 32), (int)(uint32_t)g.n.d, (int)(g.n.d >> 32));
    __syncthreads();
    for (int j = 0; j < sp.L; ++j) {
      const int w = sp.first + j;
      if (w == wave) continue;
      const int4 *o = buf + w * 256 + half + lane;
      const int4 x = o[0], y = o[64];
      g.n.a += (int64_t)(((uint64_t)(uint32_t)x.y << 32) | (uint32_t)x.x);
      g.n.b += (int64_t)(((uint64_t)(uint32_t)x.w << 32) | (uint32_t)x.z);
      g.n.c += (int64_t)(((uint64_t)(uint32_t)y.y << 32) | (uint32_t)y.x);
      g.n.d += (int64_t)(((uint64_t)(uint32_t)y.w << 32) | (uint32_t)y.z);
    }
    return;
  }
  if (xs & 2u) __syncthreads();  // the last 64-bit exchange's reads are over
  xs &= 1u;
  int4 *mine = buf + wave * 256 + lane;
  const i128_t v[4] = {g.acc.a, g.acc.b, g.acc.c, g.acc.d};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const u128_t u = (u128_t)v[k];
    mine[64 * k] = make_int4((int)(uint32_t)u, (int)(uint32_t)(u >> 32), (int)(uint32_t)(u >> 64),
                             (int)(uint32_t)(u >> 96));
  }
  __syncthreads();
  for (int j = 0; j < sp.L; ++j) {
    const int w = sp.first + j;
    if (w == wave) continue;
    const int4 *o = buf + w * 256 + lane;
    i128_t r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int4 x = o[64 * k];
      r[k] = (i128_t)(((u128_t)(uint32_t)x.w << 96) | ((u128_t)(uint32_t)x.z << 64) |
                      ((u128_t)(uint32_t)x.y << 32) | (u128_t)(uint32_t)x.x);
    }
    g.acc.a += r[0]; g.acc.b += r[1]; g.acc.c += r[2]; g.acc.d += r[3];
  }
  __syncthreads();
}

// |amplitude| of one graph (float64) for this lane's parameter words: the
// value magnitude() gives on the exact term sum.  SPLIT: the graph's terms
// are shared by the waves of sp's group (single-term graphs are evaluated
// by every wave; they need no exchange).
template <class PW, bool SPLIT, bool WIDE = true>
__device__ __forceinline__ double graph_prob(const cword *__restrict__ P,
                                             const cword *__restrict__ gh, const PW &pw,
                                             const LdsTables *lt, const Tables &tb, uint32_t &err,
                                             const Split &sp, uint32_t &xs) {
  const uint32_t gflags = gh[GH_FLAGS];
  const int e0 = (int)gh[GH_E0];
  const uint32_t nt = gh[GH_NTERMS];
  uint32_t t0 = 0, t1 = nt, off = gh[GH_TERM_OFF];
  // single-term and (in the kernels that have them) approximate graphs are
  // evaluated whole by every wave of a group
  constexpr uint32_t kWhole = kApprox<PW> ? (GF_MONO | GF_APPROX) : GF_MONO;
  const bool share = SPLIT && sp.L > 1 && !(gflags & kWhole);
  if (share) {
    const int per = GH_PARTS / sp.L;
    const int p0 = __builtin_amdgcn_readfirstlane(sp.part * per), p1 = p0 + per;
    if (p0 > 0) {
      t0 = gh[GH_PART_T1 + p0 - 1];
      off = gh[GH_PART_O1 + p0 - 1];
    }
    if (p1 < GH_PARTS) t1 = gh[GH_PART_T1 + p1 - 1];
  }
  GraphSum g;
  graph_sum_clear<WIDE>(g, gflags);
  if constexpr (kApprox<PW> && WIDE) {
    if (gflags & GF_APPROX) {
      g.fr = g.fi = 0.0;
      sum_terms<PW, WIDE>(P + off, t1 - t0, gflags, e0, pw, lt, tb, g, err, P + gh[GH_APX_OFF]);
      return __dsqrt_rn(__dadd_rn(__dmul_rn(g.fr, g.fr), __dmul_rn(g.fi, g.fi)));
    }
  }
  uint32_t tbuf = NO_TBUF;
  if (gflags & GF_TABLE) {
    // tables are read by the register-word kernels that hold a buffer for
    // the graph's width (the validator keeps them out of every other
    // program); elsewhere the graph is flagged and its terms skipped,
    // barriers still reached
    tbuf = (WIDE ? PW::kTab : PW::kTabNarrow) ? table_buffer(sp, xs) : NO_TBUF;
    if (tbuf == NO_TBUF) {
      err |= ERR_FORMAT;
      t1 = t0;
    }
  }
  sum_terms<PW, WIDE>(P + off, t1 - t0, gflags, e0, pw, lt, tb, g, err, nullptr, tbuf);
  if (gflags & GF_MONO) return g.pm;
  const bool narrow = !WIDE || (gflags & GF_NARROW);
  if constexpr (WIDE) {
    if (narrow) {
      g.n.a = low64(g.acc.a); g.n.b = low64(g.acc.b);
      g.n.c = low64(g.acc.c); g.n.d = low64(g.acc.d);
    }
  }
  if (share) exchange_sums(g, narrow, sp, xs);
  if (narrow) return magnitude(widen(g.n), e0);
  return magnitude(g.acc, e0);
}

template <int WL, bool SPLIT = false, bool WIDE = true>
__device__ __forceinline__ double graph_prob_any(const cword *__restrict__ P,
                                                 const cword *__restrict__ gh,
                                                 const uint32_t (&pv)[WL], const LdsTables *lt,
                                                 const Tables &tb, uint32_t &err, const Split &sp,
                                                 uint32_t &xs) {
  switch (gh[GH_W32]) {
#define KSIM_GP(w)                                                                       \
  case w:                                                                                \
    if constexpr (WL >= w)                                                               \
      return graph_prob<RegWords<w, WL>, SPLIT, WIDE>(P, gh, RegWords<w, WL>{pv}, lt, tb, err, \
                                                      sp, xs);                           \
    break;
    KSIM_GP(1) KSIM_GP(2) KSIM_GP(3) KSIM_GP(4) KSIM_GP(5) KSIM_GP(6) KSIM_GP(7) KSIM_GP(8)
#undef KSIM_GP
    default: break;
  }
  return 0.0;
}

// Exact amplitude of one graph for explicit parameter words (the exact
// path of k_eval): acc = sum_terms value * 2^{-e0} in 128 bits, every term
// through the general path whatever the graph's flags.
template <class PW>
__device__ __forceinline__ void eval_graph(const cword *__restrict__ P,
                                           const cword *__restrict__ gh, const PW &pw,
                                           const LdsTables *lt, const Tables &tb, Acc4 &acc,
                                           uint32_t &err, uint32_t tbuf) {
  GraphSum g;
  graph_sum_clear<true>(g, 0u);
  const uint32_t table = gh[GH_FLAGS] & GF_TABLE;
  uint32_t nt = gh[GH_NTERMS];
  if (table && !(PW::kTab && tbuf != NO_TBUF)) {
    err |= ERR_FORMAT;
    nt = 0;
  }
  sum_terms<PW>(P + gh[GH_TERM_OFF], nt, table, (int)gh[GH_E0], pw, lt, tb, g, err, nullptr,
                tbuf);
  acc = g.acc;
}

template <int WL>
__device__ __forceinline__ void eval_graph_any(const cword *__restrict__ P,
                                               const cword *__restrict__ gh,
                                               const uint32_t (&pv)[WL], const LdsTables *lt,
                                               const Tables &tb, Acc4 &acc, uint32_t &err,
                                               uint32_t tbuf) {
  switch (gh[GH_W32]) {
#define KSIM_EG(w)                                                                    \
  case w:                                                                             \
    if constexpr (WL >= w) eval_graph(P, gh, RegWords<w, WL>{pv}, lt, tb, acc, err, tbuf); \
    break;
    KSIM_EG(1) KSIM_EG(2) KSIM_EG(3) KSIM_EG(4) KSIM_EG(5) KSIM_EG(6) KSIM_EG(7) KSIM_EG(8)
#undef KSIM_EG
    default: break;
  }
}

// set parameter bit `pos` of the lane's words to `bit` (pos is wave-uniform:
// a scalar switch picks the word).  The compiler copies the other words at
// the switch's merge point, so the plain kernel's graph loop names the word
// at compile time instead (run_graphs); this form serves the split and
// batch kernels and the noise path, where it measured faster than
// rewriting every word with a uniform word mask (one VALU op per word)
template <int WL>
__device__ __forceinline__ void set_param_bit(uint32_t (&pv)[WL], int pos, uint32_t bit) {
  const int word = __builtin_amdgcn_readfirstlane(pos >> 5);
  const uint32_t m = 1u << (pos & 31);
  const uint32_t b = 0u - (bit & 1u);
  switch (word) {
#define KSIM_SET_WORD(w)                                   \
  case w:                                                   \
    if constexpr (w < WL) pv[w] = (pv[w] & ~m) | (m & b);   \
    break;
    KSIM_SET_WORD(0) KSIM_SET_WORD(1) KSIM_SET_WORD(2) KSIM_SET_WORD(3)
    KSIM_SET_WORD(4) KSIM_SET_WORD(5) KSIM_SET_WORD(6) KSIM_SET_WORD(7)
#undef KSIM_SET_WORD
    default: break;
  }
}

// Decide parameter bit `pos`, which holds 1 (the trying bit), to `bit`, and
// set bit pos + 1 (the next graph's trying bit) in the same word update when
// `next` and it lies in the same word.  Returns whether the next trying bit
// still has to be set (it starts a new word).
template <int WL>
__device__ __forceinline__ bool decide_param_bit(uint32_t (&pv)[WL], int pos, uint32_t bit,
                                                 bool next) {
  const int word = __builtin_amdgcn_readfirstlane(pos >> 5);
  const uint32_t m = 1u << (pos & 31);
  const bool same = next && (pos & 31) != 31;
  const uint32_t mn = same ? m << 1 : 0u;
  const uint32_t clr = m & ~(0u - (bit & 1u));  // the trying bit of lanes that decided 0
  switch (word) {
#define KSIM_DECIDE_WORD(w)                             \
  case w:                                               \
    if constexpr (w < WL) pv[w] = (pv[w] & ~clr) | mn;   \
    break;
    KSIM_DECIDE_WORD(0) KSIM_DECIDE_WORD(1) KSIM_DECIDE_WORD(2) KSIM_DECIDE_WORD(3)
    KSIM_DECIDE_WORD(4) KSIM_DECIDE_WORD(5) KSIM_DECIDE_WORD(6) KSIM_DECIDE_WORD(7)
#undef KSIM_DECIDE_WORD
    default: break;
  }
  return next && !same;
}

// parameter bit `pos` (wave-uniform) of the lane's words
template <int WL>
__device__ __forceinline__ uint32_t get_param_bit(const uint32_t (&pv)[WL], int pos) {
  const int word = __builtin_amdgcn_readfirstlane(pos >> 5);
  uint32_t x = 0;
  switch (word) {
#define KSIM_GET_WORD(w)              \
  case w:                             \
    if constexpr (w < WL) x = pv[w];  \
    break;
    KSIM_GET_WORD(0) KSIM_GET_WORD(1) KSIM_GET_WORD(2) KSIM_GET_WORD(3)
    KSIM_GET_WORD(4) KSIM_GET_WORD(5) KSIM_GET_WORD(6) KSIM_GET_WORD(7)
#undef KSIM_GET_WORD
    default: break;
  }
  return (x >> (pos & 31)) & 1u;
}

// four dwords at a dword-aligned address (gfx950 loads them in one
// instruction; no 16-byte alignment is assumed)
struct __attribute__((aligned(4))) U32x4 {
  uint32_t v[4];
};

// Device view of a noise model (tables built by build_noise, noise.hpp).
struct NoiseDev {
  const uint64_t *skip;    // C + 1
  const NoiseChan *chan;   // C
  const uint64_t *cdf;
  const uint32_t *pat;     // wf words per outcome row
  uint64_t seed;
  int32_t C, wf;
};

// SplitMix64 (S. Vigna): the per-shot noise stream.  next() advances the
// state by the golden-ratio increment and returns the mixed value.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t splitmix_next(uint64_t &x) {
  x += 0x9e3779b97f4a7c15ull;
  return mix64(x);
}
// start of shot `shot`'s stream under `seed`
__host__ __device__ inline uint64_t noise_stream(uint64_t seed, uint64_t shot) {
  return mix64(seed ^ mix64(shot + 0x632be59bd9b4e019ull));
}

// Walk the channels of one shot: calls fire(channel, outcome) for every
// channel that fires, in channel order.
template <class Fire>
__device__ __forceinline__ void noise_walk(const NoiseDev &nz, uint64_t x, Fire &&fire) {
  int c = 0;
  while (c < nz.C) {
    // number of channels skipped before the next candidate: the largest
    // k <= C - c with u < skip[k]  (P(k or more skipped) = (1 - p_max)^k)
    const uint64_t u = splitmix_next(x);
    int lo = 0, hi = nz.C - c;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (u < nz.skip[mid]) lo = mid; else hi = mid - 1;
    }
    c += lo;
    if (c >= nz.C) break;
    const NoiseChan ch = nz.chan[c];
    if (splitmix_next(x) <= ch.acc) {
      const uint64_t v = splitmix_next(x);
      uint32_t o = 0;
      while (o + 1 < ch.n_out && v > nz.cdf[ch.off_cdf + o]) ++o;
      fire(c, ch.off_pat + o * (uint32_t)nz.wf);
    }
    ++c;
  }
}

struct ShotArgs {
  const uint32_t *prog;
  const uint8_t *f;   // (B, num_f)
  const double *U;    // (B, tot_out) or null
  uint8_t *out;       // (B, num_outputs) u8, or (B, out_words) u32 bit rows
  uint32_t *err;
  int64_t B, shot_base;
  uint64_t seed;
  int32_t num_f, num_outputs, tot_out;
  int32_t out_words;  // 0: one byte per output; > 0: packed rows of this many words
  int32_t out_atomic; // packed rows shared by several components: OR words in
  uint32_t prefetch_dwords;  // leading program dwords to pull into L2 first (0: none)
  NoiseDev noise;     // noise.C > 0: f bits drawn on the device, f unused
};

// Pull this block's share of the first `ndw` program dwords into the L2 of
// the block's XCD: one 4-byte load per 128-byte line, spread over the
// blocks that run on that XCD (consecutive workgroup ids go to the 8 XCDs in
// turn).  Every wave then walks the program's records with scalar loads;
// from a cold L2 the waves of a launch that fits in one round of resident
// slots would otherwise walk the misses in lockstep, one HBM round trip at
// a time.  Loads only; no result is used.
__device__ __forceinline__ void prefetch_program(const uint32_t *prog, uint32_t ndw) {
  if (ndw == 0) return;
  const uint32_t nblk = gridDim.x * gridDim.y;
  const uint32_t id = blockIdx.y * gridDim.x + blockIdx.x;
  const uint32_t on_xcd = (nblk + 7 - (id & 7)) >> 3;  // blocks on this block's XCD
  const uint32_t lines = (ndw + 31) >> 5;
  const uint32_t chunk = (lines + on_xcd - 1) / on_xcd;
  const uint32_t l0 = (id >> 3) * chunk;
  const uint32_t l1 = min(lines, l0 + chunk);
  const volatile uint32_t *vp = prog;
  for (uint32_t l = l0 + threadIdx.x; l < l1; l += blockDim.x) (void)vp[(size_t)l << 5];
}

// 32 bits of the lane's parameter words starting at bit `off` (wave-uniform,
// may be negative or run past the words: missing bits read as 0)
template <int WL>
__device__ __forceinline__ uint32_t param_bits32(const uint32_t (&pv)[WL], int off) {
  const int w0 = off >> 5, r = off & 31;
  uint32_t lo = 0, hi = 0;
#pragma unroll
  for (int w = 0; w < WL; ++w) {
    lo = (w == w0) ? pv[w] : lo;
    hi = (w == w0 + 1) ? pv[w] : hi;
  }
  return r ? (lo >> r) | (hi << (32 - r)) : lo;
}

// write the shot's output row as packed bits: bit c of the row = output
// column c (little-endian words, so the byte view is numpy's
// packbits(bitorder="little") of the u8 row, padded to whole words)
template <int WL>
__device__ __forceinline__ void write_packed_row(const ShotArgs &a, const uint32_t (&pv)[WL],
                                                 const cword *__restrict__ ch,
                                                 const cword *__restrict__ dest, int F,
                                                 int nout, int64_t s) {
  uint32_t *orow = (uint32_t *)a.out + s * (int64_t)a.out_words;
  const uint32_t dest0 = ch[CH_DEST0];
  if (dest0 != 0xffffffffu) {
    // contiguous columns [dest0, dest0 + nout): row word k holds parameter
    // bits F - dest0 + 32 k .. + 31, masked to the component's columns
    const int k0 = (int)dest0 >> 5, k1 = ((int)dest0 + nout - 1) >> 5;
    for (int k = k0; k <= k1; ++k) {
      uint32_t v = param_bits32<WL>(pv, F - (int)dest0 + 32 * k);
      const int lo = max((int)dest0 - 32 * k, 0), hi = min((int)dest0 + nout - 32 * k, 32);
      const uint32_t m = (hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
      v &= m;
      if (a.out_atomic) {
        if (v) atomicOr(orow + k, v);
      } else {
        orow[k] = v;
      }
    }
    return;
  }
  // any column order: one bit per output, OR-ed into its word
  for (int i = 0; i < nout; ++i) {
    const uint32_t c = dest[i];
    const uint32_t bit = param_bits32<WL>(pv, F + i) & 1u;
    if (bit) atomicOr(orow + (c >> 5), 1u << (c & 31));
  }
}

// ------------------------------------------------------- f rows of a wave
// A wave's 64 shots are 64 consecutive f rows, i.e. one contiguous byte
// range.  Read lane by lane (each lane its own row, e.g. 100 bytes apart),
// every load instruction touches all of the range's cache lines and the
// same lines are fetched again by the next instruction after the L2 has
// turned over (about 3x the f bytes fetched from HBM for 100-byte rows at
// full occupancy).  Staged, the range is read once in aligned 16-byte
// pieces by consecutive lanes, and bit 0 of every byte goes into a per-wave
// LDS bit string; each lane then takes its parameter words out of it with
// funnel shifts.  Used when a row is at most 32 * WL bytes and f is 16-byte
// aligned (torch allocations are); the string then needs stage_words(WL)
// dwords.
__host__ __device__ constexpr int stage_words(int WL) { return 65 * WL + 2; }

struct FStage {
  uint32_t *st;   // the wave's bit string (stage_words(WL) dwords), or nullptr
  int64_t r0;     // first row of the wave's tile
  int nrows;      // rows of the tile inside f (1..64)
};

// bit 0 of each of 16 bytes, byte k -> bit k
__device__ __forceinline__ uint32_t low_bits16(const U32x4 &q) {
  uint32_t b = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) b |= (((q.v[k] & 0x01010101u) * 0x01020408u) >> 24) << (4 * k);
  return b;
}

// Fill the wave's bit string with rows [r0, r0 + nrows) of f: bit j of the
// string is bit 0 of byte (start + j), start = the first row's address
// rounded down to 16 bytes.  Every lane of the wave calls it with the same
// arguments.  Returns the string offset of row r0's first byte.
__device__ __forceinline__ int stage_f_rows(const ShotArgs &a, const FStage &fs) {
  const int lane = threadIdx.x & 63;
  const uint8_t *beg = a.f + fs.r0 * (int64_t)a.num_f;
  const uint8_t *start = (const uint8_t *)((uintptr_t)beg & ~(uintptr_t)15);
  const uint8_t *fend = a.f + a.B * (int64_t)a.num_f;
  const int lead = (int)(beg - start);
  const int n16 = (lead + fs.nrows * a.num_f + 15) >> 4;
  uint16_t *st16 = (uint16_t *)fs.st;
  // the string is rewritten for every tile: finish the previous tile's reads
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // pieces entirely inside f, four per lane in flight before the first wait
  const int nfull = (int)min((int64_t)n16, (int64_t)((fend - start) >> 4));
  for (int i0 = 0; i0 < nfull; i0 += 4 * 64) {
    U32x4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 64 * u + lane;
      q[u] = i < nfull ? *(const U32x4 *)(start + 16 * (int64_t)i) : U32x4{{0u, 0u, 0u, 0u}};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 64 * u + lane;
      if (i < nfull) st16[i] = (uint16_t)low_bits16(q[u]);
    }
  }
  // the piece that runs past the end of f (the last row of f): byte by byte
  if (nfull < n16 && lane == 0) {
    const uint8_t *p = start + 16 * (int64_t)nfull;
    uint32_t b = 0;
    for (int k = 0; p + k < fend; ++k) b |= (uint32_t)(p[k] & 1u) << k;
    st16[nfull] = (uint16_t)b;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return lead;
}

// the lane's parameter words: F bits of the staged string from bit `off`
template <int WL>
__device__ __forceinline__ void staged_words(const uint32_t *st, int off, int F,
                                             uint32_t (&pv)[WL]) {
#pragma unroll
  for (int w = 0; w < WL; ++w) {
    uint32_t v = 0;
    if (32 * w < F) {
      const int j = off + 32 * w;
      const uint32_t lo = st[j >> 5], hi = st[(j >> 5) + 1];
      v = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(j & 31));
      const int left = F - 32 * w;
      if (left < 32) v &= (1u << left) - 1u;
    }
    pv[w] = v;
  }
}

// Graphs i, i + 1, ... of mask width W32 (up to the first of another
// width) of one shot: graph 0 gives the normaliser, graph i > 0 the
// probability that output i - 1 is 1, with its trying bit (parameter
// F + i - 1) at 1, and the drawn bit then stays in that parameter.  Graph
// i has P = F + i parameters and W32 = ceil((F + i) / 32) mask words
// (checked at upload), so parameter F + i - 1 lies in word W32 - 1: the
// bit is set and decided in a register named at compile time.  (A
// switch over a runtime word index here made the compiler copy all WL
// words at every graph.)  Returns the first graph not run.
template <int W32, int WL, bool SPLIT, bool WIDE>
__device__ __forceinline__ int run_graphs(const ShotArgs &a, const cword *__restrict__ P,
                                          const cword *__restrict__ ghs, int i, int ng, int F,
                                          int colbase, int64_t s, uint32_t (&pv)[WL],
                                          const LdsTables *lt, const Tables &tb, const Split &sp,
                                          double &prev, uint64_t &rs, uint32_t &err,
                                          uint32_t &xs) {
  static_assert(W32 <= WL, "graph wider than the lane's parameter words");
  for (; i < ng; ++i) {
    const cword *gh = ghs + (size_t)i * GH_DWORDS;
    if ((int)gh[GH_W32] != W32) break;
    const uint32_t m = 1u << ((F + i - 1) & 31);
    if (i > 0) pv[W32 - 1] |= m;  // trying bit
    const double p1 =
        graph_prob<RegWords<W32, WL>, SPLIT, WIDE>(P, gh, RegWords<W32, WL>{pv}, lt, tb, err, sp, xs);
    if (i == 0) {
      prev = p1;
      continue;
    }
    const double u = a.U ? a.U[s * (int64_t)a.tot_out + colbase + i - 1] : rng_uniform(rs);
    const uint32_t bit = (u < __ddiv_rn(p1, prev)) ? 1u : 0u;
    pv[W32 - 1] &= ~(m & (bit - 1u));  // lanes that drew 0 clear the trying bit
    prev = bit ? p1 : __dsub_rn(prev, p1);
  }
  return i;
}

// The whole autoregressive loop of one shot of component `comp`
// (src/cpp/ksim/kokkos_sim.cpp:214-230, src/python/ksim/sample/evaluate.py:76-85).
// SPLIT: the waves of sp's group share each graph's terms; every lane runs
// the loop (lanes without a shot of their own run a copy of a valid shot,
// so the group's barriers are reached by whole waves) and only `valid`
// lanes of the group's first wave write.  Lanes that are not `valid` run a
// copy of a valid shot of their wave in the plain path too, so that a
// staged f read (fs.st set) has every lane of the wave; fs then names the
// wave's tile, and s lies in it.
template <int WL, bool NOISE = false, bool SPLIT = false, bool WIDE = true>
__device__ __forceinline__ void run_shot(const ShotArgs &a, int comp, int64_t s,
                                         const LdsTables *lt, const Tables &tb,
                                         const Split &sp = Split{1, 0, 0, nullptr},
                                         bool valid = true, uint32_t *xs_io = nullptr,
                                         const FStage &fs = FStage{nullptr, 0, 0}) {
  const cword *__restrict__ P = as_const(a.prog);
  const cword *__restrict__ ch = P + P[PH_DWORDS + comp];
  const int ng = (int)ch[CH_NGRAPHS], F = (int)ch[CH_F];
  const int colbase = (int)ch[CH_COLBASE];
  const cword *__restrict__ fsel = P + ch[CH_FSEL_OFF];
  const cword *__restrict__ dest = P + ch[CH_DEST_OFF];
  const cword *__restrict__ ghs = P + ch[CH_GRAPH_OFF];
  const uint8_t *frow = a.f + s * (int64_t)a.num_f;
  uint32_t pv[WL];
  const uint32_t fsel0 = ch[CH_FSEL0];
  if constexpr (NOISE) {
    // f bits drawn here: XOR of the fired outcomes' patterns, seen through
    // this component's f_selection
#pragma unroll
    for (int w = 0; w < WL; ++w) pv[w] = 0;
    noise_walk(a.noise, noise_stream(a.noise.seed, (uint64_t)(a.shot_base + s)),
               [&](int, uint32_t row) {
                 const uint32_t *pr = a.noise.pat + row;
                 if (fsel0 != 0xffffffffu) {
#pragma unroll
                   for (int w = 0; w < WL; ++w) {
                     const int off = (int)fsel0 + 32 * w;
                     const int wi = off >> 5, r = off & 31;
                     uint32_t lo = wi < a.noise.wf ? pr[wi] : 0u;
                     uint32_t hi = wi + 1 < a.noise.wf ? pr[wi + 1] : 0u;
                     uint32_t v = r ? (lo >> r) | (hi << (32 - r)) : lo;
                     const int left = F - 32 * w;
                     v &= left >= 32 ? 0xffffffffu : (left > 0 ? (1u << left) - 1u : 0u);
                     pv[w] ^= v;
                   }
                 } else {
                   for (int j = 0; j < F; ++j) {
                     const uint32_t b = fsel[j];
                     const uint32_t bit = (pr[b >> 5] >> (b & 31)) & 1u;
                     set_param_bit<WL>(pv, j, get_param_bit<WL>(pv, j) ^ bit);
                   }
                 }
               });
  } else if (fs.st != nullptr && a.num_f <= 32 * WL && ((uintptr_t)a.f & 15u) == 0) {
    // the wave's rows through its LDS bit string (one read of each line)
    const int lead = stage_f_rows(a, fs);
    const int row = lead + (int)(s - fs.r0) * a.num_f;
    if (fsel0 != 0xffffffffu) {
      staged_words<WL>(fs.st, row + (int)fsel0, F, pv);
    } else {
      // any column order: each selected bit from the string (one LDS read),
      // a word at a time so every word index is known at compile time
#pragma unroll
      for (int w = 0; w < WL; ++w) {
        uint32_t word = 0;
        const int jn = min(32, F - 32 * w);
        for (int jj = 0; jj < jn; ++jj) {
          const int b = row + (int)fsel[32 * w + jj];
          word |= ((fs.st[b >> 5] >> (b & 31)) & 1u) << jj;
        }
        pv[w] = word;
      }
    }
  } else if (fsel0 != 0xffffffffu &&
             (((uint32_t)a.num_f | fsel0 | (uint32_t)(uintptr_t)a.f) & 3u) == 0) {
    // contiguous selection on a dword-aligned row: 16 f bytes per load,
    // bit 0 of each byte gathered by one multiply per dword (bits 24..27 of
    // the product)
    const uint8_t *base = frow + fsel0;
#pragma unroll
    for (int w = 0; w < WL; ++w) {
      uint32_t word = 0;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = 32 * w + 16 * h;
        if (j + 16 <= F) {
          const U32x4 q = *(const U32x4 *)(base + j);
          uint32_t b = 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) b |= (((q.v[k] & 0x01010101u) * 0x01020408u) >> 24) << (4 * k);
          word |= b << (16 * h);
        } else {
          for (int jj = max(j, 0); jj < min(j + 16, F); ++jj) {
            if (((jj & 3) == 0) && jj + 4 <= F) {
              const uint32_t x = *(const uint32_t *)(base + jj) & 0x01010101u;
              word |= ((x * 0x01020408u) >> 24) << (jj & 31);
              jj += 3;
            } else {
              word |= (uint32_t)(base[jj] & 1u) << (jj & 31);
            }
          }
        }
      }
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
  uint32_t xs = xs_io ? *xs_io : 0u;  // exchange state of the block (split launches)
  double prev = 0.0;
  // graph 0 gives the normaliser, graph i+1 the probability of output i = 1;
  // the decided bits stay in pv (parameter F + i) until the row is written
  // at the end.  The plain kernel runs the graphs a width at a time, in
  // loops specialised on it (run_graphs); the kernels whose waves share
  // terms keep one evaluation site and a runtime word index, because the
  // specialised loops there cost registers (spills made the 128-bit split
  // kernel up to 39 % slower)
  if constexpr (SPLIT) {
    // graph i > 0 runs with its trying bit (parameter F + i - 1) at 1; graph 0
    // does not read parameter F, so graph 1's trying bit is set up front
    if (ng > 1) set_param_bit<WL>(pv, F, 1u);
    for (int i = 0; i < ng; ++i) {
      const cword *gh = ghs + (size_t)i * GH_DWORDS;
      const double p1 = graph_prob_any<WL, SPLIT, WIDE>(P, gh, pv, lt, tb, err, sp, xs);
      if (i == 0) {
        prev = p1;
        continue;
      }
      const double u =
          a.U ? a.U[s * (int64_t)a.tot_out + colbase + i - 1] : rng_uniform(rs);
      const uint32_t bit = (u < __ddiv_rn(p1, prev)) ? 1u : 0u;
      if (decide_param_bit<WL>(pv, F + i - 1, bit, i + 1 < ng))
        set_param_bit<WL>(pv, F + i, 1u);
      prev = bit ? p1 : __dsub_rn(prev, p1);
    }
  } else {
    for (int i = 0; i < ng;) {
      switch (ghs[(size_t)i * GH_DWORDS + GH_W32]) {
#define KSIM_RUN(w)                                                                          \
  case w:                                                                                  \
    if constexpr (WL >= w)                                                                 \
      i = run_graphs<w, WL, SPLIT, WIDE>(a, P, ghs, i, ng, F, colbase, s, pv, lt, tb, sp,  \
                                         prev, rs, err, xs);                               \
    else                                                                                   \
      i = ng;                                                                              \
    break;
        KSIM_RUN(1) KSIM_RUN(2) KSIM_RUN(3) KSIM_RUN(4)
        KSIM_RUN(5) KSIM_RUN(6) KSIM_RUN(7) KSIM_RUN(8)
#undef KSIM_RUN
        default: i = ng; break;  // widths are checked at upload
      }
    }
  }
  if (xs_io) *xs_io = xs;
  if (!valid || (SPLIT && sp.part != 0)) {
    if (err) atomicOr(a.err, err);
    return;
  }
  // the output row in one burst of stores, so each cache line of the rows is
  // completed while it is resident
  if (a.out_words > 0) {
    write_packed_row<WL>(a, pv, ch, dest, F, ng - 1, s);
  } else {
    // output i is parameter F + i: walk the words that hold them with
    // indices known at compile time (no runtime word switch per output)
    uint8_t *orow = a.out + s * (int64_t)a.num_outputs;
    const int nout = ng - 1;
#pragma unroll
    for (int w = 0; w < WL; ++w) {
      const int i0 = max(32 * w - F, 0), i1 = min(32 * w + 32 - F, nout);
      for (int i = i0; i < i1; ++i) orow[dest[i]] = (uint8_t)((pv[w] >> ((F + i) & 31)) & 1u);
    }
  }
  if (err) atomicOr(a.err, err);
}

// ------------------------------------------------ parameter words in LDS
// A component of more than 32 * REG_WORDS_MAX parameters keeps each lane's
// parameter words in LDS: the wave's area holds CH_WMAX words per lane,
// word w of the lane at pvl[64 w] (pvl = area + lane).  Only the lane
// itself reads and writes its words, so no barrier orders them.

__device__ __forceinline__ uint32_t lds_bit(const uint32_t *pvl, int pos) {
  return (pvl[64 * (pos >> 5)] >> (pos & 31)) & 1u;
}

__device__ __forceinline__ void lds_set_bit(uint32_t *pvl, int pos, uint32_t bit) {
  uint32_t &w = pvl[64 * (pos >> 5)];
  const uint32_t m = 1u << (pos & 31);
  w = (w & ~m) | (m & (0u - (bit & 1u)));
}

// 32 bits of the lane's nw words from bit `off` (may be negative or run
// past the words: missing bits read as 0)
__device__ __forceinline__ uint32_t lds_bits32(const uint32_t *pvl, int off, int nw) {
  const int w0 = off >> 5, r = off & 31;
  const uint32_t lo = (w0 >= 0 && w0 < nw) ? pvl[64 * w0] : 0u;
  const uint32_t hi = (w0 + 1 >= 0 && w0 + 1 < nw) ? pvl[64 * (w0 + 1)] : 0u;
  return r ? (lo >> r) | (hi << (32 - r)) : lo;
}

// The autoregressive loop of one shot of component `comp` with the lane's
// parameter words in LDS at pvl (run_shot's contract otherwise: SPLIT
// shares the graphs' terms over sp's group, only `valid` lanes of the
// group's first wave write).
template <bool NOISE, bool SPLIT, bool WIDE>
__device__ __forceinline__ void run_shot_lds(const ShotArgs &a, int comp, int64_t s,
                                             const LdsTables *lt, const Tables &tb, uint32_t *pvl,
                                             const Split &sp, bool valid, uint32_t *xs_io) {
  const cword *__restrict__ P = as_const(a.prog);
  const cword *__restrict__ ch = P + P[PH_DWORDS + comp];
  const int ng = (int)ch[CH_NGRAPHS], F = (int)ch[CH_F], nw = (int)ch[CH_WMAX];
  const int colbase = (int)ch[CH_COLBASE];
  const cword *__restrict__ fsel = P + ch[CH_FSEL_OFF];
  const cword *__restrict__ dest = P + ch[CH_DEST_OFF];
  const cword *__restrict__ ghs = P + ch[CH_GRAPH_OFF];
  const uint32_t fsel0 = ch[CH_FSEL0];
  if constexpr (NOISE) {
    for (int w = 0; w < nw; ++w) pvl[64 * w] = 0u;
    noise_walk(a.noise, noise_stream(a.noise.seed, (uint64_t)(a.shot_base + s)),
               [&](int, uint32_t row) {
                 const uint32_t *pr = a.noise.pat + row;
                 for (int j = 0; j < F; ++j) {
                   const uint32_t b = fsel[j];
                   pvl[64 * (j >> 5)] ^= ((pr[b >> 5] >> (b & 31)) & 1u) << (j & 31);
                 }
               });
  } else {
    const uint8_t *frow = a.f + s * (int64_t)a.num_f;
    const bool aligned = fsel0 != 0xffffffffu &&
                         (((uint32_t)a.num_f | fsel0 | (uint32_t)(uintptr_t)a.f) & 3u) == 0;
    for (int w = 0; w < nw; ++w) {
      uint32_t word = 0;
      const int j0 = 32 * w, jn = min(32, F - j0);
      if (aligned && jn == 32) {
        // 32 contiguous f bytes in two 16-byte loads, bit 0 of each byte
        const U32x4 q0 = *(const U32x4 *)(frow + fsel0 + j0);
        const U32x4 q1 = *(const U32x4 *)(frow + fsel0 + j0 + 16);
        word = low_bits16(q0) | (low_bits16(q1) << 16);
      } else {
        for (int jj = 0; jj < jn; ++jj) word |= (uint32_t)(frow[fsel[j0 + jj]] & 1u) << jj;
      }
      pvl[64 * w] = word;
    }
  }
  uint64_t rs = rng_seed(a.seed + 0x9E37ull * (uint64_t)comp, (uint64_t)(a.shot_base + s));
  uint32_t err = 0;
  uint32_t xs = xs_io ? *xs_io : 0u;
  double prev = 0.0;
  // graph i > 0 runs with its trying bit (parameter F + i - 1) at 1
  if (ng > 1) lds_set_bit(pvl, F, 1u);
  for (int i = 0; i < ng; ++i) {
    const cword *gh = ghs + (size_t)i * GH_DWORDS;
    const double p1 =
        graph_prob<LdsWords, SPLIT, WIDE>(P, gh, LdsWords{pvl, (int)gh[GH_W32]}, lt, tb, err, sp, xs);
    if (i == 0) {
      prev = p1;
      continue;
    }
    const double u = a.U ? a.U[s * (int64_t)a.tot_out + colbase + i - 1] : rng_uniform(rs);
    const uint32_t bit = (u < __ddiv_rn(p1, prev)) ? 1u : 0u;
    lds_set_bit(pvl, F + i - 1, bit);
    if (i + 1 < ng) lds_set_bit(pvl, F + i, 1u);
    prev = bit ? p1 : __dsub_rn(prev, p1);
  }
  if (xs_io) *xs_io = xs;
  if (!valid || (SPLIT && sp.part != 0)) {
    if (err) atomicOr(a.err, err);
    return;
  }
  const int nout = ng - 1;
  if (a.out_words > 0) {
    uint32_t *orow = (uint32_t *)a.out + s * (int64_t)a.out_words;
    const uint32_t dest0 = ch[CH_DEST0];
    if (dest0 != 0xffffffffu) {
      const int k0 = (int)dest0 >> 5, k1 = ((int)dest0 + nout - 1) >> 5;
      for (int k = k0; k <= k1; ++k) {
        uint32_t v = lds_bits32(pvl, F - (int)dest0 + 32 * k, nw);
        const int lo = max((int)dest0 - 32 * k, 0), hi = min((int)dest0 + nout - 32 * k, 32);
        v &= (hi >= 32 ? 0xffffffffu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
        if (a.out_atomic) {
          if (v) atomicOr(orow + k, v);
        } else {
          orow[k] = v;
        }
      }
    } else {
      for (int i = 0; i < nout; ++i) {
        const uint32_t c = dest[i];
        if (lds_bit(pvl, F + i)) atomicOr(orow + (c >> 5), 1u << (c & 31));
      }
    }
  } else {
    uint8_t *orow = a.out + s * (int64_t)a.num_outputs;
    for (int i = 0; i < nout; ++i) orow[dest[i]] = (uint8_t)lds_bit(pvl, F + i);
  }
  if (err) atomicOr(a.err, err);
}

// stage the central rows of the magnitude table and the pi powers
__device__ inline void load_tables(LdsTables *t, const Tables &tb) {
  for (int i = threadIdx.x; i < PI_MAX * PI_MAX; i += blockDim.x) t->pi[i] = tb.pi[i];
  const double *msrc = tb.mag + (TAB_W - MAG_LDS_R) * 32;
  for (int i = threadIdx.x; i < MAG_LDS_N; i += blockDim.x) t->mag[i] = msrc[i];
  __syncthreads();
}

// magnitude() of every table entry (all digit planes summed), exponent 0:
// the table GF_MONO graphs read their float64 magnitude from
__device__ inline double table_magnitude(const int4 *tab, int i) {
  Acc4 v;
  v.a = v.b = v.c = v.d = 0;
  for (int d = 0; d < TAB_DIGITS; ++d) {
    const int4 e = tab[d * TAB_N + i];
    const int sh = TAB_DIGIT_BITS * d;
    v.a += (i128_t)((u128_t)(i128_t)e.x << sh);
    v.b += (i128_t)((u128_t)(i128_t)e.y << sh);
    v.c += (i128_t)((u128_t)(i128_t)e.z << sh);
    v.d += (i128_t)((u128_t)(i128_t)e.w << sh);
  }
  return magnitude(v, 0);
}

}  // namespace ksim
