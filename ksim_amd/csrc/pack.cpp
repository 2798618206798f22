// pack.cpp — FlatProgram graphs -> device program words.
//
// Replaces the reference's host packing (src/cpp/ksim/kokkos_sim.cpp:73-151,
// which concatenates the u8/i32 arrays and a 25-int meta row per graph).  Here
// each graph is rewritten, term by term, into records the gfx950 kernel
// evaluates with bit-packed masks and packed exponent adds:
//
//   PROJ  (mask, expect)          term is zero unless parity == expect
//                                 (node phase 0 or 4: factor 2 or 0)
//   PPX   (mask_psi, mask_phi,    pi-product: S += 4 (psi & phi)
//          psi_c, phi_c)
//   LIN   (mask, delta)           exponents += delta if parity is 1
//                                 (node phases, half-pi terms, one-sided pairs)
//   PAIR  (mask_a, mask_b,        phase pair with both parities live
//          d10, d01, d11)
//
// Parameter-independent factors (zero masks) and the prefactor
// w^{pf_phase_idx} * floatfactor * 2^{pf_power2} are folded into per-term
// constants.  Terms with a constant zero factor are dropped.  The per-family
// semantics follow src/cpp/ksim/zx_eval.hpp:102-165 and
// src/python/ksim/sample/evaluate.py:32-61.
#include "pack.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <thread>
#include <mutex>
#include <stdexcept>
#include <string>

namespace ksim {

namespace {

struct FacVal {
  bool zero = false;
  int s = 0, v = 0, n = 0, a = 0, b = 0;
};

FacVal g_np[8];      // 1 + w^k
FacVal g_ph[8][8];   // 1 + w^a + w^b - w^{a+b}
std::once_flag g_once;

FacVal decompose_small(const Zw &x) {
  FacVal f;
  if (zw_is_zero(x)) { f.zero = true; return f; }
  Fact ft = zw_factor(x);
  if (!ft.pure) throw std::logic_error("factor table: value does not factor");
  f.s = ft.s; f.v = ft.v; f.n = ft.n; f.a = ft.a; f.b = ft.b;
  if (!zw_eq(zw_build(f.s, f.v, f.n, f.a, f.b), x))
    throw std::logic_error("factor table: rebuild mismatch");
  return f;
}

void init_tables() {
  for (int k = 0; k < 8; ++k) {
    Zw u = zw_unit(k);
    u.a += 1;
    g_np[k] = decompose_small(u);
  }
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      Zw u = zw_unit(a), v = zw_unit(b), w = zw_unit(a + b);
      Zw t{1 + u.a + v.a - w.a, u.b + v.b - w.b, u.c + v.c - w.c, u.d + v.d - w.d};
      g_ph[a][b] = decompose_small(t);
    }
}

inline uint32_t pack_exp(const FacVal &e) {
  return (uint32_t)e.v << PK_OV | (uint32_t)(e.n + NBIAS) << PK_ON | (uint32_t)e.a << PK_OA |
         (uint32_t)e.b << PK_OB | (uint32_t)(e.s & 7) << PK_OS;
}

// A LIN delta that is, mod 2^32, s 2^30 + t 2^9 with s, t in {-1, 0, 1}
// (S += 2s mod 8 and N += t in the packed word): a sign record.
bool sign_delta(uint32_t d, int &s, int &t) {
  for (s = -1; s <= 1; ++s)
    for (t = -1; t <= 1; ++t)
      if ((s || t) && d == (uint32_t)s * (1u << 30) + (uint32_t)t * (1u << 9)) return true;
  return false;
}

struct Mask {
  std::vector<uint32_t> w;
  bool zero() const {
    for (uint32_t x : w) if (x) return false;
    return true;
  }
};

Mask make_mask(const uint8_t *row, int P, int W32) {
  Mask m;
  m.w.assign(W32, 0u);
  for (int j = 0; j < P; ++j)
    if (row[j] & 1u) m.w[j >> 5] |= 1u << (j & 31);
  return m;
}

Mask mask_xor(const Mask &x, const Mask &y) {
  Mask m = x;
  for (size_t k = 0; k < m.w.size(); ++k) m.w[k] ^= y.w[k];
  return m;
}

bool mask_eq(const Mask &x, const Mask &y) { return x.w == y.w; }

struct LinRec { Mask m; FacVal e0, e1; };
struct PairRec { Mask ma, mb; FacVal e00, e10, e01, e11; };
struct ProjRec { Mask m; uint32_t expect; };
struct PpxRec { Mask ms, mf; uint32_t sc, fc; };

struct TermBuild {
  bool zero = false;
  // constant exponents
  int S = 0, V = 0, N = 0, A = 0, B = 0, E = 0;
  Zw R{1, 0, 0, 0};
  bool residual = false;
  std::vector<ProjRec> proj;
  std::vector<PpxRec> ppx;
  std::vector<LinRec> lin;
  std::vector<PairRec> pair;

  void add_const(const FacVal &f) {
    if (f.zero) { zero = true; return; }
    S += f.s; V += f.v; N += f.n; A += f.a; B += f.b;
  }
  void add_lin(const Mask &m, const FacVal &e0, const FacVal &e1) {
    if (e0.zero || e1.zero) throw std::logic_error("LIN record with a zero value");
    if (m.zero()) { add_const(e0); return; }
    lin.push_back(LinRec{m, e0, e1});
  }
};

int ceil_log2(int64_t x) {
  int b = 0;
  while ((int64_t(1) << b) < x) ++b;
  return b;
}

}  // namespace

const int32_t *materialise_table() {
  // TAB_DIGITS planes of w^s (1+w)^r (1+sqrt2)^n, n in [-TAB_W, TAB_W], as
  // balanced base-2^31 digits
  static std::vector<int32_t> tab;
  static std::once_flag once;
  std::call_once(once, [] {
    tab.assign((size_t)TAB_DIGITS * TAB_N * 4, 0);
    for (int n = -TAB_W; n <= TAB_W; ++n)
      for (int r = 0; r < 4; ++r)
        for (int s = 0; s < 8; ++s) {
          const Zw x = zw_mul(zw_mul(zw_unit(s), zw_pow(zw_lambda(), r)), zw_unit_pow(n));
          const i128 c[4] = {x.a, x.b, x.c, x.d};
          for (int k = 0; k < 4; ++k) {
            i128 v = c[k];
            for (int d = 0; d < TAB_DIGITS; ++d) {
              const i128 lo = v & (((i128)1 << TAB_DIGIT_BITS) - 1);
              const i128 dig = lo >= ((i128)1 << (TAB_DIGIT_BITS - 1)) ? lo - ((i128)1 << TAB_DIGIT_BITS) : lo;
              if (dig >= ((i128)1 << 30) || dig < -((i128)1 << 30))
                throw std::logic_error("table digit too large");
              tab[((size_t)d * TAB_N + tab_index(n, r, s)) * 4 + k] = (int32_t)dig;
              v = (v - dig) >> TAB_DIGIT_BITS;
            }
            if (v != 0) throw std::logic_error("table entry needs more digits");
            if (std::abs(n) <= TAB_R && c[k] != (i128)tab[(size_t)tab_index(n, r, s) * 4 + k])
              throw std::logic_error("central table row is not a single digit");
          }
        }
    for (int k = 0; k < 4; ++k)
      if (tab[(size_t)TAB_ZERO * 4 + k] != 0) throw std::logic_error("zero table entry is not zero");
  });
  return tab.data();
}

// bits of the largest table coefficient over rows |n| <= nmax
static int table_bits(int nmax) {
  static std::vector<int> bits;
  static std::once_flag once;
  std::call_once(once, [] {
    bits.assign(TAB_W + 1, 0);
    for (int n = -TAB_W; n <= TAB_W; ++n)
      for (int r = 0; r < 4; ++r)
        for (int s = 0; s < 8; ++s) {
          const Zw x = zw_mul(zw_mul(zw_unit(s), zw_pow(zw_lambda(), r)), zw_unit_pow(n));
          int &b = bits[std::abs(n)];
          b = std::max(b, i128_bits(zw_maxabs(x)));
        }
    for (int n = 1; n <= TAB_W; ++n) bits[n] = std::max(bits[n], bits[n - 1]);
  });
  return bits[std::min(nmax, TAB_W)];
}

const int32_t *pi_table() {
  static std::vector<int32_t> tab;
  static std::once_flag once;
  std::call_once(once, [] {
    tab.assign(PI_MAX * PI_MAX * 4, 0);
    for (int a = 0; a < PI_MAX; ++a)
      for (int b = 0; b < PI_MAX; ++b) {
        Zw x = zw_mul(zw_pow(zw_pi(), a), zw_pow(zw_pibar(), b));
        if (zw_maxabs(x) >= ((i128)1 << 30)) throw std::logic_error("pi table entry too large");
        // pi = 1 + i sqrt2 = 1 + w - conj w, so every power lies in Z[i sqrt2]:
        // (P0, P1, 0, -P1) in the basis (1, w, i, conj w); the kernels use
        // only P0 and P1
        if (x.c != 0 || x.d != -x.b) throw std::logic_error("pi power outside Z[i sqrt2]");
        int32_t *e = &tab[(a * PI_MAX + b) * 4];
        e[0] = (int32_t)x.a; e[1] = (int32_t)x.b; e[2] = (int32_t)x.c; e[3] = (int32_t)x.d;
      }
  });
  return tab.data();
}

static int pi_table_bits(int amax, int bmax) {
  const int32_t *t = pi_table();
  int bits = 0;
  for (int a = 0; a <= std::min(amax, PI_MAX - 1); ++a)
    for (int b = 0; b <= std::min(bmax, PI_MAX - 1); ++b)
      for (int k = 0; k < 4; ++k) bits = std::max(bits, i128_bits(t[(a * PI_MAX + b) * 4 + k]));
  return bits;
}

static void build_term(const ksim_graph_view &g, int t, int W32, TermBuild &tb) {
  std::call_once(g_once, init_tables);
  const int P = g.P;
  // ---- node phases (zx_eval.hpp:106-117): prod_{j < np_counts[t]} 1 + w^{4 par + phi}
  const int ncnt = g.np_counts ? g.np_counts[t] : 0;
  if (ncnt < 0 || ncnt > g.T_np) throw std::invalid_argument("np_counts out of range");
  for (int j = 0; j < ncnt; ++j) {
    const int phi = g.np_phases[(size_t)t * g.T_np + j] & 7;
    Mask m = make_mask(g.np_params + ((size_t)t * g.T_np + j) * P, P, W32);
    if (m.zero()) { tb.add_const(g_np[phi]); continue; }
    if (phi == 0 || phi == 4) {  // 2 if parity == [phi == 4] else 0
      tb.add_const(g_np[0]);
      tb.proj.push_back(ProjRec{m, phi == 4 ? 1u : 0u});
    } else {
      tb.add_lin(m, g_np[phi], g_np[phi ^ 4]);
    }
  }
  // ---- half-pi (zx_eval.hpp:118-128): w^{sum par * c mod 8}
  for (int j = 0; j < g.T_hp; ++j) {
    const int c = g.hp_coeffs[(size_t)t * g.T_hp + j] & 7;
    if (c == 0) continue;
    Mask m = make_mask(g.hp_params + ((size_t)t * g.T_hp + j) * P, P, W32);
    if (m.zero()) continue;
    FacVal e0, e1;
    e1.s = c;
    tb.add_lin(m, e0, e1);
  }
  // ---- pi products (zx_eval.hpp:129-142): (-1)^{sum psi phi}
  for (int j = 0; j < g.T_pp; ++j) {
    const size_t o = (size_t)t * g.T_pp + j;
    const uint32_t sc = g.pp_psi_c[o] & 1u, fc = g.pp_phi_c[o] & 1u;
    Mask ms = make_mask(g.pp_psi_p + o * P, P, W32);
    Mask mf = make_mask(g.pp_phi_p + o * P, P, W32);
    FacVal e0, e4;
    e4.s = 4;
    if (ms.zero() && mf.zero()) {
      if (sc & fc) tb.S += 4;
    } else if (ms.zero()) {  // psi constant
      if (sc) tb.add_lin(mf, fc ? e4 : e0, fc ? e0 : e4);
    } else if (mf.zero()) {
      if (fc) tb.add_lin(ms, sc ? e4 : e0, sc ? e0 : e4);
    } else if (mask_eq(ms, mf)) {  // psi, phi share the parity p
      if (sc == fc) tb.add_lin(ms, sc ? e4 : e0, sc ? e0 : e4);  // (sc^p) & (sc^p)
      // sc != fc: one of psi, phi is 0 for every p
    } else {
      tb.ppx.push_back(PpxRec{ms, mf, sc, fc});
    }
  }
  // ---- phase pairs (zx_eval.hpp:143-158): prod 1 + w^a + w^b - w^{a+b}
  const int pcnt = g.ph_counts ? g.ph_counts[t] : 0;
  if (pcnt < 0 || pcnt > g.T_ph) throw std::invalid_argument("ph_counts out of range");
  for (int j = 0; j < pcnt; ++j) {
    const size_t o = (size_t)t * g.T_ph + j;
    const int al = g.ph_alpha[o] & 7, be = g.ph_beta[o] & 7;
    Mask ma = make_mask(g.ph_alpha_p + o * P, P, W32);
    Mask mb = make_mask(g.ph_beta_p + o * P, P, W32);
    if (ma.zero() && mb.zero()) { tb.add_const(g_ph[al][be]); continue; }
    if ((al & 3) == 2 && (be & 3) == 2) {
      // a', b' in {2, 6}: zero unless a' == b', i.e. par_a ^ par_b == [al != be];
      // then the value is F(a', a') with a' = al ^ 4 par_a.
      Mask mx = mask_xor(ma, mb);
      const uint32_t expect = al != be ? 1u : 0u;
      if (mx.zero()) {
        if (expect) { tb.zero = true; return; }
      } else {
        tb.proj.push_back(ProjRec{mx, expect});
      }
      if (ma.zero()) {
        // a' = al fixed; alive only when b' == al
        tb.add_const(g_ph[al][al]);
      } else {
        tb.add_lin(ma, g_ph[al][al], g_ph[al ^ 4][al ^ 4]);
      }
      continue;
    }
    if (ma.zero()) { tb.add_lin(mb, g_ph[al][be], g_ph[al][be ^ 4]); continue; }
    if (mb.zero()) { tb.add_lin(ma, g_ph[al][be], g_ph[al ^ 4][be]); continue; }
    if (mask_eq(ma, mb)) { tb.add_lin(ma, g_ph[al][be], g_ph[al ^ 4][be ^ 4]); continue; }
    tb.pair.push_back(PairRec{ma, mb, g_ph[al][be], g_ph[al ^ 4][be], g_ph[al][be ^ 4],
                              g_ph[al ^ 4][be ^ 4]});
  }
  // ---- prefactor (zx_eval.hpp:159-170): w^idx * floatfactor * 2^power2
  {
    Zw ff{g.pf_floatfactor[(size_t)t * 4 + 0], g.pf_floatfactor[(size_t)t * 4 + 1],
          g.pf_floatfactor[(size_t)t * 4 + 2], g.pf_floatfactor[(size_t)t * 4 + 3]};
    Zw c = zw_mul(zw_unit(g.pf_phase_idx[t] & 7), ff);
    Fact f = zw_factor(c);
    if (f.zero) { tb.zero = true; return; }
    tb.S += f.s; tb.V += f.v; tb.N += f.n; tb.A += f.a; tb.B += f.b;
    if (!f.pure) { tb.residual = true; tb.R = f.R; }
    tb.E = g.pf_power2[t];
  }
}

static void put_mask(std::vector<uint32_t> &w, const Mask &m) {
  w.insert(w.end(), m.w.begin(), m.w.end());
}

void pack_graph(const ksim_graph_view &g, std::vector<uint32_t> &out, uint32_t *ghdr,
                PackStats &st) {
  if (g.G < 0 || g.P < 0) throw std::invalid_argument("negative graph dimensions");
  const bool approx = g.pf_has_approx != 0;
  if (approx && g.G > 0 && !g.pf_approx_f32)
    throw std::invalid_argument("pf_has_approx without pf_approx factors");
  const int W32 = std::max(1, (g.P + 31) / 32);
  {
    // algorithmic work of the reference formulation (zx_eval.hpp:102-178):
    // per term, (2W+1) ops per mask with W = ceil(P/64) 64-bit words,
    // M = n_np + T_hp + 2 T_pp + 2 n_ph masks, 4 ops per node-phase
    // factor, 12 per phase pair and 48 for the fixed factors and the sum
    const double W = (double)((g.P + 63) / 64);
    for (int t = 0; t < g.G; ++t) {
      const double nnp = g.np_counts ? g.np_counts[t] : 0, nph = g.ph_counts ? g.ph_counts[t] : 0;
      const double M = nnp + g.T_hp + 2.0 * g.T_pp + 2.0 * nph;
      st.ops += (2 * W + 1) * M + 4 * nnp + 12 * nph + 48;
    }
  }
  std::vector<TermBuild> terms;
  std::vector<uint32_t> apx;  // approximate factors of the kept terms, (re, im) f32 bits
  terms.reserve(g.G);
  for (int t = 0; t < g.G; ++t) {
    TermBuild tb;
    build_term(g, t, W32, tb);
    if (!tb.zero) {
      terms.push_back(std::move(tb));
      if (approx) {
        uint32_t re, im;
        std::memcpy(&re, &g.pf_approx_f32[2 * (size_t)t], 4);
        std::memcpy(&im, &g.pf_approx_f32[2 * (size_t)t + 1], 4);
        apx.push_back(re);
        apx.push_back(im);
      }
    }
  }
  // ---- exponent ranges and the static alignment exponent
  struct Range { int vmin, vmax, nmin, nmax, amax, bmax, emin, emax; };
  std::vector<Range> rg(terms.size());
  int e0 = 0, emax_all = 0;
  int xbits = INT32_MIN;  // max over terms of bits + emax
  auto lo4 = [](int a, int b, int c, int d) { return std::min(std::min(a, b), std::min(c, d)); };
  auto hi4 = [](int a, int b, int c, int d) { return std::max(std::max(a, b), std::max(c, d)); };
  for (size_t k = 0; k < terms.size(); ++k) {
    const TermBuild &tb = terms[k];
    Range r{tb.V, tb.V, tb.N, tb.N, tb.A, tb.B, 0, 0};
    for (const LinRec &l : tb.lin) {
      r.vmin += std::min(l.e0.v, l.e1.v); r.vmax += std::max(l.e0.v, l.e1.v);
      r.nmin += std::min(l.e0.n, l.e1.n); r.nmax += std::max(l.e0.n, l.e1.n);
      r.amax += std::max(l.e0.a, l.e1.a); r.bmax += std::max(l.e0.b, l.e1.b);
    }
    for (const PairRec &p : tb.pair) {
      r.vmin += lo4(p.e00.v, p.e10.v, p.e01.v, p.e11.v);
      r.vmax += hi4(p.e00.v, p.e10.v, p.e01.v, p.e11.v);
      r.nmin += lo4(p.e00.n, p.e10.n, p.e01.n, p.e11.n);
      r.nmax += hi4(p.e00.n, p.e10.n, p.e01.n, p.e11.n);
      r.amax += hi4(p.e00.a, p.e10.a, p.e01.a, p.e11.a);
      r.bmax += hi4(p.e00.b, p.e10.b, p.e01.b, p.e11.b);
    }
    r.emin = tb.E + (r.vmin >> 2);
    r.emax = tb.E + (r.vmax >> 2);
    rg[k] = r;
    if (k == 0 || r.emin < e0) e0 = r.emin;
    if (k == 0 || r.emax > emax_all) emax_all = r.emax;
    // |coefficients| of w^s (1+w)^r (1+sqrt2)^n' with n' = n + 2 floor(v/4):
    // below 2^30 inside the table range, about 2.414^|n'| * 8 beyond it
    const int n2max = std::max(std::abs(r.nmin + 2 * (r.vmin >> 2)),
                               std::abs(r.nmax + 2 * (r.vmax >> 2)));
    if (n2max > TAB_W)
      throw std::invalid_argument("(1+sqrt2) exponent range beyond +-" + std::to_string(TAB_W));
    // a row beyond TAB_R adds as several digits whose partial sums stay
    // within twice the entry
    int bits = n2max <= TAB_R ? 30 : table_bits(n2max) + 1;
    st.n2_bound = std::max(st.n2_bound, n2max);
    st.wide_terms += n2max > TAB_R;
    if (r.amax > 0 || r.bmax > 0) bits += pi_table_bits(r.amax, r.bmax) + 2;
    if (tb.residual) bits += i128_bits(zw_maxabs(tb.R)) + 2;
    // |term * 2^-e0| < 2^(bits + emax - e0): the largest such bound, plus
    // the term count, must stay inside the signed 128-bit sum
    xbits = std::max(xbits, bits + r.emax);
  }
  const int max_shift = terms.empty() ? 0 : emax_all - e0;
  bool any_wide = false;
  for (size_t k = 0; k < terms.size(); ++k) {
    const Range &r = rg[k];
    any_wide |= std::max(std::abs(r.nmin + 2 * (r.vmin >> 2)),
                         std::abs(r.nmax + 2 * (r.vmax >> 2))) > TAB_R;
  }
  uint32_t gflags = 0;
  if (approx) {
    // the terms are summed in float64 one by one: no exact sum to bound,
    // each term's own value fits the 128-bit accumulator (|n'| <= TAB_W)
    gflags |= GF_APPROX;
  } else if (!terms.empty()) {
    const int need = xbits - e0 + ceil_log2((int64_t)terms.size()) + 1;
    st.sum_bits = std::max(st.sum_bits, need + 1);
    if (need > SHIFT_MAX)
      throw std::invalid_argument(
          "term magnitudes exceed the 128-bit exact sum of the sampler (" +
          std::to_string(need + 1) + " bits needed)");
    // every partial sum (and every shifted term and product) is below
    // 2^need in magnitude, so 64-bit two's-complement sums are exact
    if (need <= 62 && !any_wide) gflags |= GF_NARROW;
    if (terms.size() == 1 && !terms[0].residual && rg[0].amax == 0 && rg[0].bmax == 0)
      gflags |= GF_MONO;
  }
  // ---- term order: GH_PARTS parts of about equal work (heaviest term to
  // the lightest part), parts ordered so that parts {0,1} and {2,3} are
  // also about equal (a two-way split).  The exact sum does not depend on
  // the order of its terms.
  std::vector<size_t> order;
  uint32_t part_first[GH_PARTS + 1] = {0, 0, 0, 0, 0};
  if (approx) {
    // the float64 sum follows the reference's term order (zx_eval.hpp:178),
    // and no term split shares it: every term in part 0
    for (size_t k = 0; k < terms.size(); ++k) order.push_back(k);
    for (int p = 1; p <= GH_PARTS; ++p) part_first[p] = (uint32_t)order.size();
  } else {
    std::vector<double> wt(terms.size());
    for (size_t k = 0; k < terms.size(); ++k) {
      const TermBuild &tb = terms[k];
      wt[k] = 48.0 + (double)W32 * (tb.proj.size() + 2 * tb.ppx.size() + tb.lin.size() +
                                    2 * tb.pair.size()) +
              3.0 * (tb.proj.size() + tb.ppx.size() + tb.lin.size() + 2 * tb.pair.size());
    }
    std::vector<size_t> by(terms.size());
    for (size_t k = 0; k < by.size(); ++k) by[k] = k;
    std::stable_sort(by.begin(), by.end(), [&](size_t x, size_t y) { return wt[x] > wt[y]; });
    std::vector<std::vector<size_t>> bins(GH_PARTS);
    double load[GH_PARTS] = {0, 0, 0, 0};
    for (size_t k : by) {
      int b = 0;
      for (int j = 1; j < GH_PARTS; ++j)
        if (load[j] < load[b]) b = j;
      bins[b].push_back(k);
      load[b] += wt[k];
    }
    int bo[GH_PARTS] = {0, 1, 2, 3};
    std::stable_sort(bo, bo + GH_PARTS, [&](int x, int y) { return load[x] > load[y]; });
    const int seq[GH_PARTS] = {bo[0], bo[3], bo[1], bo[2]};  // (max, min), (2nd, 3rd)
    for (int p = 0; p < GH_PARTS; ++p) {
      std::sort(bins[seq[p]].begin(), bins[seq[p]].end());
      part_first[p] = (uint32_t)order.size();
      order.insert(order.end(), bins[seq[p]].begin(), bins[seq[p]].end());
    }
    part_first[GH_PARTS] = (uint32_t)order.size();
  }
  // ---- emit
  ghdr[GH_NTERMS] = (uint32_t)terms.size();
  ghdr[GH_W32] = (uint32_t)W32;
  ghdr[GH_E0] = (uint32_t)e0;
  ghdr[GH_TERM_OFF] = (uint32_t)out.size();
  ghdr[GH_FLAGS] = gflags;
  size_t nrec = 0;
  for (size_t q = 0; q < order.size(); ++q) {
    const size_t k = order[q];
    for (int p = 1; p < GH_PARTS; ++p)
      if (part_first[p] == q) {
        ghdr[GH_PART_T1 + p - 1] = (uint32_t)q;
        ghdr[GH_PART_O1 + p - 1] = (uint32_t)out.size();
      }
    const TermBuild &tb = terms[k];
    const size_t hdr = out.size();
    out.resize(hdr + TH_DWORDS, 0u);
    std::vector<uint32_t> rec;
    for (const ProjRec &p : tb.proj) { put_mask(rec, p.m); rec.push_back(p.expect); }
    for (const PpxRec &p : tb.ppx) {
      put_mask(rec, p.ms); put_mask(rec, p.mf); rec.push_back(p.sc); rec.push_back(p.fc);
    }
    // chunks of at most CHUNK_MAX sign, LIN and PAIR records
    std::vector<const LinRec *> pm, gen;
    std::vector<int> pm_s, pm_n;
    for (const LinRec &l : tb.lin) {
      int sS, sN;
      if (sign_delta(pack_exp(l.e1) - pack_exp(l.e0), sS, sN)) {
        pm.push_back(&l); pm_s.push_back(sS); pm_n.push_back(sN);
      } else {
        gen.push_back(&l);
      }
    }
    size_t mi = 0, li = 0, pi = 0;
    int nchunk = 0;
    const int nlp = (int)(tb.lin.size() + tb.pair.size());
    while (mi < pm.size() || li < gen.size() || pi < tb.pair.size()) {
      int nm = (int)std::min<size_t>(pm.size() - mi, CHUNK_MAX);
      int nl = (int)std::min<size_t>(gen.size() - li, (size_t)(CHUNK_MAX - nm));
      int np = (int)std::min<size_t>(tb.pair.size() - pi, (size_t)(CHUNK_MAX - nm - nl));
      uint32_t base = 0;
      for (int q = 0; q < nm; ++q) base += pack_exp(pm[mi + q]->e0);
      for (int q = 0; q < nl; ++q) base += pack_exp(gen[li + q]->e0);
      for (int q = 0; q < np; ++q) base += pack_exp(tb.pair[pi + q].e00);
      const size_t chdr = rec.size();
      rec.push_back((uint32_t)nl); rec.push_back((uint32_t)np); rec.push_back(0u);
      rec.push_back((uint32_t)nm);
      for (int g0 = 0; g0 < nm; g0 += PM_GROUP) {
        const int n = std::min(nm - g0, PM_GROUP);
        uint32_t sp = 0, sm = 0, np_ = 0, nn = 0;
        for (int j = 0; j < n; ++j) {
          const uint32_t bit = 1u << (PM_GROUP - n + j);
          const int sS = pm_s[mi + g0 + j], sN = pm_n[mi + g0 + j];
          if (sS > 0) sp |= bit;
          if (sS < 0) sm |= bit;
          if (sN > 0) np_ |= bit;
          if (sN < 0) nn |= bit;
        }
        // the popcounts over (p & X+) | (~p & X-) count X- records of parity 0
        base -= ((uint32_t)__builtin_popcount(sm) << 30) + ((uint32_t)__builtin_popcount(nn) << 9);
        rec.push_back(sp); rec.push_back(sm); rec.push_back(np_); rec.push_back(nn);
        for (int j = 0; j < n; ++j) put_mask(rec, pm[mi + g0 + j]->m);
      }
      rec[chdr + 2] = base;
      for (int q = 0; q < nl; ++q) {
        const LinRec &l = *gen[li + q];
        put_mask(rec, l.m);
        rec.push_back(pack_exp(l.e1) - pack_exp(l.e0));
      }
      for (int q = 0; q < np; ++q) {
        const PairRec &p = tb.pair[pi + q];
        put_mask(rec, p.ma); put_mask(rec, p.mb);
        const uint32_t p00 = pack_exp(p.e00), p10 = pack_exp(p.e10), p01 = pack_exp(p.e01),
                       p11 = pack_exp(p.e11);
        rec.push_back(p10 - p00); rec.push_back(p01 - p00); rec.push_back(p11 - p10 - p01 + p00);
      }
      mi += nm; li += nl; pi += np; ++nchunk;
    }
    uint32_t *h = &out[hdr];
    h[TH_NPROJ] = (uint32_t)tb.proj.size();
    h[TH_NPPX] = (uint32_t)tb.ppx.size();
    h[TH_NCHUNK] = (uint32_t)nchunk;
    uint32_t flags = 0;
    if (tb.residual) flags |= TF_RESIDUAL;
    if (rg[k].amax > 0 || rg[k].bmax > 0) flags |= TF_PI;
    h[TH_FLAGS] = flags;
    h[TH_S] = (uint32_t)(tb.S & 7);
    h[TH_V] = (uint32_t)tb.V;
    h[TH_N] = (uint32_t)(tb.N - NBIAS * nlp);
    h[TH_A] = (uint32_t)tb.A;
    h[TH_B] = (uint32_t)tb.B;
    h[TH_E] = (uint32_t)(tb.E - e0);  // relative to the graph's alignment exponent
    if (tb.residual) {
      if (zw_maxabs(tb.R) >= ((i128)1 << 31)) throw std::invalid_argument("floatfactor too large");
      h[TH_R0] = (uint32_t)(int32_t)tb.R.a; h[TH_R1] = (uint32_t)(int32_t)tb.R.b;
      h[TH_R2] = (uint32_t)(int32_t)tb.R.c; h[TH_R3] = (uint32_t)(int32_t)tb.R.d;
    } else {
      h[TH_R0] = 1; h[TH_R1] = h[TH_R2] = h[TH_R3] = 0;
    }
    h[TH_SIZE] = (uint32_t)rec.size();
    out.insert(out.end(), rec.begin(), rec.end());
    nrec += tb.proj.size() + tb.ppx.size() + tb.lin.size() + tb.pair.size();
  }
  for (int p = 1; p < GH_PARTS; ++p)
    if (part_first[p] == order.size()) {  // empty trailing parts start at the end
      ghdr[GH_PART_T1 + p - 1] = (uint32_t)order.size();
      ghdr[GH_PART_O1 + p - 1] = (uint32_t)out.size();
    }
  ghdr[GH_NREC] = (uint32_t)nrec;
  ghdr[GH_APX_OFF] = 0;
  if (approx) {
    ghdr[GH_APX_OFF] = (uint32_t)out.size();
    out.insert(out.end(), apx.begin(), apx.end());
    st.approx_graphs += 1;
  }
  st.terms += terms.size();
  st.records += nrec;
}

// Rewrite a packed graph (offsets relative to its first word, sign groups
// as masks) into parity groups (GF_TABLE; layout.hpp gives the form): each
// chunk's sign records, and in its first chunk also the term's PPX pairs
// and PROJ projectors, become forms of groups of at most 32 with nibble
// parity tables.  The words must then start at a multiple of 4 dwords in
// the program, which makes every group's tables 16-byte aligned.  A graph
// keeps its masks when it is wider than PM_TAB_W32_MAX words, has at most
// one term or approximate factors, or when its groups would hold fewer than
// PM_TAB_MIN_FILL forms on average: a group costs about the same whatever
// its form count, a mask record W32 + 2 ops or more.
constexpr int PM_TAB_MIN_FILL = 8;

namespace {

struct Form {
  const uint32_t *m;   // W32 mask words
  uint32_t kind;       // 0 sign, 1 PPX psi, 2 PPX phi, 3 PROJ
  int s = 0, t = 0;    // sign record steps
  uint32_t c = 0;      // PPX constant parity / PROJ expected parity
};

// one group of at most PM_GROUP forms (form b at bit b; a PPX psi form is
// followed by its phi form), appended to out
void emit_group(const std::vector<Form> &f, uint32_t W32, std::vector<uint32_t> &out) {
  {
    const size_t g0 = 0, n = f.size();
    uint32_t h[PM_TAB_HDR] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t b = 0; b < n; ++b) {
      const Form &x = f[g0 + b];
      const uint32_t bit = 1u << b;
      if (x.kind == 0) {
        if (x.s > 0) h[0] |= bit;
        if (x.s < 0) h[1] |= bit;
        if (x.t > 0) h[2] |= bit;
        if (x.t < 0) h[3] |= bit;
      } else if (x.kind == 1) {
        h[4] |= bit;
        if (x.c & 1u) h[5] |= bit;
      } else if (x.kind == 2) {
        if (x.c & 1u) h[5] |= bit;
      } else {
        h[6] |= bit;
        if (x.c & 1u) h[7] |= bit;
      }
    }
    out.insert(out.end(), h, h + PM_TAB_HDR);
    for (uint32_t j = 0; j < 8 * W32; ++j) {
      uint32_t col[4] = {0, 0, 0, 0};  // the entries of v = 1, 2, 4, 8
      for (size_t b = 0; b < n; ++b) {
        const uint32_t nib = (f[g0 + b].m[j / 8] >> (4 * (j % 8))) & 15u;
        for (int k = 0; k < 4; ++k)
          if (nib >> k & 1u) col[k] |= 1u << b;
      }
      for (uint32_t v = 0; v < 16; ++v) {
        uint32_t e = 0;
        for (int k = 0; k < 4; ++k)
          if (v >> k & 1u) e ^= col[k];
        out.push_back(e);
      }
    }
  }
}

}  // namespace

bool tabulate_graph(std::vector<uint32_t> &words, uint32_t *ghdr, uint32_t w32max) {
  const uint32_t W32 = ghdr[GH_W32], nt = ghdr[GH_NTERMS];
  if (W32 > w32max || nt < 2 || (ghdr[GH_FLAGS] & (GF_APPROX | GF_MONO)))
    return false;
  const std::vector<uint32_t> &w = words;
  std::vector<uint32_t> out;
  out.reserve(words.size() * 3);
  out.insert(out.end(), w.begin(), w.begin() + ghdr[GH_TERM_OFF]);
  std::vector<std::pair<uint32_t, uint32_t>> moved;  // old -> new term offsets
  size_t forms = 0, groups = 0;
  size_t to = ghdr[GH_TERM_OFF];
  for (uint32_t t = 0; t < nt; ++t) {
    const size_t nto = out.size();
    moved.emplace_back((uint32_t)to, (uint32_t)nto);
    const uint32_t *th = &w[to];
    const uint32_t nproj = th[TH_NPROJ], nppx = th[TH_NPPX], nchunk = th[TH_NCHUNK];
    // the term's PPX pairs (two adjacent forms) and projectors fill the free
    // slots of the chunks' last groups; those left over stay mask records
    // (a group of their own measured slower: more table bytes per form)
    std::vector<std::vector<Form>> pool;
    size_t rp = to + TH_DWORDS;
    std::vector<std::vector<Form>> proj;
    for (uint32_t k = 0; k < nproj; ++k, rp += W32 + 1) {
      Form f{&w[rp], 3u};
      f.c = w[rp + W32];
      proj.push_back({f});
    }
    for (uint32_t k = 0; k < nppx; ++k, rp += 2 * W32 + 2) {
      Form a{&w[rp], 1u}, b{&w[rp + W32], 2u};
      a.c = w[rp + 2 * W32];
      b.c = w[rp + 2 * W32 + 1];
      pool.push_back({a, b});
    }
    pool.insert(pool.end(), proj.begin(), proj.end());
    std::vector<char> used(pool.size(), 0);
    // the records of the pool items placed in groups are known only after
    // the chunks: emit the chunks to a side buffer first
    std::vector<uint32_t> body;
    for (uint32_t c = 0; c < nchunk; ++c) {
      const uint32_t nl = w[rp], np = w[rp + 1], base = w[rp + 2], nm = w[rp + 3];
      rp += CHUNK_HDR;
      std::vector<std::vector<Form>> gs;
      for (uint32_t g0 = 0; g0 < nm; g0 += PM_GROUP) {
        const uint32_t n = std::min<uint32_t>(nm - g0, PM_GROUP);
        const uint32_t sp = w[rp], sm = w[rp + 1], tp = w[rp + 2], tm = w[rp + 3];
        rp += PM_HDR;
        gs.emplace_back();
        for (uint32_t j = 0; j < n; ++j, rp += W32) {
          const uint32_t bit = 1u << (PM_GROUP - n + j);
          Form x{&w[rp], 0u};
          x.s = (sp & bit) ? 1 : (sm & bit) ? -1 : 0;
          x.t = (tp & bit) ? 1 : (tm & bit) ? -1 : 0;
          gs.back().push_back(x);
        }
      }
      if (!gs.empty())
        for (size_t k = 0; k < pool.size(); ++k)
          if (!used[k] && gs.back().size() + pool[k].size() <= (size_t)PM_GROUP) {
            gs.back().insert(gs.back().end(), pool[k].begin(), pool[k].end());
            used[k] = 1;
          }
      const uint32_t ng = (uint32_t)gs.size();
      body.push_back(nl); body.push_back(np); body.push_back(base);
      body.push_back(ng);  // the padding is set below, once the offset is known
      for (const auto &g : gs) {
        emit_group(g, W32, body);
        forms += g.size();
      }
      groups += ng;
      const size_t rest = (size_t)nl * (W32 + 1) + (size_t)np * (2 * W32 + 3);
      body.insert(body.end(), w.begin() + rp, w.begin() + rp + rest);
      rp += rest;
    }
    // the term: its header, the PROJ and PPX records left over, the chunks
    // (each with the zero dwords that put its first group's tables, after
    // the PM_TAB_HDR-dword header, at a multiple of 4 dwords)
    out.insert(out.end(), th, th + TH_DWORDS);
    uint32_t lp = 0, lx = 0;
    for (size_t k = 0; k < pool.size(); ++k)
      if (!used[k] && pool[k][0].kind == 3u) {
        out.insert(out.end(), pool[k][0].m, pool[k][0].m + W32);
        out.push_back(pool[k][0].c);
        ++lp;
      }
    for (size_t k = 0; k < pool.size(); ++k)
      if (!used[k] && pool[k][0].kind == 1u) {
        out.insert(out.end(), pool[k][0].m, pool[k][0].m + W32);
        out.insert(out.end(), pool[k][1].m, pool[k][1].m + W32);
        out.push_back(pool[k][0].c);
        out.push_back(pool[k][1].c);
        ++lx;
      }
    out[nto + TH_NPROJ] = lp;
    out[nto + TH_NPPX] = lx;
    for (size_t q = 0; q < body.size();) {
      const uint32_t nl = body[q], np = body[q + 1], ng = body[q + 3];
      const uint32_t pad =
          ng ? (uint32_t)((4 - (out.size() + CHUNK_HDR + PM_TAB_HDR) % 4) % 4) : 0u;
      out.insert(out.end(), body.begin() + q, body.begin() + q + 3);
      out.push_back(ng | (pad << PM_PAD_SHIFT));
      out.insert(out.end(), pad, 0u);
      const size_t len = (size_t)ng * pm_tab_group_dwords(W32) + (size_t)nl * (W32 + 1) +
                         (size_t)np * (2 * W32 + 3);
      out.insert(out.end(), body.begin() + q + CHUNK_HDR, body.begin() + q + CHUNK_HDR + len);
      q += CHUNK_HDR + len;
    }
    to += TH_DWORDS + th[TH_SIZE];
    out[nto + TH_SIZE] = (uint32_t)(out.size() - nto - TH_DWORDS);
  }
  if (groups == 0 || forms < (size_t)PM_TAB_MIN_FILL * groups) return false;
  moved.emplace_back((uint32_t)to, (uint32_t)out.size());  // the end (empty trailing parts)
  uint32_t po[GH_PARTS - 1];
  for (int p = 0; p < GH_PARTS - 1; ++p) {
    const uint32_t o = ghdr[GH_PART_O1 + p];
    bool found = false;
    for (const auto &mv : moved)
      if (mv.first == o) {
        po[p] = mv.second;
        found = true;
        break;
      }
    if (!found) throw std::logic_error("term part offset is not a term boundary");
  }
  for (int p = 0; p < GH_PARTS - 1; ++p) ghdr[GH_PART_O1 + p] = po[p];
  out.insert(out.end(), w.begin() + to, w.end());  // nothing follows the terms
  words.swap(out);
  ghdr[GH_FLAGS] |= GF_TABLE;
  return true;
}

}  // namespace ksim

namespace ksim {

std::vector<uint32_t> pack_program(const ksim_graph_view *graphs, const int32_t *graphs_per_comp,
                                   const int32_t *f_selection, const int32_t *F_per_comp,
                                   int32_t n_comp, int32_t num_f, const int32_t *dest_col,
                                   int32_t num_outputs, PackStats &st) {
  if (n_comp < 0 || num_f < 0 || num_outputs < 0) throw std::invalid_argument("negative size");
  // validate the component structure first (cheap, sequential)
  size_t total_graphs = 0;
  {
    size_t gi = 0, fi = 0;
    for (int c = 0; c < n_comp; ++c) {
      const int ng = graphs_per_comp[c];
      const int F = F_per_comp[c];
      if (ng < 1) throw std::invalid_argument("a component needs at least one graph");
      if (F < 0) throw std::invalid_argument("negative F");
      for (int j = 0; j < F; ++j)
        if (f_selection[fi + j] < 0 || f_selection[fi + j] >= num_f)
          throw std::invalid_argument("f_selection index out of range");
      // parameter count of graph i must be F + i (zx_eval.hpp:65)
      for (int i = 0; i < ng; ++i) {
        const ksim_graph_view &g = graphs[gi + i];
        if (g.G > 0 && g.P != F + i)
          throw std::invalid_argument("graph " + std::to_string(i) + " of component " +
                                      std::to_string(c) + " has P=" + std::to_string(g.P) +
                                      ", expected F+i=" + std::to_string(F + i));
      }
      if (std::max(1, (F + ng - 1 + 31) / 32) > LANE_WORDS_MAX)
        throw std::invalid_argument("more than " + std::to_string(32 * LANE_WORDS_MAX) +
                                    " parameters (F + outputs) in one component");
      gi += ng;
      fi += F;
    }
    total_graphs = gi;
  }
  // pack every graph independently, in parallel
  struct Job {
    std::vector<uint32_t> words;
    uint32_t hdr[GH_DWORDS] = {0};
    PackStats st;
    std::exception_ptr err;
  };
  std::vector<Job> jobs(total_graphs);
  {
    std::atomic<size_t> next(0);
    auto work = [&] {
      for (;;) {
        const size_t i = next.fetch_add(1);
        if (i >= total_graphs) break;
        try {
          pack_graph(graphs[i], jobs[i].words, jobs[i].hdr, jobs[i].st);
        } catch (...) {
          jobs[i].err = std::current_exception();
        }
      }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<size_t>(hw, total_graphs);
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    for (auto &j : jobs)
      if (j.err) std::rethrow_exception(j.err);
  }
  // parity tables replace the sign masks in the programs that run the
  // register-word kernels (no approximate factors): graphs of up to
  // PM_TAB_W32_MAX mask words where some graph needs 128-bit sums (at most
  // REG_WORDS_MAX lane words), of up to PM_TAB_W32_MAX_NARROW where none
  // does (at most PM_TAB_NARROW_WL lane words)
  {
    bool all_narrow = true, approx = false;
    int wl = 1;
    for (int c = 0; c < n_comp; ++c)
      wl = std::max(wl, std::max(1, (F_per_comp[c] + graphs_per_comp[c] - 1 + 31) / 32));
    for (const Job &j : jobs) {
      if (!(j.hdr[GH_FLAGS] & (GF_MONO | GF_NARROW))) all_narrow = false;
      if (j.hdr[GH_FLAGS] & GF_APPROX) approx = true;
    }
    const uint32_t w32max = approx ? 0u
                            : !all_narrow ? (wl <= REG_WORDS_MAX ? (uint32_t)PM_TAB_W32_MAX : 0u)
                            : (wl <= PM_TAB_NARROW_WL ? (uint32_t)PM_TAB_W32_MAX_NARROW : 0u);
    if (w32max) {
      std::atomic<size_t> next(0);
      auto work = [&] {
        for (;;) {
          const size_t i = next.fetch_add(1);
          if (i >= total_graphs) break;
          tabulate_graph(jobs[i].words, jobs[i].hdr, w32max);
        }
      };
      const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      const unsigned nt = (unsigned)std::min<size_t>(hw, total_graphs);
      std::vector<std::thread> pool;
      for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
      work();
      for (auto &th : pool) th.join();
    }
  }
  // lay out: program header, then per component its header, f_selection,
  // output columns, graph headers and the graphs' term records
  std::vector<uint32_t> w(PH_DWORDS + (size_t)n_comp, 0u);
  w[PH_NCOMP] = (uint32_t)n_comp;
  w[PH_NUMF] = (uint32_t)num_f;
  w[PH_NUMOUT] = (uint32_t)num_outputs;
  size_t gi = 0, fi = 0, oi = 0;
  int wmax = 1;
  for (int c = 0; c < n_comp; ++c) {
    const int ng = graphs_per_comp[c];
    const int F = F_per_comp[c];
    const int n_out = ng - 1;
    const int wc = std::max(1, (F + n_out + 31) / 32);
    wmax = std::max(wmax, wc);
    const size_t ch = w.size();
    w[PH_DWORDS + c] = (uint32_t)ch;
    w.resize(ch + CH_DWORDS, 0u);
    w[ch + CH_NGRAPHS] = (uint32_t)ng;
    w[ch + CH_F] = (uint32_t)F;
    w[ch + CH_NOUT] = (uint32_t)n_out;
    w[ch + CH_WMAX] = (uint32_t)wc;
    w[ch + CH_COLBASE] = (uint32_t)oi;
    w[ch + CH_FSEL_OFF] = (uint32_t)w.size();
    bool contig = true;
    for (int j = 0; j < F; ++j) {
      w.push_back((uint32_t)f_selection[fi + j]);
      if (f_selection[fi + j] != f_selection[fi] + j) contig = false;
    }
    w[ch + CH_FSEL0] = (F > 0 && contig) ? (uint32_t)f_selection[fi] : 0xffffffffu;
    w[ch + CH_DEST_OFF] = (uint32_t)w.size();
    bool dcontig = true;
    for (int j = 0; j < n_out; ++j) {
      const int d = dest_col[oi + j];
      if (d < 0 || d >= num_outputs) throw std::invalid_argument("dest_col out of range");
      if (d != dest_col[oi] + j) dcontig = false;
      w.push_back((uint32_t)d);
    }
    w[ch + CH_DEST0] = (n_out > 0 && dcontig) ? (uint32_t)dest_col[oi] : 0xffffffffu;
    const size_t gh = w.size();
    w[ch + CH_GRAPH_OFF] = (uint32_t)gh;
    w.resize(gh + (size_t)ng * GH_DWORDS, 0u);
    for (int i = 0; i < ng; ++i) {
      Job &j = jobs[gi + i];
      // graph i has P = F + i parameters, so its width is ceil((F + i) / 32);
      // a graph without terms (whose P is not checked) takes that width too,
      // so that every graph's last parameter, the trying bit of the
      // sampler, lies in its last mask word
      if (j.hdr[GH_NTERMS] == 0) j.hdr[GH_W32] = (uint32_t)std::max(1, (F + i + 31) / 32);
      if (j.hdr[GH_FLAGS] & GF_TABLE) w.resize((w.size() + 3) / 4 * 4, 0u);  // 16-byte tables
      j.hdr[GH_TERM_OFF] += (uint32_t)w.size();  // relative -> absolute
      for (int p = 0; p < GH_PARTS - 1; ++p) j.hdr[GH_PART_O1 + p] += (uint32_t)w.size();
      if (j.hdr[GH_FLAGS] & GF_APPROX) j.hdr[GH_APX_OFF] += (uint32_t)w.size();
      std::memcpy(&w[gh + (size_t)i * GH_DWORDS], j.hdr, sizeof(j.hdr));
      if (w.size() + j.words.size() >= (size_t)1 << 32)
        throw std::invalid_argument("program larger than 16 GiB");
      w.insert(w.end(), j.words.begin(), j.words.end());
      std::vector<uint32_t>().swap(j.words);
      st.graphs += 1;
      st.terms += j.st.terms;
      st.records += j.st.records;
      st.n2_bound = std::max(st.n2_bound, j.st.n2_bound);
      st.wide_terms += j.st.wide_terms;
      st.sum_bits = std::max(st.sum_bits, j.st.sum_bits);
      st.ops += j.st.ops;
      st.approx_graphs += j.st.approx_graphs;
    }
    // exact magnitude, compare and running-probability update per output
    st.ops += 32.0 * std::max(0, ng - 1);
    gi += ng;
    fi += F;
    oi += n_out;
  }
  w[PH_TOTOUT] = (uint32_t)oi;
  w[PH_WMAX] = (uint32_t)wmax;
  st.wmax = wmax;
  // the 128-bit kernels read the records of multi-term graphs ahead of
  // their scalar loads (measured 1.1-1.5x faster on programs of 1.8-225 MB,
  // no slower on small ones), touching up to STREAM_AHEAD dwords past the
  // last record
  for (int c = 0; c < n_comp; ++c) {
    const uint32_t ng = w[w[PH_DWORDS + c] + CH_NGRAPHS];
    const uint32_t g0 = w[w[PH_DWORDS + c] + CH_GRAPH_OFF];
    for (uint32_t i = 0; i < ng; ++i) {
      uint32_t *gh = &w[g0 + (size_t)i * GH_DWORDS];
      if (gh[GH_NTERMS] > 1) gh[GH_FLAGS] |= GF_STREAM;
    }
  }
  w.resize(w.size() + STREAM_PAD, 0u);
  return w;
}

}  // namespace ksim
