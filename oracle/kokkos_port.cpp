// kokkos_port.cpp — CPU restatement of the reference's native sampler.
//
// TEST INFRASTRUCTURE ONLY: the "Kokkos-OpenMP" CPU baseline that bench.py
// times beside the GPU, and a float64-path checker.  Never linked into the
// product.
//
// Restates, as plain C++ with OpenMP (one thread per shot), the algorithm of
//   src/cpp/ksim/zx_eval.hpp:24-47   Z[w] multiply, one halving step, w^k
//   src/cpp/ksim/zx_eval.hpp:75-181  per-graph evaluation: u8 masks walked
//                                    byte by byte, a Z[w] multiply per factor,
//                                    float64 cast and sum over terms
//   src/cpp/ksim/kokkos_sim.cpp:43-46,214-230  per-shot xorshift64 stream and
//                                    the autoregressive loop
// The reference itself needs Kokkos, nanobind and MPI (CMakeLists.txt:8-11),
// none of which exist here, so it cannot be built; this port keeps its data
// layout (u8 per mask bit) and arithmetic so its timing stands in for it.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../include/ksim_hip.h"

namespace {

struct Z4 {
  long long a, b, c, d;
};

inline Z4 mul(const Z4 &x, const Z4 &y) {
  return Z4{x.a * y.a + x.b * y.d - x.c * y.c + x.d * y.b,
            x.a * y.b + x.b * y.a + x.c * y.d + x.d * y.c,
            x.a * y.c + x.b * y.b + x.c * y.a - x.d * y.d,
            x.a * y.d - x.b * y.c - x.c * y.b + x.d * y.a};
}

inline void halve_once(Z4 &x, long long &p) {
  if ((x.a | x.b | x.c | x.d) != 0 && !(x.a & 1) && !(x.b & 1) && !(x.c & 1) && !(x.d & 1)) {
    x.a >>= 1; x.b >>= 1; x.c >>= 1; x.d >>= 1;
    ++p;
  }
}

inline Z4 unit(int k) {
  static const long long t[8][4] = {{1, 0, 0, 0},  {0, 1, 0, 0},  {0, 0, 1, 0},  {0, 0, 0, -1},
                                    {-1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, 1}};
  k &= 7;
  return Z4{t[k][0], t[k][1], t[k][2], t[k][3]};
}

// one graph, copied into owned buffers
struct Graph {
  int G = 0, P = 0, Tnp = 0, Thp = 0, Tpp = 0, Tph = 0;
  std::vector<uint8_t> np_ph, np_pa, hp_c, hp_pa, ps_c, ps_pa, pf_c, pf_pa, al, al_pa, be, be_pa,
      pf_idx;
  std::vector<int32_t> np_cnt, ph_cnt, ff, p2;
  std::vector<double> apx;  // (re, im) per term, from the complex64 factors
  bool has_apx = false;
};

template <class T>
std::vector<T> copy(const T *p, size_t n) {
  return n && p ? std::vector<T>(p, p + n) : std::vector<T>(std::max<size_t>(n, 1), T(0));
}

struct Comp {
  int F = 0, n_out = 0;
  std::vector<Graph> g;
};

// parameter bit j of one shot: f, then the sampled prefix, then "trying"
struct Params {
  const uint8_t *f;
  const uint8_t *m;
  int F, nm, trying;
  inline int bit(int j) const {
    if (j < F) return f[j] & 1;
    if (j < F + nm) return m[j - F] & 1;
    return trying;
  }
};

inline int parity(const uint8_t *row, int P, const Params &pr) {
  int acc = 0;
  for (int j = 0; j < P; ++j) acc ^= (row[j] & 1) & pr.bit(j);
  return acc;
}

void eval_graph(const Graph &g, const Params &pr, double &ore, double &oim) {
  const double SQ = 0.70710678118654752440;
  double sre = 0, sim = 0;
  const int P = g.P;
  for (int t = 0; t < g.G; ++t) {
    Z4 acc{1, 0, 0, 0};
    long long pw = 0;
    for (int j = 0; j < g.np_cnt[t]; ++j) {
      const size_t o = (size_t)t * g.Tnp + j;
      const int k = (4 * parity(&g.np_pa[o * P], P, pr) + g.np_ph[o]) & 7;
      Z4 u = unit(k);
      u.a += 1;
      acc = mul(acc, u);
      halve_once(acc, pw);
    }
    int tot = 0;
    for (int j = 0; j < g.Thp; ++j) {
      const size_t o = (size_t)t * g.Thp + j;
      tot += (parity(&g.hp_pa[o * P], P, pr) * g.hp_c[o]) & 7;
    }
    acc = mul(acc, unit(tot & 7));
    halve_once(acc, pw);
    int e = 0;
    for (int j = 0; j < g.Tpp; ++j) {
      const size_t o = (size_t)t * g.Tpp + j;
      const int psi = (g.ps_c[o] + parity(&g.ps_pa[o * P], P, pr)) & 1;
      const int phi = (g.pf_c[o] + parity(&g.pf_pa[o * P], P, pr)) & 1;
      e ^= psi & phi;
    }
    if (e) { acc.a = -acc.a; acc.b = -acc.b; acc.c = -acc.c; acc.d = -acc.d; }
    for (int j = 0; j < g.ph_cnt[t]; ++j) {
      const size_t o = (size_t)t * g.Tph + j;
      const int a = (g.al[o] + 4 * parity(&g.al_pa[o * P], P, pr)) & 7;
      const int b = (g.be[o] + 4 * parity(&g.be_pa[o * P], P, pr)) & 7;
      const Z4 u = unit(a), v = unit(b), w = unit(a + b);
      acc = mul(acc, Z4{1 + u.a + v.a - w.a, u.b + v.b - w.b, u.c + v.c - w.c, u.d + v.d - w.d});
      halve_once(acc, pw);
    }
    acc = mul(acc, unit(g.pf_idx[t]));
    acc = mul(acc, Z4{g.ff[4 * t], g.ff[4 * t + 1], g.ff[4 * t + 2], g.ff[4 * t + 3]});
    halve_once(acc, pw);
    double re = (double)acc.a + ((double)acc.b + (double)acc.d) * SQ;
    double im = (double)acc.c + ((double)acc.b - (double)acc.d) * SQ;
    const double sc = std::exp2((double)(pw + g.p2[t]));
    re *= sc;
    im *= sc;
    if (g.has_apx) {  // approximate floatfactor (zx_eval.hpp:172-177)
      const double ar = g.apx[2 * t], ai = g.apx[2 * t + 1];
      const double r2 = re * ar - im * ai;
      im = re * ai + im * ar;
      re = r2;
    }
    sre += re;
    sim += im;
  }
  ore = sre;
  oim = sim;
}

inline double xorshift_uniform(uint64_t &s) {
  s ^= s << 13;
  s ^= s >> 7;
  s ^= s << 17;
  return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" {

void *kport_create(const ksim_graph_view *gv, int32_t n_graphs, int32_t F) {
  Comp *c = new Comp();
  c->F = F;
  c->n_out = n_graphs - 1;
  for (int i = 0; i < n_graphs; ++i) {
    const ksim_graph_view &v = gv[i];
    Graph g;
    g.G = v.G; g.P = v.P; g.Tnp = v.T_np; g.Thp = v.T_hp; g.Tpp = v.T_pp; g.Tph = v.T_ph;
    const size_t G = v.G, P = v.P;
    g.np_ph = copy(v.np_phases, G * v.T_np);
    g.np_pa = copy(v.np_params, G * v.T_np * P);
    g.np_cnt = copy(v.np_counts, G);
    g.hp_c = copy(v.hp_coeffs, G * v.T_hp);
    g.hp_pa = copy(v.hp_params, G * v.T_hp * P);
    g.ps_c = copy(v.pp_psi_c, G * v.T_pp);
    g.ps_pa = copy(v.pp_psi_p, G * v.T_pp * P);
    g.pf_c = copy(v.pp_phi_c, G * v.T_pp);
    g.pf_pa = copy(v.pp_phi_p, G * v.T_pp * P);
    g.al = copy(v.ph_alpha, G * v.T_ph);
    g.al_pa = copy(v.ph_alpha_p, G * v.T_ph * P);
    g.be = copy(v.ph_beta, G * v.T_ph);
    g.be_pa = copy(v.ph_beta_p, G * v.T_ph * P);
    g.ph_cnt = copy(v.ph_counts, G);
    g.pf_idx = copy(v.pf_phase_idx, G);
    g.ff = copy(v.pf_floatfactor, G * 4);
    g.p2 = copy(v.pf_power2, G);
    g.has_apx = v.pf_has_approx != 0;
    if (g.has_apx)
      for (size_t k = 0; k < 2 * G; ++k) g.apx.push_back((double)v.pf_approx_f32[k]);
    c->g.push_back(std::move(g));
  }
  return c;
}

// f_sel (B, F) u8 -> out (B, max(n_out,1)) u8; uniforms U (B, n_out) or the
// per-shot xorshift64 stream when U is NULL.  Returns 0.
int kport_sample(void *h, const uint8_t *f_sel, int64_t B, uint64_t seed, const double *U,
                 uint8_t *out, int32_t threads) {
  const Comp &c = *(const Comp *)h;
  const int n_out = c.n_out, W = n_out > 0 ? n_out : 1;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < B; ++s) {
    uint64_t rs = seed ^ ((uint64_t)s * 0x9e3779b97f4a7c15ull);
    rs ^= rs >> 30;
    rs *= 0xbf58476d1ce4e5b9ull;
    rs ^= rs >> 27;
    if (!rs) rs = 1;
    uint8_t *m = out + s * W;
    std::memset(m, 0, W);
    Params pr{f_sel + s * c.F, m, c.F, 0, 0};
    double re, im;
    eval_graph(c.g[0], pr, re, im);
    double prev = std::sqrt(re * re + im * im);
    for (int i = 0; i < n_out; ++i) {
      Params pi{f_sel + s * c.F, m, c.F, i, 1};
      eval_graph(c.g[i + 1], pi, re, im);
      const double p1 = std::sqrt(re * re + im * im);
      const double u = U ? U[s * n_out + i] : xorshift_uniform(rs);
      const int bit = u < p1 / prev ? 1 : 0;
      m[i] = (uint8_t)bit;
      prev = bit ? p1 : prev - p1;
    }
  }
  return 0;
}

void kport_destroy(void *h) { delete (Comp *)h; }

int32_t kport_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
