/* exact_fast.c — the exact oracle's graph evaluation in C, for large programs.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that tests compare the HIP sampler
 * against on programs too large for the Python big-integer oracle
 * (oracle/exact.py) to finish in seconds.  Never linked into the product.
 *
 * Restates, per term, the family products of src/cpp/ksim/zx_eval.hpp:102-165
 * (src/python/ksim/sample/evaluate.py:32-61):
 *   NodePhases  prod_{j < np_counts[t]} (1 + w^{(4 par + phase) mod 8})
 *   HalfPi      w^{sum (par * coeff) mod 8}
 *   PiProducts  (-1)^{sum psi * phi}
 *   PhasePairs  prod_{j < ph_counts[t]} (1 + w^a + w^b - w^{a+b})
 *   prefactor   w^{pf_phase_idx} * pf_floatfactor * 2^{pf_power2}
 * with the Z[w] multiply of zx_eval.hpp:27-33 and the halving step of
 * zx_eval.hpp:35-41 (applied until a coefficient is odd, which keeps the
 * value exact), then the exact sum over terms at a common power of two and
 * its canonical form (halve while all four coefficients are even), as
 * oracle/exact.py's graph_value.  Coefficients are NL-limb two's-complement
 * integers (512 bits): sums, differences and products by small integers are
 * exact modulo 2^512, hence exact whenever the true value fits, which every
 * operation checks (a value within 2^40 of the range is refused, never
 * wrapped).  Parities use the masks packed into 64-bit words once per graph.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ksim_hip.h"

#define NL 8

typedef struct { uint64_t l[NL]; } big;
typedef struct { big c[4]; } zw; /* basis (1, w, i, conj w) */

static const int UNIT[8][4] = {{1, 0, 0, 0},  {0, 1, 0, 0},  {0, 0, 1, 0},  {0, 0, 0, -1},
                               {-1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, 1}};

static void big_set(big *x, int64_t v) {
  const uint64_t ext = v < 0 ? ~0ull : 0ull;
  x->l[0] = (uint64_t)v;
  for (int i = 1; i < NL; ++i) x->l[i] = ext;
}

static int big_is_zero(const big *x) {
  for (int i = 0; i < NL; ++i)
    if (x->l[i]) return 0;
  return 1;
}

static int big_neg(const big *x) { return (int)(x->l[NL - 1] >> 63); }

/* |x| < 2^(64 NL - 40): the top limb is its sign extension in its upper 40
 * bits, so a later add or small product cannot reach the sign bit */
static int big_safe(const big *x) {
  const uint64_t top = x->l[NL - 1];
  const uint64_t hi = top >> 24;
  return hi == 0 || hi == (~0ull >> 24);
}

/* r = x * m for an unsigned m, modulo 2^(64 NL) */
static void big_mul_u(big *r, const big *x, uint64_t m) {
  unsigned __int128 c = 0;
  for (int i = 0; i < NL; ++i) {
    c += (unsigned __int128)x->l[i] * m;
    r->l[i] = (uint64_t)c;
    c >>= 64;
  }
}

static void big_add(big *acc, const big *y) {
  unsigned __int128 c = 0;
  for (int i = 0; i < NL; ++i) {
    c += (unsigned __int128)acc->l[i] + y->l[i];
    acc->l[i] = (uint64_t)c;
    c >>= 64;
  }
}

static void big_sub(big *acc, const big *y) {
  uint64_t borrow = 0;
  for (int i = 0; i < NL; ++i) {
    const uint64_t a = acc->l[i], b = y->l[i];
    const uint64_t d = a - b - borrow;
    borrow = (a < b) || (a - b < borrow);
    acc->l[i] = d;
  }
}

/* acc += x * k for a small signed k, modulo 2^(64 NL) */
static void big_addmul(big *acc, const big *x, int64_t k) {
  if (k == 0) return;
  big t;
  big_mul_u(&t, x, (uint64_t)(k < 0 ? -k : k));
  if (k > 0) big_add(acc, &t);
  else big_sub(acc, &t);
}

/* bit length of |x| */
static int big_bitlen(const big *x) {
  big m = *x;
  if (big_neg(x)) {
    big z;
    big_set(&z, 0);
    big_sub(&z, x);
    m = z;
  }
  for (int i = NL - 1; i >= 0; --i)
    if (m.l[i]) return 64 * i + 64 - __builtin_clzll(m.l[i]);
  return 0;
}

/* acc += x << sh (0 <= sh), modulo 2^(64 NL) */
static void big_add_shl(big *acc, const big *x, int sh) {
  big y;
  const int w = sh / 64, b = sh % 64;
  for (int i = 0; i < NL; ++i) {
    const int j = i - w;
    const uint64_t hi = j >= 0 ? x->l[j] << b : 0ull;
    const uint64_t lo = (b && j - 1 >= 0) ? x->l[j - 1] >> (64 - b) : 0ull;
    y.l[i] = hi | lo;
  }
  big_add(acc, &y);
}

static int big_even(const big *x) { return (x->l[0] & 1ull) == 0; }

/* arithmetic shift right by one */
static void big_half(big *x) {
  for (int i = 0; i < NL - 1; ++i) x->l[i] = (x->l[i] >> 1) | (x->l[i + 1] << 63);
  x->l[NL - 1] = (uint64_t)((int64_t)x->l[NL - 1] >> 1);
}

static int zw_zero(const zw *x) {
  for (int k = 0; k < 4; ++k)
    if (!big_is_zero(&x->c[k])) return 0;
  return 1;
}

static int zw_safe(const zw *x) {
  for (int k = 0; k < 4; ++k)
    if (!big_safe(&x->c[k])) return 0;
  return 1;
}

/* x * y for a small y (zx_eval.hpp:27-33) */
static void zw_mul_small(zw *x, const int64_t y[4]) {
  zw r;
  for (int k = 0; k < 4; ++k) big_set(&r.c[k], 0);
  const big *a = &x->c[0], *b = &x->c[1], *c = &x->c[2], *d = &x->c[3];
  /* a' = a ya + b yd - c yc + d yb */
  big_addmul(&r.c[0], a, y[0]); big_addmul(&r.c[0], b, y[3]);
  big_addmul(&r.c[0], c, -y[2]); big_addmul(&r.c[0], d, y[1]);
  /* b' = a yb + b ya + c yd + d yc */
  big_addmul(&r.c[1], a, y[1]); big_addmul(&r.c[1], b, y[0]);
  big_addmul(&r.c[1], c, y[3]); big_addmul(&r.c[1], d, y[2]);
  /* c' = a yc + b yb + c ya - d yd */
  big_addmul(&r.c[2], a, y[2]); big_addmul(&r.c[2], b, y[1]);
  big_addmul(&r.c[2], c, y[0]); big_addmul(&r.c[2], d, -y[3]);
  /* d' = a yd - b yc - c yb + d ya */
  big_addmul(&r.c[3], a, y[3]); big_addmul(&r.c[3], b, -y[2]);
  big_addmul(&r.c[3], c, -y[1]); big_addmul(&r.c[3], d, y[0]);
  *x = r;
}

/* halve while every coefficient is even (exact: value = x 2^p) */
static void zw_reduce(zw *x, int64_t *p) {
  if (zw_zero(x)) return;
  while (big_even(&x->c[0]) && big_even(&x->c[1]) && big_even(&x->c[2]) && big_even(&x->c[3])) {
    for (int k = 0; k < 4; ++k) big_half(&x->c[k]);
    ++*p;
  }
}

/* one graph with masks packed into 64-bit words */
typedef struct {
  int G, P, W, Tnp, Thp, Tpp, Tph;
  uint64_t *np_m, *hp_m, *psi_m, *phi_m, *al_m, *be_m; /* (G, T, W) */
  uint8_t *np_ph, *hp_c, *psi_c, *phi_c, *al, *be, *pf_idx;
  int32_t *np_cnt, *ph_cnt, *ff, *p2;
} graph;

static uint64_t *pack_masks(const uint8_t *m, int G, int T, int P, int W) {
  uint64_t *out = (uint64_t *)calloc((size_t)G * T * W + 1, sizeof(uint64_t));
  for (size_t r = 0; r < (size_t)G * T; ++r)
    for (int j = 0; j < P; ++j)
      if (m[r * P + j] & 1u) out[r * W + j / 64] |= 1ull << (j % 64);
  return out;
}

static void *dup(const void *p, size_t n) {
  void *q = malloc(n ? n : 1);
  if (n) memcpy(q, p, n);
  return q;
}

void *xf_graph_create(const ksim_graph_view *v) {
  graph *g = (graph *)calloc(1, sizeof(graph));
  const int G = v->G, P = v->P;
  g->G = G; g->P = P; g->W = (P + 63) / 64;
  g->Tnp = v->T_np; g->Thp = v->T_hp; g->Tpp = v->T_pp; g->Tph = v->T_ph;
  g->np_m = pack_masks(v->np_params, G, v->T_np, P, g->W);
  g->hp_m = pack_masks(v->hp_params, G, v->T_hp, P, g->W);
  g->psi_m = pack_masks(v->pp_psi_p, G, v->T_pp, P, g->W);
  g->phi_m = pack_masks(v->pp_phi_p, G, v->T_pp, P, g->W);
  g->al_m = pack_masks(v->ph_alpha_p, G, v->T_ph, P, g->W);
  g->be_m = pack_masks(v->ph_beta_p, G, v->T_ph, P, g->W);
  g->np_ph = (uint8_t *)dup(v->np_phases, (size_t)G * v->T_np);
  g->hp_c = (uint8_t *)dup(v->hp_coeffs, (size_t)G * v->T_hp);
  g->psi_c = (uint8_t *)dup(v->pp_psi_c, (size_t)G * v->T_pp);
  g->phi_c = (uint8_t *)dup(v->pp_phi_c, (size_t)G * v->T_pp);
  g->al = (uint8_t *)dup(v->ph_alpha, (size_t)G * v->T_ph);
  g->be = (uint8_t *)dup(v->ph_beta, (size_t)G * v->T_ph);
  g->pf_idx = (uint8_t *)dup(v->pf_phase_idx, (size_t)G);
  g->np_cnt = (int32_t *)dup(v->np_counts, (size_t)G * 4);
  g->ph_cnt = (int32_t *)dup(v->ph_counts, (size_t)G * 4);
  g->ff = (int32_t *)dup(v->pf_floatfactor, (size_t)G * 16);
  g->p2 = (int32_t *)dup(v->pf_power2, (size_t)G * 4);
  return g;
}

void xf_graph_destroy(void *h) {
  graph *g = (graph *)h;
  if (!g) return;
  free(g->np_m); free(g->hp_m); free(g->psi_m); free(g->phi_m); free(g->al_m); free(g->be_m);
  free(g->np_ph); free(g->hp_c); free(g->psi_c); free(g->phi_c); free(g->al); free(g->be);
  free(g->pf_idx); free(g->np_cnt); free(g->ph_cnt); free(g->ff); free(g->p2);
  free(g);
}

/* <mask, params> mod 2 */
static int parity(const uint64_t *m, const uint64_t *pv, int W) {
  uint64_t x = 0;
  for (int w = 0; w < W; ++w) x ^= m[w] & pv[w];
  return __builtin_popcountll(x) & 1;
}

/* term t at packed params pv: 0 ok (value in *out * 2^*e), 1 zero, -1 range */
static int term_value(const graph *g, int t, const uint64_t *pv, zw *out, int64_t *e) {
  const int W = g->W;
  zw acc;
  for (int k = 0; k < 4; ++k) big_set(&acc.c[k], k == 0 ? 1 : 0);
  int64_t pw = 0;
  for (int j = 0; j < g->np_cnt[t]; ++j) {
    const size_t r = (size_t)t * g->Tnp + j;
    const int k = (4 * parity(g->np_m + r * W, pv, W) + g->np_ph[r]) & 7;
    const int64_t f[4] = {UNIT[k][0] + 1, UNIT[k][1], UNIT[k][2], UNIT[k][3]};
    zw_mul_small(&acc, f);
    zw_reduce(&acc, &pw);
    if (!zw_safe(&acc)) return -1;
  }
  int tot = 0;
  for (int j = 0; j < g->Thp; ++j) {
    const size_t r = (size_t)t * g->Thp + j;
    tot += (parity(g->hp_m + r * W, pv, W) * g->hp_c[r]) & 7;
  }
  {
    const int k = tot & 7;
    const int64_t f[4] = {UNIT[k][0], UNIT[k][1], UNIT[k][2], UNIT[k][3]};
    zw_mul_small(&acc, f);
  }
  int sg = 0;
  for (int j = 0; j < g->Tpp; ++j) {
    const size_t r = (size_t)t * g->Tpp + j;
    const int psi = (g->psi_c[r] + parity(g->psi_m + r * W, pv, W)) & 1;
    const int phi = (g->phi_c[r] + parity(g->phi_m + r * W, pv, W)) & 1;
    sg ^= psi & phi;
  }
  if (sg) {
    const int64_t f[4] = {-1, 0, 0, 0};
    zw_mul_small(&acc, f);
  }
  for (int j = 0; j < g->ph_cnt[t]; ++j) {
    const size_t r = (size_t)t * g->Tph + j;
    const int a = (g->al[r] + 4 * parity(g->al_m + r * W, pv, W)) & 7;
    const int b = (g->be[r] + 4 * parity(g->be_m + r * W, pv, W)) & 7;
    const int c = (a + b) & 7;
    const int64_t f[4] = {1 + UNIT[a][0] + UNIT[b][0] - UNIT[c][0],
                          UNIT[a][1] + UNIT[b][1] - UNIT[c][1],
                          UNIT[a][2] + UNIT[b][2] - UNIT[c][2],
                          UNIT[a][3] + UNIT[b][3] - UNIT[c][3]};
    zw_mul_small(&acc, f);
    zw_reduce(&acc, &pw);
    if (!zw_safe(&acc)) return -1;
  }
  {
    const int k = g->pf_idx[t] & 7;
    const int64_t f[4] = {UNIT[k][0], UNIT[k][1], UNIT[k][2], UNIT[k][3]};
    zw_mul_small(&acc, f);
    const int32_t *ff = g->ff + (size_t)t * 4;
    const int64_t q[4] = {ff[0], ff[1], ff[2], ff[3]};
    zw_mul_small(&acc, q);
    if (!zw_safe(&acc)) return -1;
  }
  if (zw_zero(&acc)) return 1;
  zw_reduce(&acc, &pw);
  *out = acc;
  *e = pw + g->p2[t];
  return 0;
}

/* Exact value of the graph at params rows (B, P) u8: canonical coefficients
 * (B, 4, NL) u64 limbs (two's complement, least significant first) and
 * exponents (B).  Returns 0, or -1 when a value leaves the 512-bit range. */
int xf_graph_values(void *h, const uint8_t *params, int64_t B, uint64_t *coeffs, int64_t *exps) {
  const graph *g = (const graph *)h;
  const int W = g->W;
  uint64_t *pv = (uint64_t *)calloc((size_t)W + 1, sizeof(uint64_t));
  zw *tv = (zw *)malloc(sizeof(zw) * (size_t)(g->G ? g->G : 1));
  int64_t *te = (int64_t *)malloc(sizeof(int64_t) * (size_t)(g->G ? g->G : 1));
  int rc = 0;
  for (int64_t s = 0; s < B && rc == 0; ++s) {
    memset(pv, 0, sizeof(uint64_t) * (size_t)W);
    for (int j = 0; j < g->P; ++j)
      if (params[s * g->P + j] & 1u) pv[j / 64] |= 1ull << (j % 64);
    int n = 0;
    int64_t emin = 0;
    for (int t = 0; t < g->G; ++t) {
      const int r = term_value(g, t, pv, &tv[n], &te[n]);
      if (r < 0) { rc = -1; break; }
      if (r == 0) {
        if (n == 0 || te[n] < emin) emin = te[n];
        ++n;
      }
    }
    if (rc) break;
    zw sum;
    for (int k = 0; k < 4; ++k) big_set(&sum.c[k], 0);
    for (int i = 0; i < n && rc == 0; ++i) {
      const int64_t sh = te[i] - emin;
      for (int k = 0; k < 4; ++k) {
        /* the shifted term and the partial sum stay 40 bits inside the range */
        if (sh + big_bitlen(&tv[i].c[k]) > 64 * NL - 41) { rc = -1; break; }
        big_add_shl(&sum.c[k], &tv[i].c[k], (int)sh);
      }
      if (rc == 0 && !zw_safe(&sum)) rc = -1;
    }
    if (rc) break;
    int64_t e = emin;
    if (zw_zero(&sum)) e = 0;
    zw_reduce(&sum, &e);
    for (int k = 0; k < 4; ++k) memcpy(coeffs + ((size_t)s * 4 + k) * NL, sum.c[k].l, 8 * NL);
    exps[s] = e;
  }
  free(pv);
  free(tv);
  free(te);
  return rc;
}

int xf_limbs(void) { return NL; }
