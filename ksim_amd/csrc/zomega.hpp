// zomega.hpp — exact Z[w] arithmetic on the host (w = e^{i pi/4}).
//
// Basis (1, w, i, conj w) as in src/cpp/ksim/zx_eval.hpp:14,24-47, with
// 128-bit coefficients so the packer's factorisation and the lookup-table
// generation cannot wrap.  Used only on the host: to factor each factor value
// and each prefactor into w^s (1+w)^v (1+sqrt2)^n pi^a pibar^b * residual, and
// to build the materialisation tables the kernels read.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>

namespace ksim {

typedef __int128 i128;

struct Zw {
  i128 a, b, c, d;
};

inline bool zw_is_zero(const Zw &x) { return x.a == 0 && x.b == 0 && x.c == 0 && x.d == 0; }
inline bool zw_eq(const Zw &x, const Zw &y) {
  return x.a == y.a && x.b == y.b && x.c == y.c && x.d == y.d;
}

// product in the basis (1, w, i, conj w); same formula as zx_eval.hpp:27-33
inline Zw zw_mul(const Zw &x, const Zw &y) {
  return Zw{x.a * y.a + x.b * y.d - x.c * y.c + x.d * y.b,
            x.a * y.b + x.b * y.a + x.c * y.d + x.d * y.c,
            x.a * y.c + x.b * y.b + x.c * y.a - x.d * y.d,
            x.a * y.d - x.b * y.c - x.c * y.b + x.d * y.a};
}

inline Zw zw_unit(int k) {
  static const int t[8][4] = {{1, 0, 0, 0},  {0, 1, 0, 0},  {0, 0, 1, 0},  {0, 0, 0, -1},
                              {-1, 0, 0, 0}, {0, -1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, 1}};
  k &= 7;
  return Zw{t[k][0], t[k][1], t[k][2], t[k][3]};
}

inline Zw zw_one() { return Zw{1, 0, 0, 0}; }
inline Zw zw_lambda() { return Zw{1, 1, 0, 0}; }     // 1 + w
inline Zw zw_uplus() { return Zw{1, 1, 0, 1}; }      // 1 + sqrt2   (sqrt2 = w + conj w)
inline Zw zw_uminus() { return Zw{-1, 1, 0, 1}; }    // sqrt2 - 1 = (1 + sqrt2)^-1
inline Zw zw_pi() { return Zw{1, 1, 0, -1}; }        // 1 + i sqrt2  (i sqrt2 = w - conj w)
inline Zw zw_pibar() { return Zw{1, -1, 0, 1}; }     // 1 - i sqrt2

inline Zw zw_pow(Zw x, int n) {
  Zw r = zw_one();
  for (int k = 0; k < n; ++k) r = zw_mul(r, x);
  return r;
}

// (1+sqrt2)^n for any integer n
inline Zw zw_unit_pow(int n) { return n >= 0 ? zw_pow(zw_uplus(), n) : zw_pow(zw_uminus(), -n); }

inline i128 i128_abs(i128 x) { return x < 0 ? -x : x; }

inline i128 zw_maxabs(const Zw &x) {
  i128 m = i128_abs(x.a);
  if (i128_abs(x.b) > m) m = i128_abs(x.b);
  if (i128_abs(x.c) > m) m = i128_abs(x.c);
  if (i128_abs(x.d) > m) m = i128_abs(x.d);
  return m;
}

inline int i128_bits(i128 x) {  // bits of |x| (0 for 0)
  i128 m = i128_abs(x);
  int b = 0;
  while (m > 0) { m >>= 1; ++b; }
  return b;
}

// exact division x / y when y * ycof = den (an integer); returns false if not exact
inline bool zw_div_by(const Zw &x, const Zw &ycof, i128 den, Zw *q) {
  Zw t = zw_mul(x, ycof);
  if (t.a % den || t.b % den || t.c % den || t.d % den) return false;
  *q = Zw{t.a / den, t.b / den, t.c / den, t.d / den};
  return true;
}

// (1+w)(1+w^3)(1+w^5)(1+w^7) = 2, so 1/(1+w) = (1+w^3)(1+w^5)(1+w^7) / 2
inline Zw zw_lambda_cof() {
  Zw l3{1, 0, 0, -1}, l5{1, -1, 0, 0}, l7{1, 0, 0, 1};
  return zw_mul(zw_mul(l3, l5), l7);
}

// complex embedding w -> e^{i pi j/4}, j odd; returns |.| in double
inline double zw_abs_embed(const Zw &x, int j) {
  const double c = std::cos(M_PI * j / 4.0), s = std::sin(M_PI * j / 4.0);
  // w^j = c + i s ; conj(w)^j = c - i s ; i^j = cos(pi j/2) + i sin(pi j/2)
  const double ci = std::cos(M_PI * j / 2.0), si = std::sin(M_PI * j / 2.0);
  double re = (double)x.a + (double)x.b * c + (double)x.c * ci + (double)x.d * c;
  double im = (double)x.b * s + (double)x.c * si - (double)x.d * s;
  return std::sqrt(re * re + im * im);
}

// Factorisation x = w^s (1+w)^v (1+sqrt2)^n pi^a pibar^b * R.
// R = 1 when x factors completely; otherwise R is what is left after the
// (1+w), pi and pibar powers are removed (s = n = 0 then).
struct Fact {
  bool zero = false;
  int s = 0, v = 0, n = 0, a = 0, b = 0;
  bool pure = true;
  Zw R{1, 0, 0, 0};
};

inline Fact zw_factor(Zw x) {
  Fact f;
  if (zw_is_zero(x)) { f.zero = true; return f; }
  const Zw lc = zw_lambda_cof();
  Zw q;
  while (zw_div_by(x, lc, 2, &q)) { x = q; ++f.v; }
  while (zw_div_by(x, zw_pibar(), 3, &q)) { x = q; ++f.a; }   // x / pi = x pibar / 3
  while (zw_div_by(x, zw_pi(), 3, &q)) { x = q; ++f.b; }      // x / pibar = x pi / 3
  // unit?  |sigma_1(x)| = (1+sqrt2)^n for a unit w^s (1+sqrt2)^n
  double m1 = zw_abs_embed(x, 1);
  if (m1 > 0) {
    int n = (int)std::lround(std::log(m1) / std::log(1.0 + std::sqrt(2.0)));
    if (n > -60 && n < 60) {
      Zw y = zw_mul(x, zw_unit_pow(-n));
      for (int s = 0; s < 8; ++s) {
        if (zw_eq(y, zw_unit(s))) {
          f.s = s;
          f.n = n;
          return f;
        }
      }
    }
  }
  f.pure = false;
  f.R = x;
  return f;
}

inline Zw zw_build(int s, int v, int n, int a, int b) {
  Zw x = zw_unit(s);
  x = zw_mul(x, zw_pow(zw_lambda(), v));
  x = zw_mul(x, zw_unit_pow(n));
  x = zw_mul(x, zw_pow(zw_pi(), a));
  x = zw_mul(x, zw_pow(zw_pibar(), b));
  return x;
}

}  // namespace ksim
