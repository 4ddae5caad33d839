// Operation-counting scalars for the oracle's counting build (test
// infrastructure, oracle/Makefile target `flopcount`).  The oracle's C source
// is compiled as C++ with its `double` and `double complex` rewritten onto F
// and CF, which tally every floating-point operation they execute:
//   add     additions and subtractions
//   mul     multiplications
//   div     divisions
//   sqrt, exp, log, atan2, sincos   one each per call (a complex exponential
//           is one exp, one sincos and two multiplications)
// Negation, fabs and comparisons are not counted.  Complex arithmetic follows
// the plain formulas the oracle is compiled with (-fcx-limited-range):
// (a + bi)(c + di) = 4 mul + 2 add, division 6 mul + 3 add + 2 div, and a
// real operand touches only the parts it meets.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

struct OpCount {
  uint64_t add, mul, div, sqrt, exp, log, atan2, sincos;
};
inline OpCount g_ops = {0, 0, 0, 0, 0, 0, 0, 0};

struct F {
  double v;
  F() = default;
  constexpr F(double x) : v(x) {}
  constexpr F(int x) : v(x) {}
  constexpr F(long x) : v((double)x) {}
  constexpr F(long long x) : v((double)x) {}
  constexpr F(unsigned x) : v(x) {}
  constexpr F(unsigned long x) : v((double)x) {}
  explicit constexpr operator double() const { return v; }
  explicit constexpr operator int() const { return (int)v; }
  explicit constexpr operator long() const { return (long)v; }
  F &operator+=(F b) { g_ops.add++; v += b.v; return *this; }
  F &operator-=(F b) { g_ops.add++; v -= b.v; return *this; }
  F &operator*=(F b) { g_ops.mul++; v *= b.v; return *this; }
  F &operator/=(F b) { g_ops.div++; v /= b.v; return *this; }
};
inline F operator+(F a, F b) { g_ops.add++; return F(a.v + b.v); }
inline F operator-(F a, F b) { g_ops.add++; return F(a.v - b.v); }
inline F operator*(F a, F b) { g_ops.mul++; return F(a.v * b.v); }
inline F operator/(F a, F b) { g_ops.div++; return F(a.v / b.v); }
// exact matches for double and int operands (a double could otherwise reach
// both F and CF by one conversion)
#define BEM_F_MIXED(T)                                       \
  inline F operator+(F a, T b) { return a + F(b); }          \
  inline F operator+(T a, F b) { return F(a) + b; }          \
  inline F operator-(F a, T b) { return a - F(b); }          \
  inline F operator-(T a, F b) { return F(a) - b; }          \
  inline F operator*(F a, T b) { return a * F(b); }          \
  inline F operator*(T a, F b) { return F(a) * b; }          \
  inline F operator/(F a, T b) { return a / F(b); }          \
  inline F operator/(T a, F b) { return F(a) / b; }          \
  inline bool operator<(F a, T b) { return a.v < (double)b; }  \
  inline bool operator>(F a, T b) { return a.v > (double)b; }  \
  inline bool operator<=(F a, T b) { return a.v <= (double)b; } \
  inline bool operator>=(F a, T b) { return a.v >= (double)b; } \
  inline bool operator==(F a, T b) { return a.v == (double)b; } \
  inline bool operator!=(F a, T b) { return a.v != (double)b; } \
  inline bool operator<(T a, F b) { return (double)a < b.v; }  \
  inline bool operator>(T a, F b) { return (double)a > b.v; }
BEM_F_MIXED(double)
BEM_F_MIXED(int)
#undef BEM_F_MIXED
inline F operator-(F a) { return F(-a.v); }
inline F operator+(F a) { return a; }
inline bool operator<(F a, F b) { return a.v < b.v; }
inline bool operator<=(F a, F b) { return a.v <= b.v; }
inline bool operator>(F a, F b) { return a.v > b.v; }
inline bool operator>=(F a, F b) { return a.v >= b.v; }
inline bool operator==(F a, F b) { return a.v == b.v; }
inline bool operator!=(F a, F b) { return a.v != b.v; }
inline F sqrt(F a) { g_ops.sqrt++; return F(std::sqrt(a.v)); }
inline F exp(F a) { g_ops.exp++; return F(std::exp(a.v)); }
inline F log(F a) { g_ops.log++; return F(std::log(a.v)); }
inline F atan2(F a, F b) { g_ops.atan2++; return F(std::atan2(a.v, b.v)); }
inline F fabs(F a) { return F(std::fabs(a.v)); }

struct CF {
  F re, im;
  CF() = default;
  constexpr CF(F r) : re(r), im(0.0) {}
  constexpr CF(F r, F i) : re(r), im(i) {}
  constexpr CF(double r) : re(r), im(0.0) {}
  constexpr CF(int r) : re(r), im(0.0) {}
};
inline CF operator+(CF a, CF b) { return CF(a.re + b.re, a.im + b.im); }
inline CF operator-(CF a, CF b) { return CF(a.re - b.re, a.im - b.im); }
inline CF operator*(CF a, CF b) { return CF(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
inline CF operator/(CF a, CF b) {
  const F d = b.re * b.re + b.im * b.im;
  return CF((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
inline CF operator-(CF a) { return CF(-a.re, -a.im); }
// a real operand meets only the parts it touches
inline CF operator+(CF a, F b) { return CF(a.re + b, a.im); }
inline CF operator+(F a, CF b) { return CF(a + b.re, b.im); }
inline CF operator-(CF a, F b) { return CF(a.re - b, a.im); }
inline CF operator-(F a, CF b) { return CF(a - b.re, -b.im); }
inline CF operator*(CF a, F b) { return CF(a.re * b, a.im * b); }
inline CF operator*(F a, CF b) { return CF(a * b.re, a * b.im); }
inline CF operator/(CF a, F b) { return CF(a.re / b, a.im / b); }
inline CF operator/(F a, CF b) {
  const F d = b.re * b.re + b.im * b.im;
  return CF((a * b.re) / d, (-a * b.im) / d);
}
// exact matches for double and int operands, so that no overload is ambiguous
#define BEM_CF_MIXED(T)                                            \
  inline CF operator+(CF a, T b) { return a + F(b); }              \
  inline CF operator+(T a, CF b) { return F(a) + b; }              \
  inline CF operator-(CF a, T b) { return a - F(b); }              \
  inline CF operator-(T a, CF b) { return F(a) - b; }              \
  inline CF operator*(CF a, T b) { return a * F(b); }              \
  inline CF operator*(T a, CF b) { return F(a) * b; }              \
  inline CF operator/(CF a, T b) { return a / F(b); }              \
  inline CF operator/(T a, CF b) { return F(a) / b; }
BEM_CF_MIXED(double)
BEM_CF_MIXED(int)
#undef BEM_CF_MIXED
inline CF &operator+=(CF &a, CF b) { return a = a + b; }
inline CF &operator-=(CF &a, CF b) { return a = a - b; }
inline CF &operator*=(CF &a, CF b) { return a = a * b; }
inline CF &operator+=(CF &a, F b) { return a = a + b; }
inline CF &operator-=(CF &a, F b) { return a = a - b; }
inline CF &operator*=(CF &a, F b) { return a = a * b; }
inline CF &operator*=(CF &a, double b) { return a = a * F(b); }
inline CF &operator/=(CF &a, F b) { return a = a / b; }

inline F creal(CF a) { return a.re; }
inline F cimag(CF a) { return a.im; }
inline CF conj(CF a) { return CF(a.re, -a.im); }
inline CF cexp(CF a) {
  g_ops.exp++;
  g_ops.sincos++;
  const F e = F(std::exp(a.re.v));
  return CF(e * F(std::cos(a.im.v)), e * F(std::sin(a.im.v)));
}
// principal square root: r = |z|, sqrt((r + |x|) / 2) and its partner
inline CF csqrt(CF a) {
  const F r = sqrt(a.re * a.re + a.im * a.im);
  if (a.re.v >= 0) {
    const F t = sqrt((r + a.re) * F(0.5));
    return CF(t, a.im / (F(2.0) * t));
  }
  const F t = sqrt((r - a.re) * F(0.5));
  return CF(fabs(a.im) / (F(2.0) * t), a.im.v < 0 ? -t : t);
}
#define I CF(0.0, 1.0)

// the oracle's thread controls, without OpenMP
inline void omp_set_num_threads(int) {}
inline int omp_get_max_threads() { return 1; }
