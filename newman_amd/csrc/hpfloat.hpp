// hpfloat.hpp - host high-precision binary float used for the reference orbit,
// the series coefficients, pixel coordinates and the cardioid test.
//
// Semantics follow GMP mpf as the reference uses it (complex.h:4, 14-17, 33-35):
//   * every arithmetic result is truncated toward zero (mpf_add / mpf_mul truncate);
//   * conversion to double truncates toward zero (mpf_get_d, complex.h:33-35);
//   * each object carries its own precision; assignment keeps the destination's
//     precision (mpf_set), copy construction copies the source's (mpf_init2+mpf_set);
//   * new objects take a process-wide default precision (mpf_set_default_prec,
//     mandelbrot.cpp:47).
// Where mpf's effective precision varies with limb alignment (it keeps prec+1 limbs
// with a limb-granular exponent), this type keeps a bit-normalised mantissa of
// L = (bits + 127) / 64 + 1 limbs, i.e. at least as many bits as mpf ever holds,
// and truncates results exactly (the exact result, then toward zero).
//
// value = sign * 0.m * 2^exp, with m = m[L-1] .. m[0] (m[L-1] most significant,
// top bit of m[L-1] set unless the value is zero).
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>

namespace nm {

constexpr int HP_MAXL = 40;  // 2560 mantissa bits: views down to about 1e-700

inline int hp_limbs_for_bits(long bits) {
  if (bits < 1) bits = 1;
  long L = (bits + 127) / 64 + 1;
  if (L > HP_MAXL) L = HP_MAXL;
  return (int)L;
}

// Process-wide default precision in bits (mpf_set_default_prec analogue).
long hp_default_bits();
void hp_set_default_bits(long bits);

struct hpf {
  int32_t L = 0;     // precision, limbs
  int32_t sign = 0;  // -1, 0, +1
  int64_t exp = 0;
  uint64_t m[HP_MAXL];

  hpf();                            // zero at the default precision
  explicit hpf(int L_) : L(L_), sign(0), exp(0) { std::memset(m, 0, sizeof(m)); }
  hpf(const hpf& o) { std::memcpy(this, &o, sizeof(hpf)); }
  hpf(double d);                    // default precision
  hpf(const char* s);               // default precision
  hpf& operator=(const hpf& o);     // keeps this precision (truncates)
  hpf& operator=(double d);
  hpf& operator=(const char* s);
  hpf& operator=(const std::string& s) { return *this = s.c_str(); }

  static hpf with_prec(int L_) { return hpf(L_); }
  long prec_bits() const { return 64L * L; }
  void set_prec_bits(long bits);    // mpf_set_prec: truncates the value
  double get_d() const;             // truncating (mpf_get_d)
  double get_d_2exp(long* e) const; // mpf_get_d_2exp: |d| in [0.5, 1), truncated
  std::string to_string(int digits = 0) const;  // "-0.ddddde-N", rounded away from zero
  bool is_zero() const { return sign == 0; }
};

// r = trunc(a op b) at r.L limbs. r may alias a or b.
void hp_add(hpf& r, const hpf& a, const hpf& b);
void hp_sub(hpf& r, const hpf& a, const hpf& b);
void hp_mul(hpf& r, const hpf& a, const hpf& b);
void hp_mul_si(hpf& r, const hpf& a, long v);
void hp_mul_d(hpf& r, const hpf& a, double v);
void hp_div_ui(hpf& r, const hpf& a, unsigned long v);
void hp_set(hpf& r, const hpf& a);              // truncating copy into r.L
void hp_set_d(hpf& r, double d);
bool hp_set_str(hpf& r, const char* s);         // false on a malformed string
int hp_cmp(const hpf& a, const hpf& b);         // -1, 0, +1 (exact)
void hp_neg(hpf& r, const hpf& a);
// copy value and precision (copy construction semantics into an existing object)
inline void hp_copy_prec(hpf& d, const hpf& s) { std::memcpy((void*)&d, (const void*)&s, sizeof(hpf)); }

// Expression helpers with gmpxx-like temporary precision (max of the operands).
hpf operator+(const hpf& a, const hpf& b);
hpf operator-(const hpf& a, const hpf& b);
hpf operator*(const hpf& a, const hpf& b);
hpf operator*(const hpf& a, long v);
hpf operator*(long v, const hpf& a);
hpf operator*(const hpf& a, double v);
hpf operator*(double v, const hpf& a);
hpf operator+(const hpf& a, double v);
hpf operator-(const hpf& a, double v);
hpf operator/(const hpf& a, unsigned long v);
inline hpf operator*(const hpf& a, int v) { return a * (long)v; }
inline hpf operator*(int v, const hpf& a) { return a * (long)v; }
inline hpf operator/(const hpf& a, int v) {
  hpf r = a / (unsigned long)(v < 0 ? -(long)v : v);
  if (v < 0) hp_neg(r, r);
  return r;
}
hpf& operator*=(hpf& a, double v);
bool operator<(const hpf& a, const hpf& b);
bool operator>(const hpf& a, const hpf& b);

// HPComplex of complex.h:14-17 with our HP type.
struct HPComplex {
  hpf re, im;
};

// LPComplex and its fixed-order operations (complex.h:6-31).
struct LPComplex {
  double re = 0.0, im = 0.0;
  LPComplex() = default;
  LPComplex(double r, double i) : re(r), im(i) {}
};
inline LPComplex sq(const LPComplex& a) { return LPComplex(a.re * a.re - a.im * a.im, 2.0 * a.re * a.im); }
inline double sqMag(const LPComplex& a) { return a.re * a.re + a.im * a.im; }
inline LPComplex operator+(const LPComplex& a, const LPComplex& b) { return LPComplex(a.re + b.re, a.im + b.im); }
inline LPComplex operator*(const LPComplex& a, const LPComplex& b) {
  return LPComplex(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
inline LPComplex descend(const HPComplex& a) { return LPComplex(a.re.get_d(), a.im.get_d()); }

}  // namespace nm
