// hpfloat.cpp - truncating multi-limb binary float (see hpfloat.hpp).
#include "hpfloat.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>

namespace nm {

typedef unsigned __int128 u128;

static std::atomic<long> g_default_bits{64};
long hp_default_bits() { return g_default_bits.load(); }
void hp_set_default_bits(long bits) { g_default_bits.store(bits < 1 ? 1 : bits); }

hpf::hpf() : L(hp_limbs_for_bits(hp_default_bits())), sign(0), exp(0) { std::memset(m, 0, sizeof(m)); }
hpf::hpf(double d) : hpf() { hp_set_d(*this, d); }
hpf::hpf(const char* s) : hpf() { hp_set_str(*this, s); }
hpf& hpf::operator=(const hpf& o) {
  if (this != &o) hp_set(*this, o);
  return *this;
}
hpf& hpf::operator=(double d) { hp_set_d(*this, d); return *this; }
hpf& hpf::operator=(const char* s) { hp_set_str(*this, s); return *this; }

// ---------------------------------------------------------------- limb helpers
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// Shift an n-limb magnitude right by sh bits (sh >= 0); returns true if a nonzero
// bit fell off the bottom.
static bool shr_limbs(uint64_t* w, int n, long sh) {
  if (sh <= 0) return false;
  bool sticky = false;
  long ls = sh / 64;
  int bs = (int)(sh % 64);
  if (ls >= n) {
    for (int i = 0; i < n; i++) { sticky |= w[i] != 0; w[i] = 0; }
    return sticky;
  }
  for (long i = 0; i < ls; i++) sticky |= w[i] != 0;
  if (bs) sticky |= (w[ls] << (64 - bs)) != 0;
  for (int i = 0; i < n; i++) {
    long src = i + ls;
    uint64_t lo = src < n ? w[src] : 0;
    uint64_t hi = src + 1 < n ? w[src + 1] : 0;
    w[i] = bs ? (lo >> bs) | (hi << (64 - bs)) : lo;
  }
  return sticky;
}

static void shl_limbs(uint64_t* w, int n, int sh) {  // 0 <= sh < 64*n
  if (sh <= 0) return;
  int ls = sh / 64, bs = sh % 64;
  for (int i = n - 1; i >= 0; i--) {
    long src = i - ls;
    uint64_t hi = src >= 0 ? w[src] : 0;
    uint64_t lo = src - 1 >= 0 ? w[src - 1] : 0;
    w[i] = bs ? (hi << bs) | (lo >> (64 - bs)) : hi;
  }
}

// Normalise an n-limb magnitude w (top-aligned value 0.w * 2^e) and store its top
// r.L limbs into r (truncation).  Returns false if w is zero.
static bool store_trunc(hpf& r, uint64_t* w, int n, int64_t e, int sign) {
  int top = n - 1;
  while (top >= 0 && w[top] == 0) top--;
  if (top < 0) { r.sign = 0; r.exp = 0; std::memset(r.m, 0, sizeof(r.m)); return false; }
  int lead = (n - 1 - top) * 64 + clz64(w[top]);
  if (lead) shl_limbs(w, n, lead);
  e -= lead;
  int L = r.L;
  for (int i = 0; i < L; i++) {
    int src = n - L + i;
    r.m[i] = src >= 0 ? w[src] : 0;
  }
  for (int i = L; i < HP_MAXL; i++) r.m[i] = 0;
  r.exp = e;
  r.sign = sign;
  return true;
}

// ---------------------------------------------------------------- conversions
void hp_set(hpf& r, const hpf& a) {
  if (&r == &a) return;
  int L = r.L;
  if (a.sign == 0) { r.sign = 0; r.exp = 0; std::memset(r.m, 0, sizeof(r.m)); return; }
  for (int i = 0; i < L; i++) {
    int src = a.L - L + i;
    r.m[i] = src >= 0 ? a.m[src] : 0;
  }
  for (int i = L; i < HP_MAXL; i++) r.m[i] = 0;
  r.exp = a.exp;
  r.sign = a.sign;
}

void hp_set_d(hpf& r, double d) {
  std::memset(r.m, 0, sizeof(r.m));
  if (d == 0.0 || !std::isfinite(d)) { r.sign = 0; r.exp = 0; return; }
  int e;
  double f = std::frexp(std::fabs(d), &e);             // f in [0.5, 1)
  uint64_t mant = (uint64_t)std::ldexp(f, 64);         // exact: 53 significant bits
  r.m[r.L - 1] = mant;
  r.exp = e;
  r.sign = d < 0 ? -1 : 1;
}

void hp_neg(hpf& r, const hpf& a) {
  if (&r != &a) hp_set(r, a);
  r.sign = -r.sign;
}

void hpf::set_prec_bits(long bits) {
  hpf t(hp_limbs_for_bits(bits));
  hp_set(t, *this);
  *this = hpf(t);  // copy precision too
  std::memcpy(this, &t, sizeof(hpf));
}

// Truncating conversion to double, including subnormal results.
double hpf::get_d() const {
  if (sign == 0) return 0.0;
  // value = 0.m * 2^exp ; the top 64 bits of m are m[L-1].
  uint64_t top = m[L - 1];
  int64_t e = exp;  // value in [2^(e-1), 2^e)
  if (e > 1024) return sign * HUGE_VAL;
  int keep = 53;
  // the smallest subnormal is 2^-1074: bits of weight < 2^-1074 are dropped.
  // bit k of `top` (from the top, k = 1..64) has weight 2^(e-k).
  if (e - 53 < -1074) keep = (int)(e + 1074);
  if (keep <= 0) return 0.0 * sign;
  uint64_t q = top >> (64 - keep);  // truncation toward zero
  double d = std::ldexp((double)q, (int)(e - keep));
  return sign < 0 ? -d : d;
}

double hpf::get_d_2exp(long* e) const {
  if (sign == 0) { *e = 0; return 0.0; }
  *e = (long)exp;
  uint64_t q = m[L - 1] >> 11;
  double d = std::ldexp((double)q, -53);
  return sign < 0 ? -d : d;
}

// ---------------------------------------------------------------- comparison
static int cmp_mag(const hpf& a, const hpf& b) {
  if (a.exp != b.exp) return a.exp > b.exp ? 1 : -1;
  int n = std::max(a.L, b.L);
  for (int k = 0; k < n; k++) {  // from most significant
    int ia = a.L - 1 - k, ib = b.L - 1 - k;
    uint64_t x = ia >= 0 ? a.m[ia] : 0, y = ib >= 0 ? b.m[ib] : 0;
    if (x != y) return x > y ? 1 : -1;
  }
  return 0;
}

int hp_cmp(const hpf& a, const hpf& b) {
  if (a.sign != b.sign) return a.sign > b.sign ? 1 : -1;
  if (a.sign == 0) return 0;
  int c = cmp_mag(a, b);
  return a.sign > 0 ? c : -c;
}

// ---------------------------------------------------------------- add / sub
static void add_signed(hpf& r, const hpf& a, int sa, const hpf& b, int sb) {
  if (sb == 0 || b.sign == 0) {
    hpf t(r.L); hp_set(t, a); t.sign = a.sign == 0 ? 0 : sa;
    std::memcpy(r.m, t.m, sizeof(t.m)); r.exp = t.exp; r.sign = t.sign; return;
  }
  if (sa == 0 || a.sign == 0) {
    hpf t(r.L); hp_set(t, b); t.sign = sb;
    std::memcpy(r.m, t.m, sizeof(t.m)); r.exp = t.exp; r.sign = t.sign; return;
  }
  const hpf* x = &a; const hpf* y = &b; int sx = sa, sy = sb;
  if (cmp_mag(a, b) < 0) { std::swap(x, y); std::swap(sx, sy); }
  // |x| >= |y|.  Work buffer: max(Lx, Ly, L) + 2 guard limbs + 1 carry limb.
  const int n = std::max(std::max(x->L, y->L), r.L) + 3;
  uint64_t X[HP_MAXL + 3] = {0}, Y[HP_MAXL + 3] = {0};
  // place mantissas below the carry limb (index n-1 stays 0)
  for (int i = 0; i < x->L; i++) X[n - 1 - x->L + i] = x->m[i];
  for (int i = 0; i < y->L; i++) Y[n - 1 - y->L + i] = y->m[i];
  bool sticky = shr_limbs(Y, n, (long)(x->exp - y->exp));
  int64_t e = x->exp + 64;  // X is 0.X * 2^(exp) with one extra top limb
  if (sx == sy) {
    u128 c = 0;
    for (int i = 0; i < n; i++) { c += (u128)X[i] + Y[i]; X[i] = (uint64_t)c; c >>= 64; }
    // exact = X + sticky fraction: truncation ignores the fraction.
  } else {
    u128 br = 0;
    for (int i = 0; i < n; i++) {
      u128 d = (u128)X[i] - Y[i] - br;
      X[i] = (uint64_t)d;
      br = (d >> 127) & 1;
    }
    if (sticky) {  // exact = X - frac  ->  floor is X - 1 (in units of the last limb)
      for (int i = 0; i < n; i++) { if (X[i]--) break; }
    }
  }
  store_trunc(r, X, n, e, sx);
}

void hp_add(hpf& r, const hpf& a, const hpf& b) { add_signed(r, a, a.sign, b, b.sign); }
void hp_sub(hpf& r, const hpf& a, const hpf& b) { add_signed(r, a, a.sign, b, -b.sign); }

// ---------------------------------------------------------------- multiply
void hp_mul(hpf& r, const hpf& a, const hpf& b) {
  if (a.sign == 0 || b.sign == 0) { r.sign = 0; r.exp = 0; std::memset(r.m, 0, sizeof(r.m)); return; }
  const int la = a.L, lb = b.L, n = la + lb;
  uint64_t t[2 * HP_MAXL] = {0};
  for (int i = 0; i < la; i++) {
    u128 c = 0;
    uint64_t ai = a.m[i];
    if (!ai) { t[i + lb] = 0 + t[i + lb]; continue; }
    for (int j = 0; j < lb; j++) {
      c += (u128)ai * b.m[j] + t[i + j];
      t[i + j] = (uint64_t)c;
      c >>= 64;
    }
    t[i + lb] = (uint64_t)c;
  }
  int sign = a.sign * b.sign;
  store_trunc(r, t, n, a.exp + b.exp, sign);
}

void hp_mul_si(hpf& r, const hpf& a, long v) {
  hpf t(1);
  hp_set_d(t, 0.0);
  if (v != 0) {
    unsigned long u = v < 0 ? (unsigned long)(-(v + 1)) + 1ul : (unsigned long)v;
    int lz = clz64(u);
    t.m[0] = u << lz;
    t.exp = 64 - lz;
    t.sign = v < 0 ? -1 : 1;
  }
  hp_mul(r, a, t);
}

void hp_mul_d(hpf& r, const hpf& a, double v) {
  hpf t(1);
  hp_set_d(t, v);
  hp_mul(r, a, t);
}

void hp_div_ui(hpf& r, const hpf& a, unsigned long v) {
  if (v == 0) { r.sign = 0; return; }  // callers never divide by zero
  if (a.sign == 0) { r.sign = 0; r.exp = 0; std::memset(r.m, 0, sizeof(r.m)); return; }
  // numerator 0.m * 2^exp, extended by L+1 zero limbs below: floor(M * 2^(64k) / v)
  const int k = r.L + 1, n = a.L + k;
  uint64_t num[3 * HP_MAXL + 2] = {0}, q[3 * HP_MAXL + 2] = {0};
  for (int i = 0; i < a.L; i++) num[k + i] = a.m[i];
  u128 rem = 0;
  for (int i = n - 1; i >= 0; i--) {
    u128 cur = (rem << 64) | num[i];
    q[i] = (uint64_t)(cur / v);
    rem = cur % v;
  }
  store_trunc(r, q, n, a.exp, a.sign);
}

// ---------------------------------------------------------------- big integers
// Little-endian vectors of 64-bit limbs used for decimal conversion only.
typedef std::vector<uint64_t> big;
static void big_trim(big& x) { while (!x.empty() && x.back() == 0) x.pop_back(); }
static void big_mul_small(big& x, uint64_t v) {
  u128 c = 0;
  for (auto& w : x) { c += (u128)w * v; w = (uint64_t)c; c >>= 64; }
  if (c) x.push_back((uint64_t)c);
}
static void big_add_small(big& x, uint64_t v) {
  u128 c = v;
  for (auto& w : x) { if (!c) break; c += w; w = (uint64_t)c; c >>= 64; }
  if (c) x.push_back((uint64_t)c);
}
static uint64_t big_divmod_small(big& x, uint64_t v) {  // x = floor(x / v)
  u128 rem = 0;
  for (long i = (long)x.size() - 1; i >= 0; i--) {
    u128 cur = (rem << 64) | x[i];
    x[i] = (uint64_t)(cur / v);
    rem = cur % v;
  }
  big_trim(x);
  return (uint64_t)rem;
}
static long big_bitlen(const big& x) {
  if (x.empty()) return 0;
  return (long)(x.size() - 1) * 64 + (64 - clz64(x.back()));
}
static void big_shl(big& x, long sh) {
  if (x.empty() || sh <= 0) return;
  long ls = sh / 64; int bs = (int)(sh % 64);
  big r(x.size() + ls + 1, 0);
  for (size_t i = 0; i < x.size(); i++) {
    r[i + ls] |= bs ? x[i] << bs : x[i];
    if (bs) r[i + ls + 1] |= x[i] >> (64 - bs);
  }
  x.swap(r);
  big_trim(x);
}
static bool big_shr(big& x, long sh) {  // returns sticky
  if (x.empty() || sh <= 0) return false;
  bool sticky = shr_limbs(x.data(), (int)x.size(), sh);
  big_trim(x);
  return sticky;
}
static const uint64_t P10_19 = 10000000000000000000ull;
static uint64_t pow10_small(int k) { uint64_t p = 1; while (k-- > 0) p *= 10; return p; }

// Store the positive integer x * 2^e2 into r with truncation.
static void store_big(hpf& r, big x, int64_t e2, int sign) {
  big_trim(x);
  if (x.empty()) { r.sign = 0; r.exp = 0; std::memset(r.m, 0, sizeof(r.m)); return; }
  long bl = big_bitlen(x);
  // value = x * 2^e2 = 0.x * 2^(bl + e2)
  int n = (int)x.size();
  store_trunc(r, x.data(), n, (int64_t)n * 64 + e2, sign);
  (void)bl;
}

bool hp_set_str(hpf& r, const char* s) {
  const char* p = s;
  while (*p == ' ' || *p == '\t' || *p == '\n') p++;
  int sign = 1;
  if (*p == '-') { sign = -1; p++; } else if (*p == '+') p++;
  big D;
  long frac_digits = 0, ndig = 0;
  bool dot = false, any = false;
  for (; *p; p++) {
    if (*p >= '0' && *p <= '9') {
      big_mul_small(D, 10);
      big_add_small(D, (uint64_t)(*p - '0'));
      if (D.empty() && *p != '0') D.push_back((uint64_t)(*p - '0'));
      if (dot) frac_digits++;
      ndig++; any = true;
    } else if (*p == '.' && !dot) {
      dot = true;
    } else break;
  }
  if (!any) { hp_set_d(r, 0.0); return false; }
  long e10 = 0;
  if (*p == 'e' || *p == 'E' || *p == '@') {
    p++;
    char* endp;
    e10 = std::strtol(p, &endp, 10);
    if (endp == p) { hp_set_d(r, 0.0); return false; }
    p = endp;
  }
  while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
  if (*p) { hp_set_d(r, 0.0); return false; }
  big_trim(D);
  if (D.empty()) { hp_set_d(r, 0.0); return true; }
  e10 -= frac_digits;
  if (e10 >= 0) {
    long k = e10;
    while (k >= 19) { big_mul_small(D, P10_19); k -= 19; }
    if (k) big_mul_small(D, pow10_small((int)k));
    store_big(r, D, 0, sign);
  } else {
    long k = -e10;
    long need = 64L * r.L + 4 + (long)std::ceil(k * 3.3219280948873623) + 1;
    long sh = need - big_bitlen(D);
    if (sh < 0) sh = 0;
    big_shl(D, sh);
    long kk = k;
    while (kk >= 19) { big_divmod_small(D, P10_19); kk -= 19; }
    if (kk) big_divmod_small(D, pow10_small((int)kk));
    store_big(r, D, -sh, sign);
  }
  return true;
}

// Decimal output "[-]0.DDDDe<E>" (mpf_out_str layout), digits rounded away from
// zero so that parsing the string back at the same precision restores the value.
std::string hpf::to_string(int digits) const {
  if (sign == 0) return "0";
  if (digits <= 0) digits = (int)std::ceil(64.0 * L * 0.30102999566398120) + 3;
  // value = M * 2^(exp - 64L), M integer
  big M(m, m + L);
  big_trim(M);
  long e2 = exp - 64L * L;
  // choose k so that v*10^k has about `digits` integer digits
  long log10v = (long)std::floor((double)exp * 0.30102999566398120);
  long k = digits - log10v;
  for (int attempt = 0; attempt < 3; attempt++) {
    big I = M;
    if (k >= 0) {
      long kk = k;
      while (kk >= 19) { big_mul_small(I, P10_19); kk -= 19; }
      if (kk) big_mul_small(I, pow10_small((int)kk));
    }
    bool inexact = false;
    if (e2 >= 0) big_shl(I, e2);
    else inexact = big_shr(I, -e2);
    if (k < 0) {
      long kk = -k;
      while (kk >= 19) { inexact |= big_divmod_small(I, P10_19) != 0; kk -= 19; }
      if (kk) inexact |= big_divmod_small(I, pow10_small((int)kk)) != 0;
    }
    if (inexact) big_add_small(I, 1);  // round away from zero
    // decimal digits of I
    std::string ds;
    big t = I;
    while (!t.empty()) {
      uint64_t chunk = big_divmod_small(t, P10_19);
      char buf[24];
      if (t.empty()) std::snprintf(buf, sizeof buf, "%llu", (unsigned long long)chunk);
      else std::snprintf(buf, sizeof buf, "%019llu", (unsigned long long)chunk);
      ds = std::string(buf) + ds;
    }
    if (ds.empty()) ds = "0";
    long nd = (long)ds.size();
    if (nd < digits - 1) { k += digits - nd; continue; }
    // value = I * 10^-k = 0.ds * 10^(nd - k)
    while (ds.size() > 1 && ds.back() == '0') ds.pop_back();
    std::string out = sign < 0 ? "-0." : "0.";
    out += ds;
    out += "e" + std::to_string(nd - k);
    return out;
  }
  return "0";
}

// ---------------------------------------------------------------- expressions
static inline int maxL(const hpf& a, const hpf& b) { return std::max(a.L, b.L); }
hpf operator+(const hpf& a, const hpf& b) { hpf r(maxL(a, b)); hp_add(r, a, b); return r; }
hpf operator-(const hpf& a, const hpf& b) { hpf r(maxL(a, b)); hp_sub(r, a, b); return r; }
hpf operator*(const hpf& a, const hpf& b) { hpf r(maxL(a, b)); hp_mul(r, a, b); return r; }
hpf operator*(const hpf& a, long v) { hpf r(a.L); hp_mul_si(r, a, v); return r; }
hpf operator*(long v, const hpf& a) { hpf r(a.L); hp_mul_si(r, a, v); return r; }
hpf operator*(const hpf& a, double v) { hpf r(a.L); hp_mul_d(r, a, v); return r; }
hpf operator*(double v, const hpf& a) { hpf r(a.L); hp_mul_d(r, a, v); return r; }
hpf operator+(const hpf& a, double v) { hpf t(a.L); hp_set_d(t, v); hpf r(a.L); hp_add(r, a, t); return r; }
hpf operator-(const hpf& a, double v) { hpf t(a.L); hp_set_d(t, v); hpf r(a.L); hp_sub(r, a, t); return r; }
hpf operator/(const hpf& a, unsigned long v) { hpf r(a.L); hp_div_ui(r, a, v); return r; }
hpf& operator*=(hpf& a, double v) { hp_mul_d(a, a, v); return a; }
bool operator<(const hpf& a, const hpf& b) { return hp_cmp(a, b) < 0; }
bool operator>(const hpf& a, const hpf& b) { return hp_cmp(a, b) > 0; }

}  // namespace nm
