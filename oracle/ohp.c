/* ohp.c - the oracle's truncating high-precision float (see ohp.h).
 * TEST INFRASTRUCTURE ONLY.
 */
#include "ohp.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

int ohp_limbs_for_bits(long bits) {
  if (bits < 1) bits = 1;
  long n = 2 * ((bits + 127) / 64 + 1);
  if (n > OHP_MAXN) n = OHP_MAXN;
  return (int)n;
}

void ohp_zero(ohp* r, int n) {
  r->n = n;
  r->sign = 0;
  r->exp = 0;
  memset(r->d, 0, sizeof(r->d));
}

static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* buf[0..nb-1], most significant first, value 0.buf * 2^e.  Normalise and keep
 * the top r->n limbs (truncation toward zero of the magnitude). */
static void store(ohp* r, uint32_t* buf, int nb, long e, int sign) {
  int first = 0;
  while (first < nb && buf[first] == 0) first++;
  if (first == nb) {
    int n = r->n;
    ohp_zero(r, n);
    return;
  }
  int lz = clz32(buf[first]);
  e -= 32L * first + lz;
  int n = r->n;
  for (int i = 0; i < n; i++) {
    int src = first + i;
    uint32_t hi = src < nb ? buf[src] : 0;
    uint32_t lo = src + 1 < nb ? buf[src + 1] : 0;
    r->d[i] = lz ? (hi << lz) | (lo >> (32 - lz)) : hi;
  }
  for (int i = n; i < OHP_MAXN; i++) r->d[i] = 0;
  r->exp = e;
  r->sign = sign;
}

void ohp_copy(ohp* r, const ohp* a) {
  if (r == a) return;
  if (a->sign == 0) {
    ohp_zero(r, r->n);
    return;
  }
  uint32_t buf[OHP_MAXN];
  memcpy(buf, a->d, sizeof(uint32_t) * a->n);
  store(r, buf, a->n, a->exp, a->sign);
}

void ohp_set_d(ohp* r, double v) {
  int n = r->n;
  ohp_zero(r, n);
  if (v == 0.0 || !isfinite(v)) return;
  int e;
  double f = frexp(fabs(v), &e);
  uint64_t mant = (uint64_t)ldexp(f, 64);
  r->d[0] = (uint32_t)(mant >> 32);
  r->d[1] = (uint32_t)mant;
  r->exp = e;
  r->sign = v < 0 ? -1 : 1;
}

double ohp_get_d(const ohp* a) {
  if (a->sign == 0) return 0.0;
  uint64_t top = ((uint64_t)a->d[0] << 32) | a->d[1];
  long e = a->exp;
  if (e > 1024) return a->sign > 0 ? HUGE_VAL : -HUGE_VAL;
  int keep = 53;
  if (e - 53 < -1074) keep = (int)(e + 1074);
  if (keep <= 0) return a->sign > 0 ? 0.0 : -0.0;
  uint64_t q = top >> (64 - keep);
  double d = ldexp((double)q, (int)(e - keep));
  return a->sign < 0 ? -d : d;
}

double ohp_get_d_2exp(long* e, const ohp* a) {
  if (a->sign == 0) {
    *e = 0;
    return 0.0;
  }
  *e = a->exp;
  uint64_t top = ((uint64_t)a->d[0] << 32) | a->d[1];
  double d = ldexp((double)(top >> 11), -53);
  return a->sign < 0 ? -d : d;
}

static int cmp_mag(const ohp* a, const ohp* b) {
  if (a->exp != b->exp) return a->exp > b->exp ? 1 : -1;
  int n = a->n > b->n ? a->n : b->n;
  for (int i = 0; i < n; i++) {
    uint32_t x = i < a->n ? a->d[i] : 0, y = i < b->n ? b->d[i] : 0;
    if (x != y) return x > y ? 1 : -1;
  }
  return 0;
}

int ohp_cmp(const ohp* a, const ohp* b) {
  if (a->sign != b->sign) return a->sign > b->sign ? 1 : -1;
  if (a->sign == 0) return 0;
  int c = cmp_mag(a, b);
  return a->sign > 0 ? c : -c;
}

/* Shift buf (most significant first, nb limbs) right by sh bits; return 1 if any
 * nonzero bit is lost. */
static int shift_right(uint32_t* buf, int nb, long sh) {
  if (sh <= 0) return 0;
  int lost = 0;
  long ls = sh / 32;
  int bs = (int)(sh % 32);
  uint32_t out[2 * OHP_MAXN + 8];
  for (int i = 0; i < nb; i++) {
    long src = i - ls; /* limb feeding the high bits */
    uint32_t hi = (src >= 0 && src < nb) ? buf[src] : 0;
    uint32_t lo = (src - 1 >= 0 && src - 1 < nb) ? buf[src - 1] : 0;
    out[i] = bs ? (hi >> bs) | (lo << (32 - bs)) : hi;
  }
  /* bits that fall off: limbs with index >= nb - ls, plus the low bs bits of
   * limb nb-1-ls */
  for (long i = nb - ls; i < nb; i++)
    if (i >= 0 && buf[i]) lost = 1;
  long j = nb - 1 - ls;
  if (bs && j >= 0 && j < nb && (buf[j] << (32 - bs))) lost = 1;
  memcpy(buf, out, sizeof(uint32_t) * nb);
  return lost;
}

static void add_signed(ohp* r, const ohp* a, int sa, const ohp* b, int sb) {
  if (b->sign == 0 || sb == 0) {
    ohp t;
    t.n = r->n;
    ohp_copy(&t, a);
    t.sign = a->sign == 0 ? 0 : sa;
    *r = t;
    return;
  }
  if (a->sign == 0 || sa == 0) {
    ohp t;
    t.n = r->n;
    ohp_copy(&t, b);
    t.sign = sb;
    *r = t;
    return;
  }
  const ohp* x = a;
  const ohp* y = b;
  int sx = sa, sy = sb;
  if (cmp_mag(a, b) < 0) {
    x = b; y = a;
    sx = sb; sy = sa;
  }
  int m = x->n > y->n ? x->n : y->n;
  if (r->n > m) m = r->n;
  int nb = m + 3; /* carry limb + mantissa + two guard limbs */
  uint32_t X[OHP_MAXN + 3], Y[OHP_MAXN + 3];
  memset(X, 0, sizeof(X));
  memset(Y, 0, sizeof(Y));
  memcpy(X + 1, x->d, sizeof(uint32_t) * x->n);
  memcpy(Y + 1, y->d, sizeof(uint32_t) * y->n);
  int lost = shift_right(Y, nb, x->exp - y->exp);
  long e = x->exp + 32;
  if (sx == sy) {
    uint64_t c = 0;
    for (int i = nb - 1; i >= 0; i--) {
      c += (uint64_t)X[i] + Y[i];
      X[i] = (uint32_t)c;
      c >>= 32;
    }
  } else {
    int64_t br = 0;
    for (int i = nb - 1; i >= 0; i--) {
      int64_t dd = (int64_t)X[i] - Y[i] - br;
      br = dd < 0;
      X[i] = (uint32_t)dd;
    }
    if (lost) { /* exact = X - fraction: floor is X - 1 */
      for (int i = nb - 1; i >= 0; i--) {
        if (X[i]--) break;
      }
    }
  }
  store(r, X, nb, e, sx);
}

void ohp_add(ohp* r, const ohp* a, const ohp* b) { add_signed(r, a, a->sign, b, b->sign); }
void ohp_sub(ohp* r, const ohp* a, const ohp* b) { add_signed(r, a, a->sign, b, -b->sign); }

void ohp_mul(ohp* r, const ohp* a, const ohp* b) {
  if (a->sign == 0 || b->sign == 0) {
    ohp_zero(r, r->n);
    return;
  }
  int na = a->n, nb = b->n, nt = na + nb;
  uint32_t t[2 * OHP_MAXN];
  memset(t, 0, sizeof(uint32_t) * nt);
  /* t is most-significant-first: limb i of a (weight 2^-32(i+1)) times limb j of
   * b lands at position i + j + 1. */
  for (int i = na - 1; i >= 0; i--) {
    uint64_t c = 0;
    uint64_t ai = a->d[i];
    for (int j = nb - 1; j >= 0; j--) {
      c += ai * b->d[j] + t[i + j + 1];
      t[i + j + 1] = (uint32_t)c;
      c >>= 32;
    }
    t[i] = (uint32_t)c;
  }
  store(r, t, nt, a->exp + b->exp, a->sign * b->sign);
}

void ohp_mul_si(ohp* r, const ohp* a, long v) {
  ohp t;
  ohp_zero(&t, 4);
  if (v != 0) {
    unsigned long u = v < 0 ? 0ul - (unsigned long)v : (unsigned long)v;
    uint32_t buf[2] = {(uint32_t)(u >> 32), (uint32_t)u};
    store(&t, buf, 2, 64, v < 0 ? -1 : 1);
  }
  ohp_mul(r, a, &t);
}

void ohp_mul_d(ohp* r, const ohp* a, double v) {
  ohp t;
  t.n = 4;
  ohp_set_d(&t, v);
  ohp_mul(r, a, &t);
}

void ohp_add_d(ohp* r, const ohp* a, double v) {
  ohp t;
  t.n = 4;
  ohp_set_d(&t, v);
  ohp_add(r, a, &t);
}

void ohp_div_ui(ohp* r, const ohp* a, unsigned long v) {
  if (a->sign == 0 || v == 0) {
    ohp_zero(r, r->n);
    return;
  }
  int ext = r->n + 2;
  int nb = a->n + ext;
  uint32_t q[3 * OHP_MAXN + 4];
  u128 rem = 0;
  for (int i = 0; i < nb; i++) {
    uint32_t limb = i < a->n ? a->d[i] : 0;
    u128 cur = (rem << 32) | limb;
    q[i] = (uint32_t)(cur / v);
    rem = cur % v;
  }
  store(r, q, nb, a->exp, a->sign);
}

/* ---- decimal parsing: value = D * 10^e10 = D * 5^e10 * 2^e10 ---------------- */
/* little-endian 32-bit big integers */
typedef struct {
  int n;
  uint32_t* w;
  int cap;
} bigi;

static void bi_init(bigi* b, int cap) {
  b->n = 0;
  b->cap = cap;
  b->w = (uint32_t*)calloc((size_t)cap, sizeof(uint32_t));
}
static void bi_grow(bigi* b, int need) {
  if (need <= b->cap) return;
  int cap = need * 2;
  b->w = (uint32_t*)realloc(b->w, sizeof(uint32_t) * (size_t)cap);
  memset(b->w + b->cap, 0, sizeof(uint32_t) * (size_t)(cap - b->cap));
  b->cap = cap;
}
static void bi_muladd(bigi* b, uint32_t m, uint32_t add) {
  uint64_t c = add;
  for (int i = 0; i < b->n; i++) {
    c += (uint64_t)b->w[i] * m;
    b->w[i] = (uint32_t)c;
    c >>= 32;
  }
  if (c) {
    bi_grow(b, b->n + 1);
    b->w[b->n++] = (uint32_t)c;
  }
}
static void bi_div(bigi* b, uint32_t v) {
  uint64_t rem = 0;
  for (int i = b->n - 1; i >= 0; i--) {
    uint64_t cur = (rem << 32) | b->w[i];
    b->w[i] = (uint32_t)(cur / v);
    rem = cur % v;
  }
  while (b->n > 0 && b->w[b->n - 1] == 0) b->n--;
}
static long bi_bits(const bigi* b) {
  if (b->n == 0) return 0;
  return 32L * (b->n - 1) + (32 - clz32(b->w[b->n - 1]));
}
static void bi_shl(bigi* b, long sh) {
  if (b->n == 0 || sh <= 0) return;
  long ls = sh / 32;
  int bs = (int)(sh % 32);
  int nn = b->n + (int)ls + 1;
  bi_grow(b, nn);
  for (int i = nn - 1; i >= 0; i--) {
    long src = i - ls;
    uint32_t hi = (src >= 0 && src < b->n) ? b->w[src] : 0;
    uint32_t lo = (src - 1 >= 0 && src - 1 < b->n) ? b->w[src - 1] : 0;
    b->w[i] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
  }
  b->n = nn;
  while (b->n > 0 && b->w[b->n - 1] == 0) b->n--;
}

int ohp_set_str(ohp* r, const char* s) {
  const char* p = s;
  while (*p == ' ' || *p == '\t' || *p == '\n') p++;
  int sign = 1;
  if (*p == '-') {
    sign = -1;
    p++;
  } else if (*p == '+')
    p++;
  bigi D;
  bi_init(&D, 64);
  long frac = 0;
  int dot = 0, any = 0;
  for (; *p; p++) {
    if (*p >= '0' && *p <= '9') {
      bi_muladd(&D, 10, (uint32_t)(*p - '0'));
      if (D.n == 0 && *p != '0') D.w[D.n++] = (uint32_t)(*p - '0');
      if (dot) frac++;
      any = 1;
    } else if (*p == '.' && !dot)
      dot = 1;
    else
      break;
  }
  if (!any) {
    free(D.w);
    ohp_zero(r, r->n);
    return -1;
  }
  long e10 = 0;
  if (*p == 'e' || *p == 'E' || *p == '@') {
    char* end;
    p++;
    e10 = strtol(p, &end, 10);
    if (end == p) {
      free(D.w);
      ohp_zero(r, r->n);
      return -1;
    }
    p = end;
  }
  while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
  if (*p) {
    free(D.w);
    ohp_zero(r, r->n);
    return -1;
  }
  e10 -= frac;
  if (D.n == 0) {
    free(D.w);
    ohp_zero(r, r->n);
    return 0;
  }
  const uint32_t P5_13 = 1220703125u; /* 5^13 */
  long bexp;                          /* value = D * 2^bexp after the 5-power step */
  if (e10 >= 0) {
    long k = e10;
    while (k >= 13) {
      bi_muladd(&D, P5_13, 0);
      k -= 13;
    }
    uint32_t p5 = 1;
    while (k-- > 0) p5 *= 5;
    bi_muladd(&D, p5, 0);
    bexp = e10;
  } else {
    long k = -e10;
    long P = 32L * r->n;
    long sh = P + 8 + (long)ceil(k * 2.321928094887362) + 1 - bi_bits(&D);
    if (sh < 0) sh = 0;
    bi_shl(&D, sh);
    long kk = k;
    while (kk >= 13) {
      bi_div(&D, P5_13);
      kk -= 13;
    }
    uint32_t p5 = 1;
    while (kk-- > 0) p5 *= 5;
    bi_div(&D, p5);
    bexp = -k - sh;
  }
  /* D (little-endian) -> most-significant-first buffer */
  int nb = D.n;
  uint32_t* buf = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nb > 0 ? nb : 1));
  for (int i = 0; i < nb; i++) buf[i] = D.w[nb - 1 - i];
  /* value = D * 2^bexp = 0.buf * 2^(32 nb + bexp) */
  ohp t;
  t.n = r->n;
  store(&t, buf, nb, 32L * nb + bexp, sign);
  *r = t;
  free(buf);
  free(D.w);
  return 0;
}
