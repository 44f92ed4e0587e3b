/* oracle.c - CPU oracle for newman's per-pixel deep-zoom iteration (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may build, load or call this file.  Each function cites the
 * reference file:line it restates.  Built with -O2 -ffp-contract=off (no FMA
 * contraction, like makefile:3, which has no -march on x86-64).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------ state */
#define FX_ZERO_E (-0x40000000) /* the exponent of a floatexp zero */
struct orc_view {
  int nr, nc, N;
  double tol;     /* error_tolerance, mandelbrot.cpp:9 */
  double bailout; /* mandelbrot.cpp:58 (1024) */
  long bits;      /* setPrecision result */
  int n;          /* ohp limbs at `bits` */
  ohp cre, cim, szre, szim;
  /* reference orbit and series (precompute) */
  int len;            /* len(X) */
  ohp* X;             /* 2*(len+1): X[0..len-1] and one extra element X[len] */
  double* Xt;         /* 2*(len+1) truncated doubles of X */
  double* Xn;         /* 2*(len+1) nearest doubles of X */
  double* Xlo;        /* 2*(len+1) nearest(X - Xn) */
  double *At, *Bt, *Ct; /* 2*len truncated A, B, C (mandelbrot.cpp:166-168) */
  /* floatexp series (an extension, not the reference's arithmetic): per index and
   * component the truncated 53-bit mantissa in [1/2, 1) and its exponent */
  double *Am, *Bm, *Cm;
  int *Ae, *Be, *Ce;
  int fxs; /* 1: the series test and skip in floatexp */
  int winner;
};

static int g_nthreads(int want) {
  if (want > 0) return want;
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

/* -------------------------------------------------------------- threading */
typedef struct {
  atomic_long next;
  long total;
  long chunk;
  void (*fn)(void* ctx, long i);
  void* ctx;
} pool_job;

static void* pool_worker(void* arg) {
  pool_job* j = (pool_job*)arg;
  for (;;) {
    long s = atomic_fetch_add(&j->next, j->chunk);
    if (s >= j->total) break;
    long e = s + j->chunk < j->total ? s + j->chunk : j->total;
    for (long i = s; i < e; i++) j->fn(j->ctx, i);
  }
  return NULL;
}

static void parallel_for(long total, int nthreads, long chunk, void (*fn)(void*, long), void* ctx) {
  pool_job j;
  atomic_init(&j.next, 0);
  j.total = total;
  j.chunk = chunk > 0 ? chunk : 1;
  j.fn = fn;
  j.ctx = ctx;
  int nt = g_nthreads(nthreads);
  if (nt > 64) nt = 64;
  if (nt <= 1 || total <= 1) {
    pool_worker(&j);
    return;
  }
  pthread_t th[64];
  for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, pool_worker, &j);
  for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------ small helpers */
static void hp_init(ohp* x, int n) { ohp_zero(x, n); }

static void free_series(orc_view* v) {
  free(v->X);
  free(v->Xt);
  free(v->Xn);
  free(v->Xlo);
  free(v->At);
  free(v->Bt);
  free(v->Ct);
  free(v->Am); free(v->Bm); free(v->Cm);
  free(v->Ae); free(v->Be); free(v->Ce);
  v->X = NULL;
  v->Xt = v->Xn = v->Xlo = v->At = v->Bt = v->Ct = NULL;
  v->Am = v->Bm = v->Cm = NULL;
  v->Ae = v->Be = v->Ce = NULL;
  v->len = 0;
  v->winner = -1;
}

/* mandelbrot.cpp:37-56.  bits = max(64, (int)(64 - e + log2(1e-20))) with
 * sz.re = m * 2^e, m in [1/2, 1) (mpf_get_d_2exp). */
void orc_set_precision(orc_view* v) {
  const double log_alpha = log2(1.0e-20);
  const int beta = 64;
  long e;
  ohp_get_d_2exp(&e, &v->szre);
  int bits = (int)(beta - e + log_alpha);
  if (bits < 64) bits = 64;
  v->bits = bits;
  int n = ohp_limbs_for_bits(bits);
  ohp t;
  t.n = n; ohp_copy(&t, &v->cre); v->cre = t;
  t.n = n; ohp_copy(&t, &v->cim); v->cim = t;
  t.n = n; ohp_copy(&t, &v->szre); v->szre = t;
  t.n = n; ohp_copy(&t, &v->szim); v->szim = t;
  v->n = n;
  free_series(v);
}

long orc_precision_bits(const orc_view* v) { return v->bits; }

/* mandelbrot.cpp:8-17 */
orc_view* orc_create(int nr, int nc) {
  orc_view* v = (orc_view*)calloc(1, sizeof(orc_view));
  v->nr = nr;
  v->nc = nc;
  v->tol = 1e-10;
  v->N = 256;
  v->bailout = 1024.0;
  v->winner = -1;
  int n = ohp_limbs_for_bits(64);
  v->n = n;
  hp_init(&v->cre, n); hp_init(&v->cim, n); hp_init(&v->szre, n); hp_init(&v->szim, n);
  ohp_set_d(&v->cre, -0.5);
  ohp_set_d(&v->cim, 0.0);
  ohp_set_d(&v->szre, 4.0 / nc);
  ohp_set_d(&v->szim, 3.0 / nr);
  orc_set_precision(v);
  return v;
}

void orc_destroy(orc_view* v) {
  if (!v) return;
  free_series(v);
  free(v);
}

void orc_set_N(orc_view* v, int N) { v->N = N; }
void orc_set_tol(orc_view* v, double tol) { v->tol = tol; }
void orc_set_bailout(orc_view* v, double b) { v->bailout = b; }
void orc_set_series_mode(orc_view* v, int fxs) { v->fxs = fxs != 0; }

int orc_set_center(orc_view* v, const char* re, const char* im) {
  int a = ohp_set_str(&v->cre, re);
  int b = ohp_set_str(&v->cim, im);
  return a || b ? -1 : 0;
}

int orc_set_size(orc_view* v, const char* re, const char* im) {
  int a = ohp_set_str(&v->szre, re);
  int b = ohp_set_str(&v->szim, im);
  return a || b ? -1 : 0;
}

/* sz = width / cols (square samples) at the view's current precision, without
 * setPrecision: the frame sizes of a zoom-video batch, which share one reference
 * and the deepest frame's precision (the batch builder parses each frame's width
 * at that precision and divides by the column count) */
int orc_set_width(orc_view* v, const char* width) {
  ohp w;
  w.n = v->szre.n;
  if (ohp_set_str(&w, width)) return -1;
  ohp_div_ui(&v->szre, &w, (unsigned long)v->nc);
  ohp_copy(&v->szim, &v->szre);
  return 0;
}

int orc_set_view(orc_view* v, int N, const char* cre, const char* cim, const char* width) {
  /* parse the width generously, derive the precision the view needs, then parse
   * the centre at that precision */
  ohp w;
  w.n = ohp_limbs_for_bits(256);
  if (ohp_set_str(&w, width)) return -1;
  ohp sz;
  sz.n = w.n;
  ohp_div_ui(&sz, &w, (unsigned long)v->nc);
  long e;
  ohp_get_d_2exp(&e, &sz);
  int bits = (int)(64 - e + log2(1.0e-20));
  if (bits < 64) bits = 64;
  int n = ohp_limbs_for_bits(bits);
  v->N = N;
  v->cre.n = n; v->cim.n = n; v->szre.n = n; v->szim.n = n;
  if (ohp_set_str(&v->cre, cre) || ohp_set_str(&v->cim, cim)) return -1;
  ohp wn;
  wn.n = n;
  ohp_set_str(&wn, width);
  ohp_div_ui(&v->szre, &wn, (unsigned long)v->nc);
  ohp_copy(&v->szim, &v->szre);
  orc_set_precision(v);
  return 0;
}

/* mandelbrot.cpp:256-259 */
int orc_use_hardware(const orc_view* v) {
  const double minpreview = 1.5e-16;
  return ohp_get_d(&v->szre) >= minpreview && ohp_get_d(&v->szim) >= minpreview;
}

/* mandelbrot.cpp:19-35: values parsed at the current precision, sz rescaled by
 * 800/cols and 600/rows (double factors), then setPrecision. */
int orc_load_legacy(orc_view* v, const char* path) {
  FILE* fp = fopen(path, "r");
  if (!fp) return -1;
  char buf[4096];
  int N;
  if (fscanf(fp, "%d\n", &N) != 1) { fclose(fp); return -1; }
  v->N = N;
  if (fscanf(fp, "%4095s\n", buf) == 1) ohp_set_str(&v->cre, buf);
  if (fscanf(fp, "%4095s\n", buf) == 1) ohp_set_str(&v->cim, buf);
  if (fscanf(fp, "%4095s\n", buf) == 1) ohp_set_str(&v->szre, buf);
  if (fscanf(fp, "%4095s\n", buf) == 1) ohp_set_str(&v->szim, buf);
  fclose(fp);
  ohp t;
  t.n = v->szre.n; ohp_mul_d(&t, &v->szre, 800.0 / v->nc); v->szre = t;
  t.n = v->szim.n; ohp_mul_d(&t, &v->szim, 600.0 / v->nr); v->szim = t;
  orc_set_precision(v);
  return 0;
}

/* mandelbrot.cpp:297-302: sz *= (1.0 / scale) with a float scale */
void orc_zoom(orc_view* v, float scale) {
  double f = 1.0 / scale;
  ohp t;
  t.n = v->szre.n; ohp_mul_d(&t, &v->szre, f); v->szre = t;
  t.n = v->szim.n; ohp_mul_d(&t, &v->szim, f); v->szim = t;
  orc_set_precision(v);
}

/* mandelbrot.cpp:304-316 */
void orc_zoom_at(orc_view* v, float scale, int r, int c, int sc) {
  int n = v->n;
  ohp pre, pim, t;
  pre.n = n; pim.n = n; t.n = n;
  ohp_mul_si(&t, &v->szre, (long)(sc * c - v->nc / 2));
  ohp_add(&pre, &v->cre, &t);
  ohp_mul_si(&t, &v->szim, (long)(v->nr / 2 - sc * r - 1));
  ohp_add(&pim, &v->cim, &t);
  double f = 1.0 / scale;
  ohp u;
  u.n = v->szre.n; ohp_mul_d(&u, &v->szre, f); v->szre = u;
  u.n = v->szim.n; ohp_mul_d(&u, &v->szim, f); v->szim = u;
  t.n = v->cre.n;
  ohp_mul_si(&t, &v->szre, (long)(sc * c - v->nc / 2));
  ohp_sub(&v->cre, &pre, &t);
  t.n = v->cim.n;
  ohp_mul_si(&t, &v->szim, (long)(v->nr / 2 - sc * r - 1));
  ohp_sub(&v->cim, &pim, &t);
  orc_set_precision(v);
}

/* mandelbrot.cpp:292-295 (no setPrecision) */
void orc_translate(orc_view* v, int dr, int dc, int sc) {
  ohp t;
  t.n = v->n;
  ohp_mul_si(&t, &v->szre, (long)dc * sc);
  ohp_sub(&v->cre, &v->cre, &t);
  ohp_mul_si(&t, &v->szim, (long)dr * sc);
  ohp_add(&v->cim, &v->cim, &t);
}

/* scaleUp: sz = sz / sc (mandelbrot.cpp:342-343); scaleDown: sz = sz * sc
 * (mandelbrot.cpp:356-357); both then setPrecision */
void orc_scale_size(orc_view* v, int mul, int div) {
  ohp t;
  if (mul != 1) {
    t.n = v->szre.n; ohp_mul_si(&t, &v->szre, mul); v->szre = t;
    t.n = v->szim.n; ohp_mul_si(&t, &v->szim, mul); v->szim = t;
  }
  if (div != 1) {
    t.n = v->szre.n; ohp_div_ui(&t, &v->szre, (unsigned long)div); v->szre = t;
    t.n = v->szim.n; ohp_div_ui(&t, &v->szim, (unsigned long)div); v->szim = t;
  }
  orc_set_precision(v);
}

static void hp_to_string(const ohp* x, char* out, int cap);

void orc_get_center(const orc_view* v, char* re, char* im, int cap) {
  hp_to_string(&v->cre, re, cap);
  hp_to_string(&v->cim, im, cap);
}

void orc_get_size_d(const orc_view* v, double* re, double* im) {
  *re = ohp_get_d(&v->szre);
  *im = ohp_get_d(&v->szim);
}

/* ------------------------------------------------------- pixel map, cardioid */
/* computeRow / findProbe pixel map (mandelbrot.cpp:271, 275, 86-87):
 * re = center.re + (c - cols/2) * sz.re ; im = center.im + (rows/2 - r - 1) * sz.im */
static void pixel_coord(const orc_view* v, int r, int c, ohp* re, ohp* im) {
  int n = v->n;
  ohp t;
  t.n = n; re->n = n; im->n = n;
  ohp_mul_si(&t, &v->szre, (long)(c - v->nc / 2));
  ohp_add(re, &v->cre, &t);
  ohp_mul_si(&t, &v->szim, (long)(v->nr / 2 - r - 1));
  ohp_add(im, &v->cim, &t);
}

/* mandelbrot.cpp:63-71, all in HP at the working precision */
static int in_cardioid(const ohp* zr, const ohp* zi, int n) {
  ohp fourth, xmf, y2, q, t, lhs, rhs;
  fourth.n = xmf.n = y2.n = q.n = t.n = lhs.n = rhs.n = n;
  ohp_set_d(&fourth, 0.25);
  ohp_sub(&xmf, zr, &fourth);
  ohp_mul(&y2, zi, zi);
  ohp_mul(&q, &xmf, &xmf);
  ohp_add(&q, &q, &y2);
  ohp_add(&t, &q, &xmf);
  ohp_mul(&lhs, &q, &t);
  ohp_mul(&rhs, &fourth, &y2);
  if (ohp_cmp(&lhs, &rhs) < 0) return 1; /* cardioid */
  ohp_add_d(&q, zr, 1.0);
  ohp_mul(&lhs, &q, &q);
  ohp_add(&lhs, &lhs, &y2);
  ohp_mul(&rhs, &fourth, &fourth);
  return ohp_cmp(&lhs, &rhs) < 0; /* second disk */
}

int orc_pixel_in_cardioid(orc_view* v, int r, int c) {
  ohp re, im;
  pixel_coord(v, r, c, &re, &im);
  return in_cardioid(&re, &im, v->n);
}

void orc_pixel_coord_d(orc_view* v, int r, int c, double* z) {
  ohp re, im;
  pixel_coord(v, r, c, &re, &im);
  z[0] = ohp_get_d(&re);
  z[1] = ohp_get_d(&im);
}

/* ------------------------------------------------------------ LP helpers */
/* complex.h:19-31, fixed operation order, no FMA (this file is built with
 * -ffp-contract=off) */
typedef struct {
  double re, im;
} lpc;
typedef struct {
  double m;
  int e;
} fxr;
typedef struct {
  fxr re, im;
} fxc;
static inline lpc lp_sq(lpc a) {
  lpc r = {a.re * a.re - a.im * a.im, 2.0 * a.re * a.im};
  return r;
}
static inline double lp_sqmag(lpc a) { return a.re * a.re + a.im * a.im; }
static inline lpc lp_add(lpc a, lpc b) {
  lpc r = {a.re + b.re, a.im + b.im};
  return r;
}
static inline lpc lp_mul(lpc a, lpc b) {
  lpc r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

/* mandelbrot.cpp:133-136 */
static double smoothing_of(double r2, double bailout) { return 1.0 - log2(0.5 * log(r2) / log(bailout)); }

/* ------------------------------------------------------- plain double path */
/* mandelbrot.cpp:231-254 */
static orc_escape hw_iterations(const ohp* y0r, const ohp* y0i, int n, int N, double bailout) {
  orc_escape esc;
  lpc z0 = {ohp_get_d(y0r), ohp_get_d(y0i)};
  if (in_cardioid(y0r, y0i, n)) {
    esc.iterations = N;
    esc.smoothing = 0.0f;
    return esc;
  }
  const double b2 = bailout * bailout;
  lpc z = z0;
  int it;
  for (it = 0; it < N; it++) {
    z = lp_add(lp_sq(z), z0);
    if (lp_sqmag(z) > b2) break;
  }
  esc.iterations = it;
  esc.smoothing = it < N ? (float)smoothing_of(lp_sqmag(z), bailout) : 0.0f;
  return esc;
}

orc_escape orc_point_hw(const char* re, const char* im, int N, long bits) {
  int n = ohp_limbs_for_bits(bits);
  ohp zr, zi;
  zr.n = zi.n = n;
  ohp_set_str(&zr, re);
  ohp_set_str(&zi, im);
  return hw_iterations(&zr, &zi, n, N, 1024.0);
}

int orc_point_in_cardioid(const char* re, const char* im, long bits) {
  int n = ohp_limbs_for_bits(bits);
  ohp zr, zi;
  zr.n = zi.n = n;
  ohp_set_str(&zr, re);
  ohp_set_str(&zi, im);
  return in_cardioid(&zr, &zi, n);
}

/* ------------------------------------------------------- reference orbit */
static int bailed_out(const ohp* re, const ohp* im, double b2) {
  lpc z = {ohp_get_d(re), ohp_get_d(im)};
  return lp_sqmag(z) > b2;
}

/* mandelbrot.cpp:97-110 length only: X[0] = X0, X[i] = X[i-1]^2 + X0, size = first
 * i with bailedOut(X[i]), else N */
static int orbit_length(const ohp* x0r, const ohp* x0i, int n, int N, double b2) {
  ohp xr = *x0r, xi = *x0i, t1, t2, nr, ni;
  t1.n = t2.n = nr.n = ni.n = n;
  for (int i = 1; i < N; i++) {
    ohp_mul(&t1, &xr, &xr);
    ohp_mul(&t2, &xi, &xi);
    ohp_sub(&t1, &t1, &t2);
    ohp_add(&nr, &t1, x0r);
    ohp_mul(&t1, &xr, &xi);
    if (t1.sign) t1.exp += 1; /* 2.0 * (...) is exact */
    ohp_add(&ni, &t1, x0i);
    if (bailed_out(&nr, &ni, b2)) return i;
    xr = nr;
    xi = ni;
  }
  return N;
}

/* the stored orbit plus one extra element X[len] */
static void orbit_store(orc_view* v, const ohp* x0r, const ohp* x0i, int len) {
  int n = v->n;
  v->len = len;
  v->X = (ohp*)malloc(sizeof(ohp) * 2 * (size_t)(len + 1));
  v->X[0] = *x0r;
  v->X[1] = *x0i;
  ohp t1, t2;
  t1.n = t2.n = n;
  for (int i = 1; i <= len; i++) {
    ohp* pr = &v->X[2 * (i - 1)];
    ohp* pi = &v->X[2 * (i - 1) + 1];
    ohp* xr = &v->X[2 * i];
    ohp* xi = &v->X[2 * i + 1];
    xr->n = xi->n = n;
    ohp_mul(&t1, pr, pr);
    ohp_mul(&t2, pi, pi);
    ohp_sub(&t1, &t1, &t2);
    ohp_add(xr, &t1, x0r);
    ohp_mul(&t1, pr, pi);
    if (t1.sign) t1.exp += 1;
    ohp_add(xi, &t1, x0i);
  }
}

/* nearest-rounded double of an HP value */
static double get_d_nearest(const ohp* a) {
  double t = ohp_get_d(a);
  if (!isfinite(t) || a->sign == 0) return t;
  ohp r, td;
  r.n = a->n;
  td.n = 4;
  ohp_set_d(&td, t);
  ohp_sub(&r, a, &td); /* remainder, same sign as a (or zero) */
  if (r.sign == 0) return t;
  double up = nextafter(t, a->sign > 0 ? INFINITY : -INFINITY);
  ohp half;
  half.n = 4;
  ohp_set_d(&half, (up - t) * 0.5); /* exact: half an ulp */
  ohp ar = r, ah = half;
  ar.sign = 1;
  ah.sign = 1;
  int c = ohp_cmp(&ar, &ah);
  if (c > 0) return up;
  if (c < 0) return t;
  /* tie: even mantissa */
  uint64_t bits;
  memcpy(&bits, &t, 8);
  return (bits & 1) ? up : t;
}

static void store_doubles(orc_view* v) {
  int len = v->len;
  v->Xt = (double*)malloc(sizeof(double) * 2 * (size_t)(len + 1));
  v->Xn = (double*)malloc(sizeof(double) * 2 * (size_t)(len + 1));
  v->Xlo = (double*)malloc(sizeof(double) * 2 * (size_t)(len + 1));
  for (int i = 0; i < 2 * (len + 1); i++) {
    const ohp* x = &v->X[i];
    v->Xt[i] = ohp_get_d(x);
    v->Xn[i] = get_d_nearest(x);
    ohp r, td;
    r.n = x->n;
    td.n = 4;
    ohp_set_d(&td, v->Xn[i]);
    ohp_sub(&r, x, &td);
    v->Xlo[i] = get_d_nearest(&r);
  }
}

/* mandelbrot.cpp:112-131, the quirky recurrence with the new-index A[i], B[i] */
static void compute_series(orc_view* v) {
  int len = v->len, n = v->n;
  v->At = (double*)malloc(sizeof(double) * 2 * (size_t)len);
  v->Bt = (double*)malloc(sizeof(double) * 2 * (size_t)len);
  v->Ct = (double*)malloc(sizeof(double) * 2 * (size_t)len);
  double** fm[3] = {&v->Am, &v->Bm, &v->Cm};
  int** fe[3] = {&v->Ae, &v->Be, &v->Ce};
  for (int k = 0; k < 3; k++) {
    *fm[k] = (double*)calloc(2 * (size_t)len, sizeof(double));
    *fe[k] = (int*)malloc(sizeof(int) * 2 * (size_t)len);
    for (size_t j = 0; j < 2 * (size_t)len; j++) (*fe[k])[j] = FX_ZERO_E;
  }
  v->Am[0] = 0.5;
  v->Ae[0] = 1;
  ohp Ar, Ai, Br, Bi, Cr, Ci, nAr, nAi, nBr, nBi, nCr, nCi, t, u;
  ohp* all[] = {&Ar, &Ai, &Br, &Bi, &Cr, &Ci, &nAr, &nAi, &nBr, &nBi, &nCr, &nCi, &t, &u};
  for (int k = 0; k < 14; k++) ohp_zero(all[k], n);
  ohp_set_d(&Ar, 1.0);
  v->At[0] = 1.0; v->At[1] = 0.0;
  v->Bt[0] = v->Bt[1] = v->Ct[0] = v->Ct[1] = 0.0;
  for (int i = 1; i < len; i++) {
    const ohp* xr = &v->X[2 * (i - 1)];
    const ohp* xi = &v->X[2 * (i - 1) + 1];
    /* A.re = 2*(xr*Ar - xi*Ai) + 1 */
    ohp_mul(&nAr, xr, &Ar); ohp_mul(&t, xi, &Ai); ohp_sub(&nAr, &nAr, &t);
    if (nAr.sign) nAr.exp += 1;
    ohp_add_d(&nAr, &nAr, 1.0);
    /* A.im = 2*(xr*Ai + xi*Ar) */
    ohp_mul(&nAi, xr, &Ai); ohp_mul(&t, xi, &Ar); ohp_add(&nAi, &nAi, &t);
    if (nAi.sign) nAi.exp += 1;
    /* B.re = 2*(xr*Br - xi*Bi) + nAr*nAr - nAi*nAi */
    ohp_mul(&nBr, xr, &Br); ohp_mul(&t, xi, &Bi); ohp_sub(&nBr, &nBr, &t);
    if (nBr.sign) nBr.exp += 1;
    ohp_mul(&t, &nAr, &nAr); ohp_add(&nBr, &nBr, &t);
    ohp_mul(&t, &nAi, &nAi); ohp_sub(&nBr, &nBr, &t);
    /* B.im = 2*(xr*Bi + xi*Br + nAr*nAi) */
    ohp_mul(&nBi, xr, &Bi); ohp_mul(&t, xi, &Br); ohp_add(&nBi, &nBi, &t);
    ohp_mul(&t, &nAr, &nAi); ohp_add(&nBi, &nBi, &t);
    if (nBi.sign) nBi.exp += 1;
    /* C.re = 2*(xr*Cr - xi*Ci + nAr*nBr - nAi*nBi) */
    ohp_mul(&nCr, xr, &Cr); ohp_mul(&t, xi, &Ci); ohp_sub(&nCr, &nCr, &t);
    ohp_mul(&t, &nAr, &nBr); ohp_add(&nCr, &nCr, &t);
    ohp_mul(&t, &nAi, &nBi); ohp_sub(&nCr, &nCr, &t);
    if (nCr.sign) nCr.exp += 1;
    /* C.im = 2*(xr*Ci + xi*Cr + nAr*nBi + nAi*nBr) */
    ohp_mul(&nCi, xr, &Ci); ohp_mul(&t, xi, &Cr); ohp_add(&nCi, &nCi, &t);
    ohp_mul(&t, &nAr, &nBi); ohp_add(&nCi, &nCi, &t);
    ohp_mul(&t, &nAi, &nBr); ohp_add(&nCi, &nCi, &t);
    if (nCi.sign) nCi.exp += 1;
    Ar = nAr; Ai = nAi; Br = nBr; Bi = nBi; Cr = nCr; Ci = nCi;
    v->At[2 * i] = ohp_get_d(&Ar); v->At[2 * i + 1] = ohp_get_d(&Ai);
    v->Bt[2 * i] = ohp_get_d(&Br); v->Bt[2 * i + 1] = ohp_get_d(&Bi);
    v->Ct[2 * i] = ohp_get_d(&Cr); v->Ct[2 * i + 1] = ohp_get_d(&Ci);
    const ohp* hv[6] = {&Ar, &Ai, &Br, &Bi, &Cr, &Ci};
    for (int k = 0; k < 6; k++) {
      long e = 0;
      double m = ohp_get_d_2exp(&e, hv[k]);
      (*fm[k / 2])[2 * i + k % 2] = m;
      (*fe[k / 2])[2 * i + k % 2] = m == 0.0 ? FX_ZERO_E : (int)e;
    }
  }
  (void)u;
}

/* findProbe, mandelbrot.cpp:73-95: probe order (rows/4, c), (rows/2, c),
 * (3*rows/4, c) for c = 0, 2, ...; then (r, cols/2) for r = 0, 2, ...;
 * the longest orbit wins, the first one on ties (strict >). */
typedef struct {
  const orc_view* v;
  const int* pts;
  int* lens;
} probe_ctx;

static void probe_fn(void* vp, long i) {
  probe_ctx* P = (probe_ctx*)vp;
  ohp re, im;
  pixel_coord(P->v, P->pts[2 * i], P->pts[2 * i + 1], &re, &im);
  P->lens[i] = orbit_length(&re, &im, P->v->n, P->v->N, P->v->bailout * P->v->bailout);
}

static int probe_list(const orc_view* v, int** out) {
  int cnt = 0;
  int cap = 3 * ((v->nc + 1) / 2) + (v->nr + 1) / 2 + 4;
  int* pts = (int*)malloc(sizeof(int) * 2 * (size_t)cap);
  for (int c = 0; c < v->nc; c += 2) {
    pts[2 * cnt] = v->nr / 4; pts[2 * cnt + 1] = c; cnt++;
    pts[2 * cnt] = v->nr / 2; pts[2 * cnt + 1] = c; cnt++;
    pts[2 * cnt] = 3 * v->nr / 4; pts[2 * cnt + 1] = c; cnt++;
  }
  for (int r = 0; r < v->nr; r += 2) {
    pts[2 * cnt] = r; pts[2 * cnt + 1] = v->nc / 2; cnt++;
  }
  *out = pts;
  return cnt;
}

int orc_precompute(orc_view* v, int nthreads) {
  free_series(v);
  int* pts;
  int P = probe_list(v, &pts);
  int* lens = (int*)malloc(sizeof(int) * (size_t)P);
  probe_ctx ctx = {v, pts, lens};
  parallel_for(P, nthreads, 4, probe_fn, &ctx);
  int best = -1, bestlen = 0;
  for (int i = 0; i < P; i++)
    if (lens[i] > bestlen) {
      bestlen = lens[i];
      best = i;
    }
  ohp re, im;
  pixel_coord(v, pts[2 * best], pts[2 * best + 1], &re, &im);
  orbit_store(v, &re, &im, bestlen);
  store_doubles(v);
  compute_series(v);
  v->winner = best;
  free(lens);
  free(pts);
  return v->len;
}

int orc_precompute_at(orc_view* v, const char* re_s, const char* im_s) {
  free_series(v);
  ohp re, im;
  re.n = im.n = v->n;
  ohp_set_str(&re, re_s);
  ohp_set_str(&im, im_s);
  int len = orbit_length(&re, &im, v->n, v->N, v->bailout * v->bailout);
  orbit_store(v, &re, &im, len);
  store_doubles(v);
  compute_series(v);
  v->winner = -1;
  return len;
}

int orc_probe_winner(const orc_view* v) { return v->winner; }
int orc_ref_len(const orc_view* v) { return v->len; }

void orc_get_series(const orc_view* v, double* Xt, double* Xn, double* A, double* B, double* C) {
  size_t nx = 2 * (size_t)(v->len + 1), ns = 2 * (size_t)v->len;
  if (Xt) memcpy(Xt, v->Xt, sizeof(double) * nx);
  if (Xn) memcpy(Xn, v->Xn, sizeof(double) * nx);
  if (A) memcpy(A, v->At, sizeof(double) * ns);
  if (B) memcpy(B, v->Bt, sizeof(double) * ns);
  if (C) memcpy(C, v->Ct, sizeof(double) * ns);
}

/* the floatexp coefficients (mantissas and exponents, 2*len each) */
void orc_get_series_fx(const orc_view* v, double* Am, int* Ae, double* Bm, int* Be, double* Cm, int* Ce) {
  size_t ns = 2 * (size_t)v->len;
  memcpy(Am, v->Am, sizeof(double) * ns); memcpy(Ae, v->Ae, sizeof(int) * ns);
  memcpy(Bm, v->Bm, sizeof(double) * ns); memcpy(Be, v->Be, sizeof(int) * ns);
  memcpy(Cm, v->Cm, sizeof(double) * ns); memcpy(Ce, v->Ce, sizeof(int) * ns);
}

/* --------------------------------------------------------- SA + HP path */
/* Floatexp series (an extension): the reference's series test and skip
 * (mandelbrot.cpp:138-142, 160-178) in the same operation order, each real held
 * as m * 2^e with m = 0 (e = FX_ZERO_E) or 1/2 <= |m| < 1.  A product or sum
 * rounds the mantissa once, to nearest, as a double operation on the scaled
 * operands does: where the reference's doubles are normal numbers the results
 * are the same, and nothing under- or overflows where they are not. */
static fxr fxr_make(double d, int e) {
  fxr r;
  if (d == 0.0) {
    r.m = 0.0;
    r.e = FX_ZERO_E;
  } else {
    int k = 0;
    r.m = frexp(d, &k);
    r.e = e + k;
  }
  return r;
}
static fxr fxr_mul(fxr a, fxr b) {
  if (a.m == 0.0 || b.m == 0.0) return fxr_make(0.0, 0);
  return fxr_make(a.m * b.m, a.e + b.e);
}
static fxr fxr_add(fxr a, fxr b) {
  if (a.m == 0.0) return b;
  if (b.m == 0.0) return a;
  if (a.e >= b.e) return fxr_make(a.m + ldexp(b.m, b.e - a.e), a.e);
  return fxr_make(ldexp(a.m, a.e - b.e) + b.m, b.e);
}
static fxr fxr_neg(fxr a) {
  a.m = -a.m;
  return a;
}
/* a < b for a, b >= 0 */
static int fxr_lt(fxr a, fxr b) {
  if (b.m == 0.0) return 0;
  if (a.m == 0.0) return 1;
  return a.e < b.e || (a.e == b.e && a.m < b.m);
}
static double fxr_get(fxr a) { return a.m == 0.0 ? 0.0 : ldexp(a.m, a.e); }
static fxc fxc_of(lpc z) {
  fxc r = {fxr_make(z.re, 0), fxr_make(z.im, 0)};
  return r;
}
static fxc fxc_mul(fxc a, fxc b) { /* complex.h:29-31 */
  fxc r = {fxr_add(fxr_mul(a.re, b.re), fxr_neg(fxr_mul(a.im, b.im))),
           fxr_add(fxr_mul(a.re, b.im), fxr_mul(a.im, b.re))};
  return r;
}
static fxc fxc_sq(fxc a) { /* complex.h:19-21: (re^2 - im^2, 2 re im) */
  fxr two = {0.5, 2};
  fxc r = {fxr_add(fxr_mul(a.re, a.re), fxr_neg(fxr_mul(a.im, a.im))), fxr_mul(fxr_mul(two, a.re), a.im)};
  return r;
}
static fxr fxc_sqmag(fxc a) { return fxr_add(fxr_mul(a.re, a.re), fxr_mul(a.im, a.im)); }
static fxc fx_coef(const double* m, const int* e, int i) {
  fxc r = {{m[2 * i], e[2 * i]}, {m[2 * i + 1], e[2 * i + 1]}};
  return r;
}
/* isUnstable(B e^2, C e^3) in floatexp */
static int is_unstable_fx(fxc b, fxc c, double tol) {
  return fxr_lt(fxr_mul(fxc_sqmag(b), fxr_make(tol, 0)), fxc_sqmag(c));
}

/* mandelbrot.cpp:138-142 */
static int is_unstable(lpc b, lpc c, double tol) {
  double bmag = b.re * b.re + b.im * b.im;
  double cmag = c.re * c.re + c.im * c.im;
  return bmag * tol < cmag;
}

static lpc sa_term(const orc_view* v, int i, lpc eps, lpc eps2, lpc eps3) {
  if (i == 0) return eps;
  if (v->fxs) {
    fxc e = fxc_of(eps), e2 = fxc_sq(e), e3 = fxc_mul(e, e2);
    fxc a = fxc_mul(fx_coef(v->Am, v->Ae, i), e);
    fxc b = fxc_mul(fx_coef(v->Bm, v->Be, i), e2);
    fxc c = fxc_mul(fx_coef(v->Cm, v->Ce, i), e3);
    fxc s = {fxr_add(fxr_add(a.re, b.re), c.re), fxr_add(fxr_add(a.im, b.im), c.im)};
    lpc d = {fxr_get(s.re), fxr_get(s.im)};
    return d;
  }
  lpc a = {v->At[2 * i], v->At[2 * i + 1]};
  lpc b = {v->Bt[2 * i], v->Bt[2 * i + 1]};
  lpc c = {v->Ct[2 * i], v->Ct[2 * i + 1]};
  a = lp_mul(a, eps);
  b = lp_mul(b, eps2);
  c = lp_mul(c, eps3);
  return lp_add(lp_add(a, b), c); /* d[i] = a + b + c */
}

/* The SA scan of mandelbrot.cpp:165-181.  Returns d.size() and fills unstable_at
 * with the first unstable i (0 if none). */
static int sa_scan(const orc_view* v, lpc eps, int* unstable_at) {
  lpc eps2 = lp_sq(eps), eps3 = lp_mul(eps, eps2);
  *unstable_at = 0;
  if (v->fxs) {
    fxc e = fxc_of(eps), e2 = fxc_sq(e), e3 = fxc_mul(e, e2);
    for (int i = 1; i < v->len; i++)
      if (is_unstable_fx(fxc_mul(fx_coef(v->Bm, v->Be, i), e2), fxc_mul(fx_coef(v->Cm, v->Ce, i), e3), v->tol)) {
        *unstable_at = i;
        return i - 3 < 1 ? 1 : i - 3;
      }
    return v->len;
  }
  for (int i = 1; i < v->len; i++) {
    lpc a = {v->At[2 * i], v->At[2 * i + 1]};
    lpc b = {v->Bt[2 * i], v->Bt[2 * i + 1]};
    lpc c = {v->Ct[2 * i], v->Ct[2 * i + 1]};
    (void)a;
    b = lp_mul(b, eps2);
    c = lp_mul(c, eps3);
    if (is_unstable(b, c, v->tol)) {
      int good = i - 3;
      if (good < 1) good = 1;
      *unstable_at = i;
      return good;
    }
  }
  return v->len;
}

/* pixel epsilon = trunc(Y0 - X[0]) (mandelbrot.cpp:155-159) */
static lpc pixel_eps(const orc_view* v, const ohp* y0r, const ohp* y0i) {
  ohp yr, yi;
  yr.n = yi.n = v->n;
  ohp_sub(&yr, y0r, &v->X[0]);
  ohp_sub(&yi, y0i, &v->X[1]);
  lpc e = {ohp_get_d(&yr), ohp_get_d(&yi)};
  return e;
}

void orc_pixel_eps(orc_view* v, int r, int c, double* eps) {
  ohp re, im;
  pixel_coord(v, r, c, &re, &im);
  lpc e = pixel_eps(v, &re, &im);
  eps[0] = e.re;
  eps[1] = e.im;
}

/* mandelbrot.cpp:144-229 */
static orc_escape sa_iterations(const orc_view* v, const ohp* y0r, const ohp* y0i) {
  orc_escape esc;
  const int N = v->N, n = v->n;
  const double b2 = v->bailout * v->bailout;
  if (in_cardioid(y0r, y0i, n)) {
    esc.iterations = N;
    esc.smoothing = 0.0f;
    return esc;
  }
  lpc eps = pixel_eps(v, y0r, y0i);
  lpc eps2 = lp_sq(eps), eps3 = lp_mul(eps, eps2);
  int ua;
  int dsize = sa_scan(v, eps, &ua);
  int found = dsize - 1;
  ohp Yr, Yi;
  Yr.n = Yi.n = n;
  lpc d = sa_term(v, found, eps, eps2, eps3);
  ohp_add_d(&Yr, &v->X[2 * found], d.re);
  ohp_add_d(&Yi, &v->X[2 * found + 1], d.im);
  if (bailed_out(&Yr, &Yi, b2)) {
    int low = 0, high = dsize - 1, mid = dsize / 2;
    while (low <= high) {
      d = sa_term(v, mid, eps, eps2, eps3);
      ohp_add_d(&Yr, &v->X[2 * mid], d.re);
      ohp_add_d(&Yi, &v->X[2 * mid + 1], d.im);
      if (!bailed_out(&Yr, &Yi, b2))
        low = mid + 1;
      else {
        high = mid - 1;
        found = mid;
      }
      mid = (low + high) / 2;
    }
    d = sa_term(v, found, eps, eps2, eps3);
    ohp_add_d(&Yr, &v->X[2 * found], d.re);
    ohp_add_d(&Yi, &v->X[2 * found + 1], d.im);
    lpc y = {ohp_get_d(&Yr), ohp_get_d(&Yi)};
    esc.iterations = found;
    esc.smoothing = (float)smoothing_of(lp_sqmag(y), v->bailout);
    return esc;
  }
  /* HP continuation (mandelbrot.cpp:209-224) */
  ohp Ynr, Yni, t1, t2;
  Ynr.n = Yni.n = t1.n = t2.n = n;
  for (int i = dsize; i < N; i++) {
    ohp_mul(&t1, &Yr, &Yr);
    ohp_mul(&t2, &Yi, &Yi);
    ohp_sub(&t1, &t1, &t2);
    ohp_add(&Ynr, &t1, y0r);
    ohp_mul(&t1, &Yr, &Yi);
    if (t1.sign) t1.exp += 1;
    ohp_add(&Yni, &t1, y0i);
    if (bailed_out(&Ynr, &Yni, b2)) {
      lpc y = {ohp_get_d(&Ynr), ohp_get_d(&Yni)};
      esc.iterations = i;
      esc.smoothing = (float)smoothing_of(lp_sqmag(y), v->bailout);
      return esc;
    }
    Yr = Ynr;
    Yi = Yni;
  }
  esc.iterations = N;
  esc.smoothing = 0.0f;
  return esc;
}

int orc_sa_start(orc_view* v, int r, int c, int* found_out, double* dout, orc_escape* esc) {
  ohp re, im;
  pixel_coord(v, r, c, &re, &im);
  lpc eps = pixel_eps(v, &re, &im);
  lpc eps2 = lp_sq(eps), eps3 = lp_mul(eps, eps2);
  int ua;
  int dsize = sa_scan(v, eps, &ua);
  int found = dsize - 1;
  lpc d = sa_term(v, found, eps, eps2, eps3);
  *found_out = found;
  dout[0] = d.re;
  dout[1] = d.im;
  esc->iterations = -1;
  esc->smoothing = 0.0f;
  return ua;
}

/* ------------------------------------------------------ pert64 restatement */
/* Truncation toward zero of hi + lo + d, the way the HIP kernels evaluate
 * descend(X[k] + d) with X[k] held as a double-double. */
static double trunc_sum3(double hi, double lo, double d) {
  double s = hi + d;
  double bp = s - hi;
  double e = (hi - (s - bp)) + (d - bp);
  double t = e + lo;
  double r = s + t;
  double bp2 = r - s;
  double e2 = (s - (r - bp2)) + (t - bp2);
  if (r != 0.0 && e2 != 0.0 && ((e2 > 0.0) != (r > 0.0))) r = nextafter(r, 0.0);
  return r;
}

typedef struct {
  const double* Zd; /* 2*(len+2): Zd[0] = 0, Zd[k] = nearest(X[k-1]) */
  int mend;         /* len + 1 */
  int N;
  double b2, bailout;
} pert_ctx;

/* FP64 perturbation with rebasing from (delta, it0): Y_it = Zd[m] + delta,
 * delta' = delta * (Zd[m] + Y) + eps; rebase to m = 0 when |Y|^2 < |delta|^2 or at
 * the end of the reference.  Y_it0 is not tested (it was tested by the caller).
 * Every operation matches the HIP kernels. */
static orc_escape pert_from(const pert_ctx* P, double dr, double di, double er, double ei, int it0) {
  orc_escape esc;
  int N = P->N;
  if (it0 + 1 >= N) {
    esc.iterations = N;
    esc.smoothing = 0.0f;
    return esc;
  }
  int m = it0 + 1;
  for (int it = it0;; it++) {
    double zr = P->Zd[2 * m], zi = P->Zd[2 * m + 1];
    double yr = zr + dr, yi = zi + di;
    if (it > it0) {
      double r2e = yr * yr + yi * yi; /* sqMag, no FMA */
      if (r2e > P->b2) {
        esc.iterations = it;
        esc.smoothing = (float)smoothing_of(r2e, P->bailout);
        return esc;
      }
      if (it == N - 1) break;
    }
    double r2 = yr * yr + yi * yi; /* sqMag, no FMA: also the rebase test */
    double dmag = fma(dr, dr, di * di);
    double wr = zr + yr, wi = zi + yi;
    if (r2 < dmag || m == P->mend) { /* rebase: d = Y against Zd[0] = 0, so W = Y */
      dr = yr;
      di = yi;
      wr = yr;
      wi = yi;
      m = 0;
    }
    double ndr = fma(dr, wr, fma(-di, wi, er));
    double ndi = fma(dr, wi, fma(di, wr, ei));
    dr = ndr;
    di = ndi;
    m++;
  }
  esc.iterations = N;
  esc.smoothing = 0.0f;
  return esc;
}

/* The FP32X form (mantissa pair w in float, shared int exponent ex; es = e * 2^-ex)
 * of pert_from, as the HIP kernel runs it: the first step (at it0) in FP64 like
 * pert_from, then per step in float d = w 2^ex, Y = Z + d, the rebase test, W and
 * w' = w W + es; the escape test in double when the float |Y|^2 exceeds
 * b2 (1 - 2^-10) (Y = Zd + w 2^ex, sqMag); every rebase in double (d = Y,
 * renormalised); and after every K steps of a lane's blocks the mantissa is
 * renormalised when its exponent leaves [-8, 8].  The K-step schedule starts at
 * it0 + 1, where the kernel picks the sample up. */
static void fx_normalise(double yr, double yi, double er, double ei, float* wr, float* wi, int* ex, float* esr,
                         float* esi) {
  double mx = fmax(fabs(yr), fabs(yi));
  if (mx > 0.0) *ex = ilogb(mx);
  *wr = (float)ldexp(yr, -*ex);
  *wi = (float)ldexp(yi, -*ex);
  *esr = (float)ldexp(er, -*ex);
  *esi = (float)ldexp(ei, -*ex);
}

static orc_escape pert_from_fx(const pert_ctx* P, double dr, double di, double er, double ei, int it0, int K) {
  orc_escape esc;
  int N = P->N;
  if (it0 + 1 >= N) {
    esc.iterations = N;
    esc.smoothing = 0.0f;
    return esc;
  }
  /* the first step, FP64 (the start kernel) */
  int m = it0 + 1;
  {
    double zr = P->Zd[2 * m], zi = P->Zd[2 * m + 1];
    double yr = zr + dr, yi = zi + di;
    double r2 = yr * yr + yi * yi;
    double dmag = fma(dr, dr, di * di);
    double wr = zr + yr, wi = zi + yi;
    if (r2 < dmag || m == P->mend) {
      dr = yr;
      di = yi;
      wr = yr;
      wi = yi;
      m = 0;
    }
    double ndr = fma(dr, wr, fma(-di, wi, er));
    double ndi = fma(dr, wi, fma(di, wr, ei));
    dr = ndr;
    di = ndi;
    m++;
  }
  float wr, wi, esr, esi;
  int ex = 0;
  if (fmax(fabs(dr), fabs(di)) == 0.0 && fmax(fabs(er), fabs(ei)) > 0.0) ex = ilogb(fmax(fabs(er), fabs(ei)));
  fx_normalise(dr, di, er, ei, &wr, &wi, &ex, &esr, &esi);
  const double b2 = P->b2;
  const float b2g = (float)(b2 * (1.0 - 1.0 / 1024.0));
  for (int it = it0 + 1;; it++) {
    double zrd = P->Zd[2 * m], zid = P->Zd[2 * m + 1];
    float zr = (float)zrd, zi = (float)zid;
    float fdr = ldexpf(wr, ex), fdi = ldexpf(wi, ex);
    float yr = zr + fdr, yi = zi + fdi;
    float r2 = fmaf(yr, yr, yi * yi);
    if (r2 > b2g) {
      double ydr = zrd + ldexp((double)wr, ex), ydi = zid + ldexp((double)wi, ex);
      double r2d = ydr * ydr + ydi * ydi;
      if (r2d > b2) {
        esc.iterations = it;
        esc.smoothing = (float)smoothing_of(r2d, P->bailout);
        return esc;
      }
    }
    if (it == N - 1) break;
    /* the rebase test in the mantissa's scale: |Z 2^-ex + w|^2 < |w|^2 */
    float ysr = (float)ldexp(zrd, -ex) + wr, ysi = (float)ldexp(zid, -ex) + wi;
    float Wr = zr + yr, Wi = zi + yi;
    if (fmaf(ysr, ysr, ysi * ysi) < fmaf(wr, wr, wi * wi) || m == P->mend) {
      double ydr = zrd + ldexp((double)wr, ex), ydi = zid + ldexp((double)wi, ex);
      fx_normalise(ydr, ydi, er, ei, &wr, &wi, &ex, &esr, &esi);
      Wr = (float)ydr;
      Wi = (float)ydi;
      m = 0;
    }
    float nwr = fmaf(wr, Wr, fmaf(-wi, Wi, esr));
    float nwi = fmaf(wr, Wi, fmaf(wi, Wr, esi));
    wr = nwr;
    wi = nwi;
    m++;
    if ((it - it0) % K == 0) { /* a block ended: renormalise the mantissa */
      float mx = fmaxf(fabsf(wr), fabsf(wi));
      if (mx > 0.0f && mx < INFINITY) {
        int k = ilogbf(mx);
        if (k > 8 || k < -8) {
          wr = ldexpf(wr, -k);
          wi = ldexpf(wi, -k);
          ex += k;
          esr = (float)ldexp(er, -ex);
          esi = (float)ldexp(ei, -ex);
        }
      }
    }
  }
  esc.iterations = N;
  esc.smoothing = 0.0f;
  return esc;
}

static double* make_zd(const orc_view* v) {
  int len = v->len;
  double* Zd = (double*)malloc(sizeof(double) * 2 * (size_t)(len + 2));
  Zd[0] = Zd[1] = 0.0;
  for (int k = 1; k <= len + 1; k++) {
    Zd[2 * k] = v->Xn[2 * (k - 1)];
    Zd[2 * k + 1] = v->Xn[2 * (k - 1) + 1];
  }
  return Zd;
}

/* the GPU method for one sample: reference SA start (scan), SA-escape binary
 * search with double-double truncation, then pert_from.  ua_given >= 0 replaces
 * the sample's own scan by a given first unstable index (a tile's, 0 = none). */
static orc_escape pert_sample(const orc_view* v, const pert_ctx* P, const ohp* y0r, const ohp* y0i, int ua_given,
                              int fx_k) {
  orc_escape esc;
  const int N = v->N;
  const double b2 = v->bailout * v->bailout;
  if (in_cardioid(y0r, y0i, v->n)) {
    esc.iterations = N;
    esc.smoothing = 0.0f;
    return esc;
  }
  lpc eps = pixel_eps(v, y0r, y0i);
  lpc eps2 = lp_sq(eps), eps3 = lp_mul(eps, eps2);
  int ua;
  int dsize;
  if (ua_given < 0) {
    dsize = sa_scan(v, eps, &ua);
  } else {
    ua = ua_given;
    dsize = ua ? (ua - 3 < 1 ? 1 : ua - 3) : v->len;
  }
  int found = dsize - 1;
  lpc d = sa_term(v, found, eps, eps2, eps3);
  lpc y = {trunc_sum3(v->Xn[2 * found], v->Xlo[2 * found], d.re),
           trunc_sum3(v->Xn[2 * found + 1], v->Xlo[2 * found + 1], d.im)};
  if (lp_sqmag(y) > b2) {
    int low = 0, high = dsize - 1, mid = dsize / 2;
    while (low <= high) {
      lpc dm = sa_term(v, mid, eps, eps2, eps3);
      lpc ym = {trunc_sum3(v->Xn[2 * mid], v->Xlo[2 * mid], dm.re),
                trunc_sum3(v->Xn[2 * mid + 1], v->Xlo[2 * mid + 1], dm.im)};
      if (!(lp_sqmag(ym) > b2))
        low = mid + 1;
      else {
        high = mid - 1;
        found = mid;
      }
      mid = (low + high) / 2;
    }
    d = sa_term(v, found, eps, eps2, eps3);
    y.re = trunc_sum3(v->Xn[2 * found], v->Xlo[2 * found], d.re);
    y.im = trunc_sum3(v->Xn[2 * found + 1], v->Xlo[2 * found + 1], d.im);
    esc.iterations = found;
    esc.smoothing = (float)smoothing_of(lp_sqmag(y), v->bailout);
    return esc;
  }
  if (fx_k > 0) return pert_from_fx(P, d.re, d.im, eps.re, eps.im, found, fx_k);
  return pert_from(P, d.re, d.im, eps.re, eps.im, found);
}

/* ------------------------------------------------------------- renders */
typedef struct {
  orc_view* v;
  const int* rc;
  orc_escape* out;
  int mode; /* 0 reference, 1 pert64, 2 pert FP32X */
  pert_ctx P;
  int ncols; /* for full rows */
  const int* rows;
  const int* ua; /* per-sample first unstable index (tile skips), or NULL */
  int fx_k;      /* mode 2: the kernel's block size */
} render_ctx;

static orc_escape sample_value(render_ctx* R, int r, int c, long i) {
  orc_view* v = R->v;
  ohp re, im;
  pixel_coord(v, r, c, &re, &im);
  if (orc_use_hardware(v)) return hw_iterations(&re, &im, v->n, v->N, v->bailout);
  if (R->mode == 0) return sa_iterations(v, &re, &im);
  return pert_sample(v, &R->P, &re, &im, R->ua ? R->ua[i] : -1, R->mode == 2 ? R->fx_k : 0);
}

static void samples_fn(void* cp, long i) {
  render_ctx* R = (render_ctx*)cp;
  R->out[i] = sample_value(R, R->rc[2 * i], R->rc[2 * i + 1], i);
}

static void rows_fn(void* cp, long i) {
  render_ctx* R = (render_ctx*)cp;
  long k = i / R->ncols;
  int c = (int)(i % R->ncols);
  R->out[i] = sample_value(R, R->rows[k], c, i);
}

static void setup_pert(orc_view* v, render_ctx* R) {
  if (v->len > 0) {
    R->P.Zd = make_zd(v);
    R->P.mend = v->len + 1;
  } else {
    R->P.Zd = NULL;
    R->P.mend = 0;
  }
  R->P.N = v->N;
  R->P.b2 = v->bailout * v->bailout;
  R->P.bailout = v->bailout;
}

void orc_render_rows(orc_view* v, const int* rows, int nrows, orc_escape* out, int nthreads) {
  render_ctx R = {v, NULL, out, 0, {0}, 0, NULL, NULL, 0};
  R.ncols = v->nc;
  R.rows = rows;
  setup_pert(v, &R);
  parallel_for((long)nrows * v->nc, nthreads, 8, rows_fn, &R);
  free((void*)R.P.Zd);
}

void orc_render_samples(orc_view* v, const int* rc, int n, orc_escape* out, int nthreads) {
  render_ctx R = {v, rc, out, 0, {0}, 0, NULL, NULL, 0};
  setup_pert(v, &R);
  parallel_for(n, nthreads, 4, samples_fn, &R);
  free((void*)R.P.Zd);
}

void orc_pert_samples(orc_view* v, const int* rc, int n, orc_escape* out, int nthreads) {
  render_ctx R = {v, rc, out, 1, {0}, 0, NULL, NULL, 0};
  setup_pert(v, &R);
  parallel_for(n, nthreads, 4, samples_fn, &R);
  free((void*)R.P.Zd);
}

/* the GPU method with FP32X deltas (block size K of the kernel); ua as in
 * orc_pert_samples_ua, or NULL for each sample's own scan */
void orc_pert_samples_fx(orc_view* v, const int* rc, const int* ua, int n, int K, orc_escape* out, int nthreads) {
  render_ctx R = {v, rc, out, 2, {0}, 0, NULL, ua, K};
  setup_pert(v, &R);
  parallel_for(n, nthreads, 4, samples_fn, &R);
  free((void*)R.P.Zd);
}

/* the GPU method with a given first unstable index per sample (tile skips) */
void orc_pert_samples_ua(orc_view* v, const int* rc, const int* ua, int n, orc_escape* out, int nthreads) {
  render_ctx R = {v, rc, out, 1, {0}, 0, NULL, ua, 0};
  setup_pert(v, &R);
  parallel_for(n, nthreads, 4, samples_fn, &R);
  free((void*)R.P.Zd);
}

/* Per T x T tile (row-major tiles): the probe is the tile's sample farthest from
 * the reference point, a corner (epsilon is separable and monotone along rows and
 * columns; a tie keeps the first column / row); its first unstable index by the
 * reference's scan (mandelbrot.cpp:165-181), 0 = none.  probe_rc gets (r, c). */
int orc_tile_probe_ua(orc_view* v, int T, int* ua, int* probe_rc) {
  if (T < 1 || v->len < 1) return -1;
  int ntc = (v->nc + T - 1) / T, ntr = (v->nr + T - 1) / T;
  for (int tr = 0; tr < ntr; tr++)
    for (int tc = 0; tc < ntc; tc++) {
      int c0 = tc * T, c1 = (c0 + T < v->nc ? c0 + T : v->nc) - 1;
      int r0 = tr * T, r1 = (r0 + T < v->nr ? r0 + T : v->nr) - 1;
      double e0[2], e1[2], f0[2], f1[2];
      orc_pixel_eps(v, r0, c0, e0);
      orc_pixel_eps(v, r0, c1, e1);
      orc_pixel_eps(v, r0, c0, f0);
      orc_pixel_eps(v, r1, c0, f1);
      int pc = fabs(e1[0]) > fabs(e0[0]) ? c1 : c0;
      int pr = fabs(f1[1]) > fabs(f0[1]) ? r1 : r0;
      double e[2];
      orc_pixel_eps(v, pr, pc, e);
      lpc eps = {e[0], e[1]};
      int u = 0;
      sa_scan(v, eps, &u);
      ua[tr * ntc + tc] = u;
      if (probe_rc) {
        probe_rc[2 * (tr * ntc + tc)] = pr;
        probe_rc[2 * (tr * ntc + tc) + 1] = pc;
      }
    }
  return ntc * ntr;
}

/* ------------------------------------------------------- multisampling */
/* mandelbrot.cpp:320-332: float sum of iterations + smoothing, r1 outer, c1
 * inner; divided by sc*sc; split into (int) part and fraction */
void orc_downsample(const orc_escape* fine, int nr, int nc, int sc, orc_escape* out) {
  int onr = nr / sc, onc = nc / sc;
  for (int r = 0; r < onr; r++)
    for (int c = 0; c < onc; c++) {
      float sum = 0.0f;
      for (int r1 = 0; r1 < sc; r1++)
        for (int c1 = 0; c1 < sc; c1++) {
          const orc_escape* e = &fine[(size_t)(sc * r + r1) * nc + (sc * c + c1)];
          sum += e->iterations + e->smoothing;
        }
      sum /= sc * sc;
      orc_escape o;
      o.iterations = (int)sum;
      o.smoothing = sum - o.iterations;
      out[(size_t)r * onc + c] = o;
    }
}

/* mandelbrot.cpp:334-346 */
void orc_upsample(const orc_escape* coarse, int nr, int nc, int sc, orc_escape* out) {
  int fnr = nr * sc, fnc = nc * sc;
  for (int r = 0; r < fnr; r++)
    for (int c = 0; c < fnc; c++) out[(size_t)r * fnc + c] = coarse[(size_t)(r / sc) * nc + c / sc];
}

/* ------------------------------------------------------- HP op probe */
int orc_hp_op(int op, const char* a, const char* b, long iv, double dv, long bits, int* sign, long* exp,
              uint32_t* limbs) {
  int n = ohp_limbs_for_bits(bits);
  ohp x, y, r;
  x.n = y.n = r.n = n;
  ohp_zero(&x, n);
  ohp_zero(&y, n);
  ohp_zero(&r, n);
  if (a && ohp_set_str(&x, a)) return -1;
  if (b && ohp_set_str(&y, b)) return -1;
  switch (op) {
    case 0: ohp_add(&r, &x, &y); break;
    case 1: ohp_sub(&r, &x, &y); break;
    case 2: ohp_mul(&r, &x, &y); break;
    case 3: ohp_mul_si(&r, &x, iv); break;
    case 4: ohp_div_ui(&r, &x, (unsigned long)iv); break;
    case 5: r = x; break;
    case 6: ohp_mul_d(&r, &x, dv); break;
    default: return -1;
  }
  *sign = r.sign;
  *exp = r.exp;
  for (int i = 0; i < n; i++) limbs[i] = r.d[i];
  return n;
}

/* decimal output "[-]0.DDDDe<E>" with digits rounded away from zero (enough
 * digits that parsing back at the same precision restores the value) */
static void hp_to_string(const ohp* x, char* out, int cap) {
  if (x->sign == 0) {
    snprintf(out, (size_t)cap, "0");
    return;
  }
  /* value = M * 2^(exp - 32 n), M integer of n limbs */
  int n = x->n;
  int digits = (int)ceil(32.0 * n * 0.30102999566398120) + 3;
  long e2 = x->exp - 32L * n;
  long log10v = (long)floor((double)x->exp * 0.30102999566398120);
  long k = digits - log10v;
  /* I = ceil(M * 10^k * 2^e2) using little-endian 32-bit bigi */
  int capw = n + (int)((labs(k) * 4 + labs(e2)) / 32) + 8;
  uint32_t* w = (uint32_t*)calloc((size_t)capw, sizeof(uint32_t));
  int wn = n;
  for (int i = 0; i < n; i++) w[i] = x->d[n - 1 - i];
  int inexact = 0;
  /* multiply by 10^k (k >= 0) */
  for (long kk = k; kk > 0; kk--) {
    uint64_t c = 0;
    for (int i = 0; i < wn; i++) {
      c += (uint64_t)w[i] * 10u;
      w[i] = (uint32_t)c;
      c >>= 32;
    }
    if (c) w[wn++] = (uint32_t)c;
  }
  /* shift by e2 */
  if (e2 > 0) {
    long ls = e2 / 32;
    int bs = (int)(e2 % 32);
    for (int i = wn + (int)ls; i >= 0; i--) {
      long src = i - ls;
      uint32_t hi = (src >= 0 && src < wn) ? w[src] : 0;
      uint32_t lo = (src - 1 >= 0 && src - 1 < wn) ? w[src - 1] : 0;
      w[i] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
    }
    wn += (int)ls + 1;
  } else if (e2 < 0) {
    long sh = -e2;
    long ls = sh / 32;
    int bs = (int)(sh % 32);
    for (long i = 0; i < ls && i < wn; i++)
      if (w[i]) inexact = 1;
    if (bs && ls < wn && (w[ls] << (32 - bs))) inexact = 1;
    for (int i = 0; i < wn; i++) {
      long src = i + ls;
      uint32_t lo = src < wn ? w[src] : 0;
      uint32_t hi = src + 1 < wn ? w[src + 1] : 0;
      w[i] = bs ? (lo >> bs) | (hi << (32 - bs)) : lo;
    }
    wn = wn - (int)ls;
    if (wn < 0) wn = 0;
  }
  /* divide by 10^-k (k < 0) */
  for (long kk = k; kk < 0; kk++) {
    uint64_t rem = 0;
    for (int i = wn - 1; i >= 0; i--) {
      uint64_t cur = (rem << 32) | w[i];
      w[i] = (uint32_t)(cur / 10u);
      rem = cur % 10u;
    }
    if (rem) inexact = 1;
  }
  while (wn > 0 && w[wn - 1] == 0) wn--;
  if (inexact) {
    uint64_t c = 1;
    for (int i = 0; i < wn && c; i++) {
      c += w[i];
      w[i] = (uint32_t)c;
      c >>= 32;
    }
    if (c) w[wn++] = (uint32_t)c;
  }
  /* decimal digits */
  char* ds = (char*)malloc((size_t)(wn * 10 + 16));
  int nd = 0;
  while (wn > 0) {
    uint64_t rem = 0;
    for (int i = wn - 1; i >= 0; i--) {
      uint64_t cur = (rem << 32) | w[i];
      w[i] = (uint32_t)(cur / 10u);
      rem = cur % 10u;
    }
    ds[nd++] = (char)('0' + rem);
    while (wn > 0 && w[wn - 1] == 0) wn--;
  }
  /* ds holds digits least significant first */
  int lead = nd;
  int trail = 0;
  while (trail < nd - 1 && ds[trail] == '0') trail++;
  int pos = 0;
  if (x->sign < 0) pos += snprintf(out + pos, (size_t)(cap - pos), "-");
  pos += snprintf(out + pos, (size_t)(cap - pos), "0.");
  for (int i = nd - 1; i >= trail && pos < cap - 24; i--) out[pos++] = ds[i];
  snprintf(out + pos, (size_t)(cap - pos), "e%ld", (long)lead - k);
  free(ds);
  free(w);
}

/* --------------------------------------------------------------- zoom walk */
/* Seeded zoom walk that fixes the deep-zoom configurations' centres: start from
 * the reference's full view (centre -0.5+0i, width 4, mandelbrot.cpp:13-14);
 * repeatedly render res x res at the current centre/width, pick a uniformly
 * random escaped sample with count >= 0.9 N (if none, >= 0.9 of the largest
 * escaped count) among those on a band edge (some 4-neighbour has a different
 * count; all candidates if none is), move the centre there and divide the width
 * by 10; stop when the next width would fall below the target, and use the
 * target width.
 * Samples are rendered with pert64 against the orbit of the centre itself. */
static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct {
  pert_ctx P;
  const double* eps_re;
  const double* eps_im;
  int res;
  int32_t* counts;
  const double* cre_d; /* doubles of the coordinates, for the cheap cardioid skip */
  const double* cim_d;
} walk_ctx;

static int cardioid_d(double x, double y) {
  double xmf = x - 0.25, y2 = y * y, q = xmf * xmf + y2;
  if (q * (q + xmf) < 0.25 * y2 * (1.0 - 1e-9)) return 1;
  q = x + 1.0;
  return q * q + y2 < 0.0625 * (1.0 - 1e-9);
}

static void walk_fn(void* cp, long i) {
  walk_ctx* W = (walk_ctx*)cp;
  int r = (int)(i / W->res), c = (int)(i % W->res);
  if (cardioid_d(W->cre_d[c], W->cim_d[r])) {
    W->counts[i] = W->P.N;
    return;
  }
  orc_escape e = pert_from(&W->P, W->eps_re[c], W->eps_im[r], W->eps_re[c], W->eps_im[r], 0);
  W->counts[i] = e.iterations;
}

static void center_render_hp(const ohp* cre, const ohp* cim, const ohp* sz, int n, int res, int N, int nthreads,
                             int32_t* counts) {
  /* reference orbit of the centre, plus one extra element */
  orc_view tmp;
  memset(&tmp, 0, sizeof tmp);
  tmp.n = n;
  tmp.N = N;
  tmp.bailout = 1024.0;
  int len = orbit_length(cre, cim, n, N, 1024.0 * 1024.0);
  orbit_store(&tmp, cre, cim, len);
  store_doubles(&tmp);
  double* Zd = make_zd(&tmp);
  double* er = (double*)malloc(sizeof(double) * (size_t)res);
  double* ei = (double*)malloc(sizeof(double) * (size_t)res);
  double* xr = (double*)malloc(sizeof(double) * (size_t)res);
  double* xi = (double*)malloc(sizeof(double) * (size_t)res);
  ohp t, u;
  t.n = u.n = n;
  for (int k = 0; k < res; k++) {
    ohp_mul_si(&t, sz, (long)(k - res / 2));
    er[k] = ohp_get_d(&t);
    ohp_add(&u, cre, &t);
    xr[k] = ohp_get_d(&u);
    ohp_mul_si(&t, sz, (long)(res / 2 - k - 1));
    ei[k] = ohp_get_d(&t);
    ohp_add(&u, cim, &t);
    xi[k] = ohp_get_d(&u);
  }
  walk_ctx W;
  W.P.Zd = Zd;
  W.P.mend = len + 1;
  W.P.N = N;
  W.P.b2 = 1024.0 * 1024.0;
  W.P.bailout = 1024.0;
  W.eps_re = er;
  W.eps_im = ei;
  W.res = res;
  W.counts = counts;
  W.cre_d = xr;
  W.cim_d = xi;
  parallel_for((long)res * res, nthreads, 16, walk_fn, &W);
  free(er); free(ei); free(xr); free(xi);
  free(Zd);
  free(tmp.X); free(tmp.Xt); free(tmp.Xn); free(tmp.Xlo);
}

void orc_center_render(const char* cre_s, const char* cim_s, const char* width, int res, int N, int nthreads,
                       int32_t* counts) {
  ohp w;
  w.n = ohp_limbs_for_bits(256);
  ohp_set_str(&w, width);
  ohp sz;
  sz.n = w.n;
  ohp_div_ui(&sz, &w, (unsigned long)res);
  long e;
  ohp_get_d_2exp(&e, &sz);
  int bits = (int)(64 - e + log2(1.0e-20));
  if (bits < 64) bits = 64;
  int n = ohp_limbs_for_bits(bits);
  ohp cre, cim, szn, wn;
  cre.n = cim.n = szn.n = wn.n = n;
  ohp_set_str(&cre, cre_s);
  ohp_set_str(&cim, cim_s);
  ohp_set_str(&wn, width);
  ohp_div_ui(&szn, &wn, (unsigned long)res);
  center_render_hp(&cre, &cim, &szn, n, res, N, nthreads, counts);
}

int orc_zoom_walk(uint64_t seed, int N, const char* target_width, int res, int nthreads, char* out_re, char* out_im,
                  int cap, int* steps) {
  ohp target;
  target.n = ohp_limbs_for_bits(64);
  if (ohp_set_str(&target, target_width)) return -1;
  long te;
  ohp_get_d_2exp(&te, &target);
  /* precision for the deepest step */
  int nmax = ohp_limbs_for_bits((long)(64 - (te - 9) + log2(1.0e-20)) + 64);
  ohp cre, cim, w, sz, t;
  cre.n = cim.n = w.n = sz.n = t.n = nmax;
  ohp_set_d(&cre, -0.5);
  ohp_set_d(&cim, 0.0);
  ohp_set_d(&w, 4.0);
  ohp tgt;
  tgt.n = nmax;
  ohp_set_str(&tgt, target_width);
  uint64_t state = seed;
  int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)res * res);
  int k = 0, retries = 0;
  ohp pre, pim, pw; /* the previous step's centre and width (back-off) */
  pre.n = pim.n = pw.n = nmax;
  for (;;) {
    ohp w10;
    w10.n = nmax;
    ohp_div_ui(&w10, &w, 10);
    if (ohp_cmp(&w10, &tgt) < 0) break;
    ohp_div_ui(&sz, &w, (unsigned long)res);
    long e;
    ohp_get_d_2exp(&e, &sz);
    int bits = (int)(64 - e + log2(1.0e-20));
    if (bits < 64) bits = 64;
    int n = ohp_limbs_for_bits(bits);
    ohp cr, ci, s;
    cr.n = ci.n = s.n = n;
    ohp_copy(&cr, &cre);
    ohp_copy(&ci, &cim);
    ohp_copy(&s, &sz);
    center_render_hp(&cr, &ci, &s, n, res, N, nthreads, counts);
    int thr = (int)ceil(0.9 * N);
    long ncand = 0;
    int maxesc = -1;
    for (long i = 0; i < (long)res * res; i++) {
      if (counts[i] < N && counts[i] >= thr) ncand++;
      if (counts[i] < N && counts[i] > maxesc) maxesc = counts[i];
    }
    if (ncand == 0) {
      if (maxesc < 0) {
        /* nothing escaped at this depth: back off one step and draw another
         * candidate there (the draws continue the same splitmix64 stream) */
        if (k == 0 || retries >= 32) break;
        ohp_copy(&cre, &pre);
        ohp_copy(&cim, &pim);
        ohp_copy(&w, &pw);
        k--;
        retries++;
        continue;
      }
      thr = (int)ceil(0.9 * maxesc);
    }
    /* candidates on a band edge (a 4-neighbour has another count) keep the walk
     * on structure; a 10x zoom into the inside of a band is flat forever after */
    long nedge = 0, nset = 0;
    unsigned char* edge = (unsigned char*)calloc((size_t)res * res, 1);
    for (long i = 0; i < (long)res * res; i++) {
      if (!(counts[i] < N && counts[i] >= thr)) continue;
      int r = (int)(i / res), c = (int)(i % res);
      int e = (r > 0 && counts[i - res] != counts[i]) || (r < res - 1 && counts[i + res] != counts[i]) ||
              (c > 0 && counts[i - 1] != counts[i]) || (c < res - 1 && counts[i + 1] != counts[i]);
      int s = (r > 0 && counts[i - res] == N) || (r < res - 1 && counts[i + res] == N) ||
              (c > 0 && counts[i - 1] == N) || (c < res - 1 && counts[i + 1] == N);
      edge[i] = (unsigned char)(e + s); /* 2: touches a sample that did not escape */
      nedge += e;
      nset += s;
    }
    /* level 2 (next to non-escaped samples: the set's boundary) over level 1
     * (band edge) over level 0 (any candidate) */
    int use_level = nset > 0 ? 2 : (nedge > 0 ? 1 : 0);
    for (long i = 0; i < (long)res * res; i++) edge[i] = edge[i] >= use_level;
    int use_edge = 1;
    ncand = 0;
    for (long i = 0; i < (long)res * res; i++)
      if (counts[i] < N && counts[i] >= thr && (!use_edge || edge[i])) ncand++;
    uint64_t rnd = splitmix64(&state);
    long pick = (long)(((unsigned __int128)rnd * (unsigned __int128)ncand) >> 64);
    long sel = -1;
    for (long i = 0; i < (long)res * res; i++)
      if (counts[i] < N && counts[i] >= thr && (!use_edge || edge[i])) {
        if (pick == 0) {
          sel = i;
          break;
        }
        pick--;
      }
    free(edge);
    int r = (int)(sel / res), c = (int)(sel % res);
    ohp_copy(&pre, &cre);
    ohp_copy(&pim, &cim);
    ohp_copy(&pw, &w);
    ohp_mul_si(&t, &sz, (long)(c - res / 2));
    ohp_add(&cre, &cre, &t);
    ohp_mul_si(&t, &sz, (long)(res / 2 - r - 1));
    ohp_add(&cim, &cim, &t);
    w = w10;
    k++;
  }
  free(counts);
  hp_to_string(&cre, out_re, cap);
  hp_to_string(&cim, out_im, cap);
  if (steps) *steps = k;
  return 0;
}
