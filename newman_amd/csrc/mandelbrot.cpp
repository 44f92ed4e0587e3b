// mandelbrot.cpp - host side of the drop-in render core (see mandelbrot.hpp).
// Each method cites the reference method it replaces.
#include "mandelbrot.hpp"

#include <algorithm>
#include <mutex>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <thread>

namespace nm {

static constexpr double kMinPreview = 1.5e-16;  // mandelbrot.cpp:257

// ------------------------------------------------------------------ threading
template <class F>
static void parallel_for(long total, F&& fn, long chunk = 4) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 64) nt = 64;
  if (total <= chunk || nt == 1) {
    for (long i = 0; i < total; i++) fn(i);
    return;
  }
  std::atomic<long> next{0};
  auto worker = [&]() {
    for (;;) {
      long s = next.fetch_add(chunk);
      if (s >= total) break;
      long e = std::min(total, s + chunk);
      for (long i = s; i < e; i++) fn(i);
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; t++) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
}

// ------------------------------------------------------------------ helpers
static inline void times2(hpf& x) {
  if (x.sign) x.exp += 1;  // 2.0 * x is exact
}

// nearest-rounded double of an HP value (the perturbation kernels' reference
// orbit), built on the truncating conversion
static double get_d_nearest(const hpf& a) {
  double t = a.get_d();
  if (a.sign == 0 || !std::isfinite(t)) return t;
  hpf td(2), r(a.L);
  hp_set_d(td, t);
  hp_sub(r, a, td);  // remainder, same sign as a or zero
  if (r.sign == 0) return t;
  double up = std::nextafter(t, a.sign > 0 ? INFINITY : -INFINITY);
  hpf half(2);
  hp_set_d(half, (up - t) * 0.5);
  r.sign = 1;
  half.sign = 1;
  int c = hp_cmp(r, half);
  if (c > 0) return up;
  if (c < 0) return t;
  uint64_t bits;
  std::memcpy(&bits, &t, 8);
  return (bits & 1) ? up : t;
}

static inline bool bailedOut(const hpf& re, const hpf& im, double b2) {
  LPComplex z(re.get_d(), im.get_d());
  return sqMag(z) > b2;  // mandelbrot.cpp:60
}

// mandelbrot.cpp:63-71, in HP at the working precision
bool inCardioidHP(const hpf& re, const hpf& im) {
  int L = std::max(re.L, im.L);
  hpf fourth(L), xmf(L), y2(L), q(L), t(L), lhs(L), rhs(L);
  hp_set_d(fourth, 0.25);
  hp_sub(xmf, re, fourth);
  hp_mul(y2, im, im);
  hp_mul(q, xmf, xmf);
  hp_add(q, q, y2);
  hp_add(t, q, xmf);
  hp_mul(lhs, q, t);
  hp_mul(rhs, fourth, y2);
  if (hp_cmp(lhs, rhs) < 0) return true;  // cardioid
  hpf one(2);
  hp_set_d(one, 1.0);
  hp_add(q, re, one);
  hp_mul(lhs, q, q);
  hp_add(lhs, lhs, y2);
  hp_mul(rhs, fourth, fourth);
  return hp_cmp(lhs, rhs) < 0;  // second disk
}

// The pixel map of computeRow / findProbe / pointAt (mandelbrot.cpp:271, 275,
// 86-87, 287-288): re = center.re + k*sz.re, im = center.im + j*sz.im, each step
// truncated at the working precision.
static void pixel_re(hpf& out, const hpf& cre, const hpf& szre, long k) {
  hpf t(out.L);
  hp_mul_si(t, szre, k);
  hp_add(out, cre, t);
}

// ------------------------------------------------------------------ class
Mandelbrot::Mandelbrot() : Mandelbrot(1, 1) {}

// mandelbrot.cpp:8-17
Mandelbrot::Mandelbrot(int nr, int nc) : grid(nr, nc) {
  error_tolerance = 1e-10;
  N = 256;
  center.re = -0.5;
  center.im = 0.0;
  sz.re = 4.0 / nc;
  sz.im = 3.0 / nr;
  setPrecision();
}

// mandelbrot.cpp:37-56
void Mandelbrot::setPrecision() {
  const double log_alpha = std::log2(1.0e-20);
  const int beta = 64;
  long e;
  sz.re.get_d_2exp(&e);
  int bits = (int)(beta - e + log_alpha);
  if (bits < 64) bits = 64;
  bits_ = bits;
  hp_set_default_bits(bits);  // process-wide, as mpf_set_default_prec
  center.re.set_prec_bits(bits);
  center.im.set_prec_bits(bits);
  sz.re.set_prec_bits(bits);
  sz.im.set_prec_bits(bits);
  X.clear();
  ref_ = SaReferenceHost();
  have_ref_ = false;
  rendered_ = false;
}

// mandelbrot.cpp:19-35: values parsed at the current precision, then sz rescaled
// by 800/cols, 600/rows and setPrecision.
int Mandelbrot::loadLegacy(const char* fn) {
  FILE* fp = std::fopen(fn, "r");
  if (!fp) return NM_ERR_IO;
  char buf[4096];
  int n;
  if (std::fscanf(fp, "%d\n", &n) != 1) {
    std::fclose(fp);
    return NM_ERR_IO;
  }
  N = n;
  if (std::fscanf(fp, "%4095s\n", buf) == 1) center.re = buf;
  if (std::fscanf(fp, "%4095s\n", buf) == 1) center.im = buf;
  if (std::fscanf(fp, "%4095s\n", buf) == 1) sz.re = buf;
  if (std::fscanf(fp, "%4095s\n", buf) == 1) sz.im = buf;
  std::fclose(fp);
  sz.re = sz.re * (800.0 / cols());
  sz.im = sz.im * (600.0 / rows());
  setPrecision();
  return NM_OK;
}

// Parse a view at the precision the target width needs: N, centre strings and
// width = cols * sz.re (square samples).  A harness convenience, not a reference
// method: loadLegacy would parse the centre at the previous precision.
int Mandelbrot::setView(int n, const char* cre, const char* cim, const char* width) {
  if (n <= 0 || !cre || !cim || !width) return NM_ERR_ARG;
  hpf w(hp_limbs_for_bits(256));
  if (!hp_set_str(w, width) || w.sign <= 0) return NM_ERR_ARG;
  hpf s(w.L);
  hp_div_ui(s, w, (unsigned long)cols());
  long e;
  s.get_d_2exp(&e);
  int bits = (int)(64 - e + std::log2(1.0e-20));
  if (bits < 64) bits = 64;
  int L = hp_limbs_for_bits(bits);
  hpf r(L), i(L), wl(L), szl(L);
  if (!hp_set_str(r, cre) || !hp_set_str(i, cim)) return NM_ERR_ARG;
  hp_set_str(wl, width);
  hp_div_ui(szl, wl, (unsigned long)cols());
  N = n;
  hp_copy_prec(center.re, r);
  hp_copy_prec(center.im, i);
  hp_copy_prec(sz.re, szl);
  hp_copy_prec(sz.im, szl);
  setPrecision();
  return NM_OK;
}

// mandelbrot.cpp:256-259
bool Mandelbrot::useHardware() { return sz.re.get_d() >= kMinPreview && sz.im.get_d() >= kMinPreview; }

// mandelbrot.cpp:97-110.  X[0] = X0; X[i] = X[i-1]^2 + X0, truncated at the first
// i with bailedOut(X[i]) (size i) or at N.  With store, X keeps len+1 entries:
// the extra X[len] (the escaping element, or one more step) feeds the device
// orbit so the perturbation kernel can rebase at the reference's end.
int Mandelbrot::computeOrbit(const HPComplex& X0, bool store) {
  const int L = hp_limbs_for_bits(hp_default_bits());
  const double b2 = bailout * bailout;
  hpf xr(L), xi(L), nr(L), ni(L), t1(L), t2(L);
  hp_set(xr, X0.re);
  hp_set(xi, X0.im);
  if (store) {
    X.clear();
    X.reserve((size_t)N + 1);
    HPComplex z;
    z.re = hpf(L); z.im = hpf(L);
    hp_set(z.re, xr);
    hp_set(z.im, xi);
    X.push_back(z);
  }
  int len = N;
  for (int i = 1; i <= N; i++) {
    if (!store && i == N) break;
    hp_mul(t1, xr, xr);
    hp_mul(t2, xi, xi);
    hp_sub(t1, t1, t2);
    hp_add(nr, t1, X0.re);
    hp_mul(t1, xr, xi);
    times2(t1);
    hp_add(ni, t1, X0.im);
    if (store) {
      HPComplex z;
      z.re = hpf(L); z.im = hpf(L);
      hp_set(z.re, nr);
      hp_set(z.im, ni);
      X.push_back(z);
    }
    if (i == N) break;  // the extra element past a full-length orbit
    if (bailedOut(nr, ni, b2)) {
      len = i;
      break;
    }
    hp_copy_prec(xr, nr);
    hp_copy_prec(xi, ni);
  }
  return len;
}

// The probe list of findProbe (mandelbrot.cpp:77-94), in its order.
static std::vector<std::pair<int, int>> probe_points(int nr, int nc) {
  std::vector<std::pair<int, int>> pts;
  for (int c = 0; c < nc; c += 2) {
    pts.emplace_back(nr / 4, c);
    pts.emplace_back(nr / 2, c);
    pts.emplace_back(3 * nr / 4, c);
  }
  for (int r = 0; r < nr; r += 2) pts.emplace_back(r, nc / 2);
  return pts;
}

// mandelbrot.cpp:73-95.  The probe orbits are independent: their lengths are
// computed in parallel and the first longest one is kept, as the serial
// swap-if-longer loop does; its orbit is then recomputed and stored.
void Mandelbrot::findProbe() {
  const auto pts = probe_points(rows(), cols());
  const int L = hp_limbs_for_bits(hp_default_bits());
  std::vector<int> lens(pts.size(), 0);
  parallel_for((long)pts.size(), [&](long i) {
    HPComplex p;
    p.re = hpf(L); p.im = hpf(L);
    pixel_re(p.re, center.re, sz.re, (long)(pts[i].second - cols() / 2));
    pixel_re(p.im, center.im, sz.im, (long)(rows() / 2 - pts[i].first - 1));
    Mandelbrot* self = this;
    // length only: a private orbit loop (computeOrbit's non-storing path is
    // reentrant; it touches no member state)
    lens[i] = self->computeOrbit(p, false);
  }, 2);
  int best = -1, bestlen = 0;
  for (size_t i = 0; i < pts.size(); i++)
    if (lens[i] > bestlen) {
      bestlen = lens[i];
      best = (int)i;
    }
  winner_ = best;
  probe_candidates_ = (int)pts.size();
  HPComplex p;
  p.re = hpf(L); p.im = hpf(L);
  pixel_re(p.re, center.re, sz.re, (long)(pts[best].second - cols() / 2));
  pixel_re(p.im, center.im, sz.im, (long)(rows() / 2 - pts[best].first - 1));
  int len = computeOrbit(p, true);
  X.resize((size_t)len + 1);
  ref_.len = len;
}

// mandelbrot.cpp:112-131, the recurrence exactly as written (A[i], B[i] on the
// right-hand side of B[i], C[i]).  Only the truncated doubles the per-pixel work
// reads are kept.
void Mandelbrot::computeSeries() {
  const int len = ref_.len;
  const int L = hp_limbs_for_bits(hp_default_bits());
  ref_.A.assign(2 * (size_t)len, 0.0);
  ref_.B.assign(2 * (size_t)len, 0.0);
  ref_.C.assign(2 * (size_t)len, 0.0);
  ref_.Am.assign(2 * (size_t)len, 0.0);
  ref_.Bm.assign(2 * (size_t)len, 0.0);
  ref_.Cm.assign(2 * (size_t)len, 0.0);
  ref_.Ae.assign(2 * (size_t)len, kFxZeroExp);
  ref_.Be.assign(2 * (size_t)len, kFxZeroExp);
  ref_.Ce.assign(2 * (size_t)len, kFxZeroExp);
  ref_.Am[0] = 0.5;
  ref_.Ae[0] = 1;
  hpf Ar(L), Ai(L), Br(L), Bi(L), Cr(L), Ci(L), nAr(L), nAi(L), nBr(L), nBi(L), nCr(L), nCi(L), t(L);
  hpf one(2);
  hp_set_d(one, 1.0);
  hp_set_d(Ar, 1.0);
  ref_.A[0] = 1.0;
  for (int i = 1; i < len; i++) {
    const hpf& xr = X[i - 1].re;
    const hpf& xi = X[i - 1].im;
    hp_mul(nAr, xr, Ar); hp_mul(t, xi, Ai); hp_sub(nAr, nAr, t); times2(nAr); hp_add(nAr, nAr, one);
    hp_mul(nAi, xr, Ai); hp_mul(t, xi, Ar); hp_add(nAi, nAi, t); times2(nAi);
    hp_mul(nBr, xr, Br); hp_mul(t, xi, Bi); hp_sub(nBr, nBr, t); times2(nBr);
    hp_mul(t, nAr, nAr); hp_add(nBr, nBr, t);
    hp_mul(t, nAi, nAi); hp_sub(nBr, nBr, t);
    hp_mul(nBi, xr, Bi); hp_mul(t, xi, Br); hp_add(nBi, nBi, t);
    hp_mul(t, nAr, nAi); hp_add(nBi, nBi, t); times2(nBi);
    hp_mul(nCr, xr, Cr); hp_mul(t, xi, Ci); hp_sub(nCr, nCr, t);
    hp_mul(t, nAr, nBr); hp_add(nCr, nCr, t);
    hp_mul(t, nAi, nBi); hp_sub(nCr, nCr, t); times2(nCr);
    hp_mul(nCi, xr, Ci); hp_mul(t, xi, Cr); hp_add(nCi, nCi, t);
    hp_mul(t, nAr, nBi); hp_add(nCi, nCi, t);
    hp_mul(t, nAi, nBr); hp_add(nCi, nCi, t); times2(nCi);
    hp_copy_prec(Ar, nAr); hp_copy_prec(Ai, nAi);
    hp_copy_prec(Br, nBr); hp_copy_prec(Bi, nBi);
    hp_copy_prec(Cr, nCr); hp_copy_prec(Ci, nCi);
    ref_.A[2 * i] = Ar.get_d(); ref_.A[2 * i + 1] = Ai.get_d();
    ref_.B[2 * i] = Br.get_d(); ref_.B[2 * i + 1] = Bi.get_d();
    ref_.C[2 * i] = Cr.get_d(); ref_.C[2 * i + 1] = Ci.get_d();
    const hpf* hv[6] = {&Ar, &Ai, &Br, &Bi, &Cr, &Ci};
    std::vector<double>* fm[3] = {&ref_.Am, &ref_.Bm, &ref_.Cm};
    std::vector<int32_t>* fe[3] = {&ref_.Ae, &ref_.Be, &ref_.Ce};
    for (int k = 0; k < 6; k++) {
      long e = 0;
      const double m = hv[k]->get_d_2exp(&e);
      (*fm[k / 2])[2 * i + k % 2] = m;
      (*fe[k / 2])[2 * i + k % 2] = m == 0.0 ? kFxZeroExp : (int32_t)e;
    }
  }
}

// log2 |z|^2 without overflow (long double has a 15-bit exponent)
static double log2_sqmag(double re, double im) {
  long double r = re, i = im;
  long double m = r * r + i * i;
  if (m == 0) return -INFINITY;
  return (double)log2l(m);
}

// log2 |z|^2 of a floatexp pair (mantissas mr, mi with exponents er, ei)
static double log2_sqmag_fx(double mr, int er, double mi, int ei) {
  if (mr == 0.0 && mi == 0.0) return -INFINITY;
  const int E = mr == 0.0 ? ei : mi == 0.0 ? er : std::max(er, ei);
  const double r = std::ldexp(mr, std::max(er - E, -2100)), i = std::ldexp(mi, std::max(ei - E, -2100));
  return std::log2(r * r + i * i) + 2.0 * E;
}

// Tables for the O(log N) per-sample stop-index search that reproduces the
// reference's SA scan (mandelbrot.cpp:165-181).  The stop test isUnstable
// (mandelbrot.cpp:138-142), |B e^2|^2 tol < |C e^3|^2, is |e|^2 > T_i with
// T_i = tol |B_i|^2 / |C_i|^2 up to rounding; the first unstable i is the first
// index whose prefix-minimum of T drops below |e|^2.  The kernels resolve the
// rounding band around each threshold with the exact double test, and use the
// magnitude prefixes to prove that no intermediate of the exact test under- or
// overflows (otherwise they fall back to the reference's linear scan).
void Mandelbrot::buildThresholds() {
  const int len = ref_.len;
  const double ltol = std::log2(error_tolerance);
  ref_.fxs = series_mode == SERIES_FLOATEXP ? 1 : 0;
  if (ref_.fxs) {
    buildThresholdsFx();
    return;
  }
  ref_.PM.assign(len, INFINITY);
  ref_.BX.assign(len, -INFINITY);
  ref_.BN.assign(len, INFINITY);
  ref_.CX.assign(len, -INFINITY);
  ref_.CN.assign(len, INFINITY);
  ref_.SP.assign(len, 0);
  double pm = INFINITY, bx = -INFINITY, bn = INFINITY, cx = -INFINITY, cn = INFINITY;
  int sp = 0;
  for (int i = 1; i < len; i++) {
    double br = ref_.B[2 * i], bi = ref_.B[2 * i + 1], cr = ref_.C[2 * i], ci = ref_.C[2 * i + 1];
    bool special = !std::isfinite(br) || !std::isfinite(bi) || !std::isfinite(cr) || !std::isfinite(ci) ||
                   (br == 0.0 && bi == 0.0) || (cr == 0.0 && ci == 0.0);
    double lb = special ? 0.0 : log2_sqmag(br, bi);
    double lc = special ? 0.0 : log2_sqmag(cr, ci);
    if (special) sp++;
    else {
      double T = ltol + lb - lc;
      pm = std::min(pm, T);
      bx = std::max(bx, lb); bn = std::min(bn, lb);
      cx = std::max(cx, lc); cn = std::min(cn, lc);
    }
    ref_.PM[i] = pm; ref_.BX[i] = bx; ref_.BN[i] = bn; ref_.CX[i] = cx; ref_.CN[i] = cn; ref_.SP[i] = sp;
  }
  // thresholds, their range-minimum sparse table and the magnitude suffix minima
  ref_.T.assign(len, INFINITY);
  ref_.BNs.assign(len + 1, INFINITY);
  ref_.CNs.assign(len + 1, INFINITY);
  for (int i = 1; i < len; i++) {
    double br = ref_.B[2 * i], bi = ref_.B[2 * i + 1], cr = ref_.C[2 * i], ci = ref_.C[2 * i + 1];
    bool special = !std::isfinite(br) || !std::isfinite(bi) || !std::isfinite(cr) || !std::isfinite(ci) ||
                   (br == 0.0 && bi == 0.0) || (cr == 0.0 && ci == 0.0);
    if (!special) ref_.T[i] = ltol + log2_sqmag(br, bi) - log2_sqmag(cr, ci);
  }
  for (int i = len - 1; i >= 1; i--) {
    double br = ref_.B[2 * i], bi = ref_.B[2 * i + 1], cr = ref_.C[2 * i], ci = ref_.C[2 * i + 1];
    bool special = !std::isfinite(br) || !std::isfinite(bi) || !std::isfinite(cr) || !std::isfinite(ci) ||
                   (br == 0.0 && bi == 0.0) || (cr == 0.0 && ci == 0.0);
    ref_.BNs[i] = std::min(ref_.BNs[i + 1], special ? INFINITY : log2_sqmag(br, bi));
    ref_.CNs[i] = std::min(ref_.CNs[i + 1], special ? INFINITY : log2_sqmag(cr, ci));
  }
  buildStopLookup();
}

// The tables of the floatexp series test.  Its rounding has no range limits,
// so only the thresholds matter: T_i = log2 tol + log2 |B_i|^2 - log2 |C_i|^2
// from the floatexp coefficients, -inf where B_i = 0 and C_i != 0 (the test
// fires for every e != 0), +inf where C_i = 0 (it never fires); PM is their
// prefix minimum, ST their sparse table of range minima and PMT the lookup over
// the finite range of PM.  The magnitude tables (BX, BN, CX, CN, BNs, CNs) and
// the special counts (SP) belong to the double test's certificate and are not
// read in this mode.
void Mandelbrot::buildThresholdsFx() {
  const int len = ref_.len;
  const double ltol = std::log2(error_tolerance);
  ref_.T.assign(len, INFINITY);
  ref_.PM.assign(len, INFINITY);
  ref_.BX.assign(len, -INFINITY);
  ref_.BN.assign(len, INFINITY);
  ref_.CX.assign(len, -INFINITY);
  ref_.CN.assign(len, INFINITY);
  ref_.SP.assign(len, 0);
  ref_.BNs.assign(len + 1, INFINITY);
  ref_.CNs.assign(len + 1, INFINITY);
  double pm = INFINITY;
  for (int i = 1; i < len; i++) {
    const double lb = log2_sqmag_fx(ref_.Bm[2 * i], ref_.Be[2 * i], ref_.Bm[2 * i + 1], ref_.Be[2 * i + 1]);
    const double lc = log2_sqmag_fx(ref_.Cm[2 * i], ref_.Ce[2 * i], ref_.Cm[2 * i + 1], ref_.Ce[2 * i + 1]);
    const double T = lc == -INFINITY ? INFINITY : lb == -INFINITY ? -INFINITY : ltol + lb - lc;
    ref_.T[i] = T;
    pm = std::min(pm, T);
    ref_.PM[i] = pm;
  }
  buildStopLookup();
}

// the stop-search lookup (cells of the finite PM range, each holding the first
// index whose prefix minimum reaches the cell's lower edge) and the sparse table
// of range minima of T
void Mandelbrot::buildStopLookup() {
  const int len = ref_.len;
  double pmax = -INFINITY, pmin = INFINITY;
  for (int i = 1; i < len; i++)
    if (std::isfinite(ref_.PM[i])) {
      pmax = std::max(pmax, ref_.PM[i]);
      pmin = std::min(pmin, ref_.PM[i]);
    }
  ref_.PMT.clear();
  ref_.pmt_n = 0;
  if (pmax > pmin) {
    const int n = (int)std::min<long>(65536, 8L * len);
    const double h = (pmax - pmin) / n;
    ref_.pmt_base = pmin;
    ref_.pmt_inv_h = 1.0 / h;
    ref_.pmt_n = n;
    ref_.PMT.assign((size_t)n + 1, len);
    int i = len;  // first index with PM[i] <= v; non-increasing in v
    for (int k = 0; k <= n; k++) {
      const double v = pmin + k * h;
      while (i > 1 && ref_.PM[i - 1] <= v) i--;
      ref_.PMT[k] = i;
    }
  }
  int K = 1;
  while ((1 << K) <= len) K++;
  ref_.st_levels = K;
  ref_.ST.assign((size_t)K * len, INFINITY);
  for (int i = 0; i < len; i++) ref_.ST[i] = ref_.T[i];
  for (int k = 1; k < K; k++)
    for (int i = 0; i + (1 << k) <= len; i++)
      ref_.ST[(size_t)k * len + i] =
          std::min(ref_.ST[(size_t)(k - 1) * len + i], ref_.ST[(size_t)(k - 1) * len + i + (1 << (k - 1))]);
}

int Mandelbrot::setSeriesMode(int mode) {
  if (mode != SERIES_DOUBLE && mode != SERIES_FLOATEXP) return NM_ERR_ARG;
  if (mode == series_mode) return NM_OK;
  series_mode = mode;
  if (have_ref_) {  // the coefficients are kept in both forms: only the tables change
    buildThresholds();
    rendered_ = false;
    ref_serial_ = next_serial();
  }
  return NM_OK;
}

// mandelbrot.cpp:261-267
int Mandelbrot::precompute() {
  if (useHardware()) return NM_OK;
  if (N <= 0 || rows() <= 0 || cols() <= 0) return NM_ERR_ARG;
  X.clear();
  ref_ = SaReferenceHost();
  Engine* eng = nullptr;
  if (probe_mode == PROBE_HOST || Engine::instance(&eng) != NM_OK || findProbeGPU(eng) != NM_OK) findProbe();
  computeSeries();
  return uploadReference();
}

// findProbe with the GPU doing the screening: every probe's orbit length is
// estimated by FP64 perturbation (no series, rebasing) against the HP orbit of
// the frame centre, in two small renders on the device; only the probes whose
// estimate is within kProbeMargin of the best are then iterated in HP on the
// host, in probe order, and the first with the greatest exact length wins, as in
// the reference.  The estimates agree with HP lengths on all but about 1e-4 of
// samples and then by 1 iteration, so the margin keeps the true winner among the
// candidates.  The winner's orbit is computed and stored as findProbe does.
int Mandelbrot::findProbeGPU(Engine* eng) {
  constexpr int kProbeMargin = 3;
  const auto pts = probe_points(rows(), cols());
  const int L = hp_limbs_for_bits(hp_default_bits());
  // the centre's orbit as the temporary device reference (A, B, C unused: every
  // probe starts at d0 = e)
  HPComplex c0;
  c0.re = hpf(L); c0.im = hpf(L);
  hp_set(c0.re, center.re);
  hp_set(c0.im, center.im);
  const int lenc = computeOrbit(c0, true);
  X.resize((size_t)lenc + 1);
  SaReferenceHost tmp;
  tmp.len = lenc;
  fillOrbit(tmp);
  tmp.A.assign(2 * (size_t)lenc, 0.0);
  tmp.B = tmp.A;
  tmp.C = tmp.A;
  tmp.PM.assign(lenc, 0.0);
  tmp.BX = tmp.BN = tmp.CX = tmp.CN = tmp.PM;
  tmp.SP.assign(lenc, 0);
  tmp.T = tmp.PM;
  tmp.BNs.assign((size_t)lenc + 1, 0.0);
  tmp.CNs = tmp.BNs;
  tmp.st_levels = 0;
  int st = eng->set_reference(tmp, next_serial());
  if (st != NM_OK) return st;
  // frame A: rows nr/4, nr/2, 3nr/4 x even columns; frame B: even rows x column nc/2
  const int nca = (cols() + 1) / 2, nrb = (rows() + 1) / 2;
  SaFrameHost fa, fb;
  fa.nr = 3; fa.nc = nca; fa.card_mode = CARD_NONE;
  fb.nr = nrb; fb.nc = 1; fb.card_mode = CARD_NONE;
  hpf p(L), y(L);
  auto eps_re = [&](int c) {
    pixel_re(p, center.re, sz.re, (long)(c - cols() / 2));
    hp_sub(y, p, X[0].re);
    return y.get_d();
  };
  auto eps_im = [&](int r) {
    pixel_re(p, center.im, sz.im, (long)(rows() / 2 - r - 1));
    hp_sub(y, p, X[0].im);
    return y.get_d();
  };
  for (int k = 0; k < nca; k++) fa.epsre.push_back(eps_re(2 * k));
  for (int r : {rows() / 4, rows() / 2, 3 * rows() / 4}) fa.epsim.push_back(eps_im(r));
  fb.epsre.push_back(eps_re(cols() / 2));
  for (int k = 0; k < nrb; k++) fb.epsim.push_back(eps_im(2 * k));
  RenderOpts o;
  applyLaunch(o);
  o.tile = 0;
  o.graph = false;
  o.repeats = 1;
  o.delta_mode = DELTA_FP64;
  std::vector<EscapeValue> outa((size_t)3 * nca), outb((size_t)nrb);
  RenderStats rs;
  for (int pass = 0; pass < 2; pass++) {
    const SaFrameHost& f = pass ? fb : fa;
    std::vector<int> ua((size_t)((f.nr + 63) / 64) * ((f.nc + 63) / 64), 1);
    RenderOpts op = o;
    op.tile_ua = ua.data();
    op.skip_T = 64;
    st = eng->render_sa({&f}, op, pass ? outb.data() : outa.data(), &rs);
    if (st != NM_OK) return st;
  }
  std::vector<int> est(pts.size());
  for (size_t i = 0; i < pts.size(); i++) {
    if (i < (size_t)3 * nca) est[i] = outa[(i % 3) * nca + i / 3].iterations;
    else est[i] = outb[i - (size_t)3 * nca].iterations;
  }
  const int emax = *std::max_element(est.begin(), est.end());
  std::vector<int> cand;
  for (size_t i = 0; i < pts.size(); i++)
    if (est[i] >= emax - kProbeMargin) cand.push_back((int)i);
  // exact lengths of the candidates, in order, in parallel batches; a length of
  // N cannot be beaten, so the scan stops at the first batch holding one
  int best = -1, bestlen = 0, hp_orbits = 0;
  const size_t batch = 32;
  for (size_t b0 = 0; b0 < cand.size(); b0 += batch) {
    const size_t b1 = std::min(cand.size(), b0 + batch);
    std::vector<int> lens(b1 - b0, 0);
    parallel_for((long)(b1 - b0), [&](long k) {
      const auto& pt = pts[(size_t)cand[b0 + k]];
      HPComplex q;
      q.re = hpf(L); q.im = hpf(L);
      pixel_re(q.re, center.re, sz.re, (long)(pt.second - cols() / 2));
      pixel_re(q.im, center.im, sz.im, (long)(rows() / 2 - pt.first - 1));
      lens[k] = computeOrbit(q, false);
    }, 1);
    for (size_t k = 0; k < lens.size(); k++)
      if (lens[k] > bestlen) {
        bestlen = lens[k];
        best = cand[b0 + k];
      }
    hp_orbits += (int)lens.size();
    if (bestlen >= N) break;
  }
  if (best < 0) return NM_ERR_STATE;
  probe_candidates_ = hp_orbits;
  winner_ = best;
  HPComplex q;
  q.re = hpf(L); q.im = hpf(L);
  pixel_re(q.re, center.re, sz.re, (long)(pts[(size_t)best].second - cols() / 2));
  pixel_re(q.im, center.im, sz.im, (long)(rows() / 2 - pts[(size_t)best].first - 1));
  const int len = computeOrbit(q, true);
  X.resize((size_t)len + 1);
  ref_.len = len;
  return NM_OK;
}

int Mandelbrot::precomputeAt(const char* re, const char* im) {
  if (N <= 0) return NM_ERR_ARG;
  const int L = hp_limbs_for_bits(hp_default_bits());
  HPComplex p;
  p.re = hpf(L); p.im = hpf(L);
  if (!hp_set_str(p.re, re) || !hp_set_str(p.im, im)) return NM_ERR_ARG;
  X.clear();
  ref_ = SaReferenceHost();
  int len = computeOrbit(p, true);
  X.resize((size_t)len + 1);
  ref_.len = len;
  winner_ = -1;
  probe_candidates_ = 0;
  computeSeries();
  return uploadReference();
}

static std::atomic<uint64_t> g_ref_serial{1};

// The device orbit of X (len + 1 entries): Zd[0] = 0, Zd[k + 1] = nearest(X[k]),
// Xlo[k] = nearest(X[k] - Zd[k + 1])
void Mandelbrot::fillOrbit(SaReferenceHost& r) const {
  const int len = r.len;
  r.Zd.assign(2 * (size_t)(len + 2), 0.0);
  r.Xlo.assign(2 * (size_t)len, 0.0);
  for (int k = 0; k <= len; k++) {
    for (int p = 0; p < 2; p++) {
      const hpf& x = p ? X[k].im : X[k].re;
      double nd = get_d_nearest(x);
      r.Zd[2 * (k + 1) + p] = nd;
      if (k < len) {
        hpf td(2), rem(x.L);
        hp_set_d(td, nd);
        hp_sub(rem, x, td);
        r.Xlo[2 * k + p] = get_d_nearest(rem);
      }
    }
  }
}

int Mandelbrot::uploadReference() {
  fillOrbit(ref_);
  buildThresholds();
  have_ref_ = true;
  rendered_ = false;
  ref_serial_ = next_serial();
  return NM_OK;
}

// ------------------------------------------------------------ frame builders
// Frame-level cardioid/bulb classification: returns CARD_NONE or CARD_ALL when
// the sign of both tests is provably constant over the frame, otherwise -1.
// For |x|, |y| <= 4 the gradient of q(q + x - 1/4) - y^2/4 is bounded by 4000
// and that of (x + 1)^2 + y^2 - 1/16 by 40 (mean value bound).
static int classify_frame(const HPComplex& center, const HPComplex& sz, int nr, int nc) {
  double cx = center.re.get_d(), cy = center.im.get_d();
  double hx = (std::fabs(sz.re.get_d()) * (nc / 2 + 2)), hy = (std::fabs(sz.im.get_d()) * (nr / 2 + 2));
  if (cx + hx < -1.3 || cx - hx > 0.4 || cy + hy < -0.7 || cy - hy > 0.7) return CARD_NONE;
  if (std::fabs(cx) + hx > 4.0 || std::fabs(cy) + hy > 4.0) return -1;
  int L = std::max(center.re.L, 4);
  hpf fourth(L), xmf(L), y2(L), q(L), t(L), lhs(L), rhs(L), f(L);
  hp_set_d(fourth, 0.25);
  hp_sub(xmf, center.re, fourth);
  hp_mul(y2, center.im, center.im);
  hp_mul(q, xmf, xmf);
  hp_add(q, q, y2);
  hp_add(t, q, xmf);
  hp_mul(lhs, q, t);
  hp_mul(rhs, fourth, y2);
  hp_sub(f, lhs, rhs);
  double fcard = f.get_d();
  hpf one(2);
  hp_set_d(one, 1.0);
  hp_add(q, center.re, one);
  hp_mul(lhs, q, q);
  hp_add(lhs, lhs, y2);
  hp_mul(rhs, fourth, fourth);
  hp_sub(f, lhs, rhs);
  double fdisk = f.get_d();
  double h = hx + hy;
  if (std::fabs(fcard) > 4000.0 * h * 1.01 && std::fabs(fdisk) > 40.0 * h * 1.01)
    return (fcard < 0 || fdisk < 0) ? CARD_ALL : CARD_NONE;
  return -1;
}

// Per-column / per-row truncated coordinates of the plain-double path
// (descend(Y0) in getIterationsHW, mandelbrot.cpp:234) plus the double-double
// remainder used by the device cardioid test.
int Mandelbrot::buildHwFrame(HwFrameHost& f, const HPComplex& cen, const HPComplex& s) {
  f.nr = rows();
  f.nc = cols();
  f.cre.resize(f.nc); f.crelo.resize(f.nc);
  f.cim.resize(f.nr); f.cimlo.resize(f.nr);
  fillHwTables(cen, s, f.cre.data(), f.crelo.data(), f.cim.data(), f.cimlo.data(), &f.card_mode, &f.fma_ok,
               &f.c_bound);
  return NM_OK;
}

// The tables of buildHwFrame into caller-owned arrays (cols() and rows()
// entries); allocates nothing.
void Mandelbrot::fillHwTables(const HPComplex& cen, const HPComplex& s, double* cre, double* crelo, double* cim,
                              double* cimlo, int* card_mode, bool* fma_ok, double* c_bound) const {
  const int L = hp_limbs_for_bits(hp_default_bits());
  const int nr = rows(), nc = cols();
  hpf p(L), rem(L), td(2);
  for (int c = 0; c < nc; c++) {
    pixel_re(p, cen.re, s.re, (long)(c - nc / 2));
    double t = p.get_d();
    hp_set_d(td, t);
    hp_sub(rem, p, td);
    cre[c] = t;
    crelo[c] = get_d_nearest(rem);
  }
  for (int r = 0; r < nr; r++) {
    pixel_re(p, cen.im, s.im, (long)(nr / 2 - r - 1));
    double t = p.get_d();
    hp_set_d(td, t);
    hp_sub(rem, p, td);
    cim[r] = t;
    cimlo[r] = get_d_nearest(rem);
  }
  int m = classify_frame(cen, s, nr, nc);
  *card_mode = m < 0 ? CARD_TEST_DD : m;
  auto clear = [](double v) { return v == 0.0 || std::fabs(v) >= 0x1p-200; };
  *fma_ok = std::all_of(cre, cre + nc, clear) && std::all_of(cim, cim + nr, clear);
  double mx = 0.0, my = 0.0;
  for (int c = 0; c < nc; c++) mx = std::max(mx, std::fabs(cre[c]));
  for (int r = 0; r < nr; r++) my = std::max(my, std::fabs(cim[r]));
  *c_bound = std::sqrt(mx * mx + my * my) * (1.0 + 0x1p-40);  // rounded up
}

// computeAll without device allocation (render_hw_noalloc): the plain path
// only; the coordinate tables go through a static host buffer.
int Mandelbrot::computeAllNoAlloc(EscapeValue* out, size_t n) {
  if (!out || N <= 0) return NM_ERR_ARG;
  if (!useHardware()) return NM_ERR_STATE;  // deep views need the device reference orbit
  const int nr = rows(), nc = cols();
  if ((uint64_t)nr * nc > n) return NM_ERR_ARG;
  constexpr int kSide = 8192;
  if (nr > kSide || nc > kSide || (uint64_t)nr * nc > (1ull << 19)) return NM_ERR_RANGE;
  static std::mutex mu;
  static double tab[4 * kSide];
  std::lock_guard<std::mutex> lk(mu);
  HwFrameView v;
  v.nr = nr;
  v.nc = nc;
  double *cre = tab, *crelo = tab + kSide, *cim = tab + 2 * kSide, *cimlo = tab + 3 * kSide;
  fillHwTables(center, sz, cre, crelo, cim, cimlo, &v.card_mode, &v.fma_ok, &v.c_bound);
  v.cre = cre;
  v.crelo = crelo;
  v.cim = cim;
  v.cimlo = cimlo;
  return render_hw_noalloc(v, N, bailout, out, noalloc_err_);
}

// Per-column / per-row epsilon = descend(Y0 - X[0]) (mandelbrot.cpp:155-159).
int Mandelbrot::buildSaFrame(SaFrameHost& f, const HPComplex& cen, const HPComplex& s) {
  if (!have_ref_) return NM_ERR_STATE;
  const int L = hp_limbs_for_bits(hp_default_bits());
  f.nr = rows();
  f.nc = cols();
  f.epsre.resize(f.nc);
  f.epsim.resize(f.nr);
  hpf p(L), y(L);
  for (int c = 0; c < f.nc; c++) {
    pixel_re(p, cen.re, s.re, (long)(c - cols() / 2));
    hp_sub(y, p, X[0].re);
    f.epsre[c] = y.get_d();
  }
  for (int r = 0; r < f.nr; r++) {
    pixel_re(p, cen.im, s.im, (long)(rows() / 2 - r - 1));
    hp_sub(y, p, X[0].im);
    f.epsim[r] = y.get_d();
  }
  int m = classify_frame(cen, s, f.nr, f.nc);
  if (m >= 0) {
    f.card_mode = m;
    f.mask.clear();
  } else {
    // the cardioid or bulb boundary crosses this deep frame: test every sample
    // in HP (mandelbrot.cpp:149) and hand the device a bit mask
    f.card_mode = CARD_MASK;
    size_t n = (size_t)f.nr * f.nc;
    f.mask.assign((n + 31) / 32, 0u);
    std::vector<uint8_t> in(n, 0);
    parallel_for((long)f.nr, [&](long r) {
      hpf pr(L), pi(L);
      pixel_re(pi, cen.im, s.im, (long)(rows() / 2 - r - 1));
      for (int c = 0; c < f.nc; c++) {
        pixel_re(pr, cen.re, s.re, (long)(c - cols() / 2));
        in[(size_t)r * f.nc + c] = inCardioidHP(pr, pi) ? 1 : 0;
      }
    }, 1);
    for (size_t i = 0; i < n; i++)
      if (in[i]) f.mask[i / 32] |= 1u << (i % 32);
  }
  return NM_OK;
}

// ------------------------------------------------------------------ render
std::string Mandelbrot::snapshot() const {
  std::string s;
  auto add = [&](const void* p, size_t n) { s.append((const char*)p, n); };
  for (const hpf* h : {&center.re, &center.im, &sz.re, &sz.im}) {
    add(&h->sign, sizeof(h->sign));
    add(&h->exp, sizeof(h->exp));
    add(h->m, sizeof(uint64_t) * h->L);
  }
  add(&N, sizeof N);
  add(&error_tolerance, sizeof error_tolerance);
  add(&bailout, sizeof bailout);
  add(&delta_mode, sizeof delta_mode);
  add(&skip_mode, sizeof skip_mode);
  add(&skip_tile, sizeof skip_tile);
  add(&grid.nr, sizeof grid.nr);
  add(&grid.nc, sizeof grid.nc);
  add(&ref_serial_, sizeof ref_serial_);
  return s;
}

// the view's parameters and the launch settings into the engine options
void Mandelbrot::applyLaunch(RenderOpts& o) const {
  o.N = N;
  o.bailout = bailout;
  o.tol = error_tolerance;
  o.delta_mode = delta_mode;
  if (o.block_k == 0) o.block_k = launch.block_k;
  if (o.waves_per_simd == 0) o.waves_per_simd = launch.waves_per_simd;
  o.tile = launch.tile;
  o.graph = launch.graph != 0;
  if (o.pre_iter == 0) o.pre_iter = launch.pre_iter;
  if (o.chunk_run == 0) o.chunk_run = launch.chunk_run;
}

int Mandelbrot::render(RenderOpts* opts, bool copy_back) {
  if (N <= 0 || rows() <= 0 || cols() <= 0) return NM_ERR_ARG;
  RenderOpts o = opts ? *opts : RenderOpts();
  applyLaunch(o);
  EscapeValue* out = copy_back ? grid.values.data() : nullptr;
  const bool hw = useHardware();
  if (!hw) {
    if (!have_ref_) return NM_ERR_STATE;  // precompute() first (mandelbrot.cpp:161)
    if (std::fabs(o.tol - error_tolerance) > 0) return NM_ERR_ARG;
  }
  const std::string key = snapshot() + "|view";
  if (key != cache_key_) {
    cache_key_.clear();
    cache_ = FrameSet();
    cache_.hw = hw;
    int st = NM_OK;
    if (hw) {
      cache_.h.resize(1);
      st = buildHwFrame(cache_.h[0]);
    } else {
      cache_.s.resize(1);
      st = buildSaFrame(cache_.s[0]);
    }
    if (st != NM_OK) return st;
    cache_key_ = key;
    cache_serial_ = next_serial();
  }
  o.set_serial = cache_serial_;
  int st = renderSet(cache_, o, out, 0, 1);
  if (st == NM_OK) {
    rendered_ = copy_back;
    snap_ = snapshot();
    dev_grid_ = true;
  }
  return st;
}

// mandelbrot.cpp:269-283.  The first call for a view renders the whole frame;
// later calls for the same view find their row already in the host grid.
int Mandelbrot::computeRow(int r) {
  if (r < 0 || r >= rows()) return NM_ERR_ARG;
  if (rendered_ && snap_ == snapshot()) return NM_OK;
  return render(nullptr, true);
}

int Mandelbrot::computeAll() { return render(nullptr, true); }

// process-wide serial numbers of reference orbits and frame sets (device caches)
uint64_t Mandelbrot::next_serial() { return g_ref_serial.fetch_add(1); }

// A zoom-video batch: every frame shares this object's centre, reference orbit
// and grid size; frame k has sz = width_k / cols (square samples).  All frames go
// to the device in one start launch and one persistent iterate launch.
int Mandelbrot::buildFrames(const std::vector<std::string>& widths, FrameSet& set) {
  if (widths.empty() || N <= 0) return NM_ERR_ARG;
  const int L = hp_limbs_for_bits(hp_default_bits());
  std::vector<HPComplex> sizes(widths.size());
  bool all_hw = true, any_hw = false;
  for (size_t k = 0; k < widths.size(); k++) {
    hpf w(L);
    if (!hp_set_str(w, widths[k].c_str()) || w.sign <= 0) return NM_ERR_ARG;
    sizes[k].re = hpf(L);
    hp_div_ui(sizes[k].re, w, (unsigned long)cols());
    sizes[k].im = sizes[k].re;
    bool hw = sizes[k].re.get_d() >= kMinPreview;
    all_hw = all_hw && hw;
    any_hw = any_hw || hw;
  }
  if (any_hw && !all_hw) return NM_ERR_ARG;  // one path per batch
  set.hw = all_hw;
  if (all_hw) {
    set.h.assign(widths.size(), HwFrameHost());
    parallel_for((long)set.h.size(), [&](long k) { buildHwFrame(set.h[k], center, sizes[k]); }, 1);
    return NM_OK;
  }
  if (!have_ref_) return NM_ERR_STATE;
  set.s.assign(widths.size(), SaFrameHost());
  std::vector<int> sts(set.s.size(), NM_OK);
  parallel_for((long)set.s.size(), [&](long k) { sts[k] = buildSaFrame(set.s[k], center, sizes[k]); }, 1);
  for (int v : sts)
    if (v != NM_OK) return v;
  return NM_OK;
}

int Mandelbrot::renderSet(const FrameSet& set, const RenderOpts& o, EscapeValue* out, size_t first, size_t step) {
  Engine* eng = nullptr;
  int st = Engine::instance(&eng);
  if (st != NM_OK) return st;
  dev_tag_ = 0;
  st = renderSetOn(eng, set, o, out, first, step);
  if (st == NM_OK) {
    dev_tag_ = eng->render_tag();
    dev_frames_ = (int)(((set.hw ? set.h.size() : set.s.size()) - first + step - 1) / step);
  }
  return st;
}

int Mandelbrot::renderSetOn(Engine* eng, const FrameSet& set, const RenderOpts& o, EscapeValue* out, size_t first,
                            size_t step) {
  int st = NM_OK;
  if (set.hw) {
    std::vector<const HwFrameHost*> ptr;
    for (size_t k = first; k < set.h.size(); k += step) ptr.push_back(&set.h[k]);
    return eng->render_hw(ptr, o, out, &stats_);
  }
  st = eng->set_reference(ref_, ref_serial_);
  if (st != NM_OK) return st;
  std::vector<const SaFrameHost*> ptr;
  for (size_t k = first; k < set.s.size(); k += step) ptr.push_back(&set.s[k]);
  if (skip_mode == SKIP_PIXEL) return eng->render_sa(ptr, o, out, &stats_);
  // one series skip per tile, from the tile probes (SKIP_NONE: first unstable
  // index 1 everywhere, so every sample starts at d_0 = e)
  std::vector<int> ua;
  if (skip_mode == SKIP_NONE) {
    const int T = skip_tile, ntc = (cols() + T - 1) / T, ntr = (rows() + T - 1) / T;
    ua.assign((size_t)ntc * ntr * ptr.size(), 1);
    tile_stats = TileStats();
  } else if ((st = tileSkipsFor(eng, ptr, o, ua)) != NM_OK) {
    return st;
  }
  RenderOpts ot = o;
  ot.tile_ua = ua.data();
  ot.skip_T = skip_tile;
  return eng->render_sa(ptr, ot, out, &stats_);
}

// ------------------------------------------------------------ tile skips
namespace {
// complex floatexp: (re + i im) * 2^e with max(|re|, |im|) in [1/2, 1), or zero
struct FeC {
  double re = 0.0, im = 0.0;
  int64_t e = 0;
};
FeC fe_norm(double re, double im, int64_t e) {
  const double m = std::max(std::fabs(re), std::fabs(im));
  if (m == 0.0 || !std::isfinite(m)) return FeC{};
  int k = 0;
  std::frexp(m, &k);
  return FeC{std::ldexp(re, -k), std::ldexp(im, -k), e + k};
}
FeC fe_mul(const FeC& a, const FeC& b) {
  return fe_norm(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re, a.e + b.e);
}
FeC fe_add(const FeC& a, const FeC& b) {
  if (a.re == 0.0 && a.im == 0.0) return b;
  if (b.re == 0.0 && b.im == 0.0) return a;
  const int64_t e = std::max(a.e, b.e);
  const int64_t da = a.e - e, db = b.e - e;
  const double sa = da < -1100 ? 0.0 : std::ldexp(1.0, (int)da), sb = db < -1100 ? 0.0 : std::ldexp(1.0, (int)db);
  return fe_norm(a.re * sa + b.re * sb, a.im * sa + b.im * sb, e);
}
double fe_log2(const FeC& a) {
  const double m = std::hypot(a.re, a.im);
  return m > 0.0 ? (double)a.e + std::log2(m) : -4096.0;
}
}  // namespace

// log2 |a_k(i)| of the Taylor series delta_i = sum_k a_k(i) eps^k of the
// perturbation recurrence delta_{i+1} = 2 X_i delta_i + delta_i^2 + eps with
// delta_0 = eps (a_1(0) = 1), for k = 1..8 at the orbit indices
// i_j = (j + 1) len / 17, j = 0..15 (the Taylor series, not the reference's
// three-term one: features only)
static void taylor_features(const std::vector<HPComplex>& X, int len, std::vector<double>& out) {
  constexpr int K = 8, J = 16;
  out.assign(K * J, -4096.0);
  FeC a[K + 1];
  a[1] = fe_norm(1.0, 0.0, 0);
  int j = 0;
  for (int i = 1; i < len && j < J; i++) {
    const FeC z2 = fe_norm(2.0 * X[i - 1].re.get_d(), 2.0 * X[i - 1].im.get_d(), 0);
    FeC nx[K + 1];
    for (int k = 1; k <= K; k++) {
      FeC v = fe_mul(z2, a[k]);
      for (int p = 1; p < k; p++) v = fe_add(v, fe_mul(a[p], a[k - p]));
      if (k == 1) v = fe_add(v, fe_norm(1.0, 0.0, 0));
      nx[k] = v;
    }
    for (int k = 1; k <= K; k++) a[k] = nx[k];
    while (j < J && i == (int)(((int64_t)(j + 1) * len) / 17)) {
      for (int k = 1; k <= K; k++) out[(size_t)(k - 1) * J + j] = fe_log2(a[k]);
      j++;
    }
  }
}

static double hp_log10(const hpf& v) {
  long e = 0;
  const double m = v.get_d_2exp(&e);
  return m != 0.0 ? std::log10(std::fabs(m)) + (double)e * std::log10(2.0) : -4096.0;
}

int Mandelbrot::tileFeatures(int T, std::vector<float>& feats, const SaFrameHost* frame) {
  if (!have_ref_ || useHardware()) return NM_ERR_STATE;
  if (T < 8) return NM_ERR_ARG;
  if (taylor_serial_ != ref_serial_) {
    taylor_features(X, ref_.len, taylor_feats_);
    taylor_serial_ = ref_serial_;
  }
  SaFrameHost own;
  if (!frame) {
    int st = buildSaFrame(own);
    if (st != NM_OK) return st;
    frame = &own;
  }
  const SaFrameHost& f = *frame;
  hpf w(sz.re.L);
  hp_mul_si(w, sz.re, (long)cols());
  const double lw = hp_log10(w);
  const double wd = std::pow(10.0, lw);  // width as a double exponent-safe scale below
  const int ntc = (cols() + T - 1) / T, ntr = (rows() + T - 1) / T;
  feats.assign((size_t)ntc * ntr * SkipNet::kIn, 0.0f);
  for (int tr = 0; tr < ntr; tr++)
    for (int tc = 0; tc < ntc; tc++) {
      const int c0 = tc * T, c1 = std::min(cols(), c0 + T) - 1, r0 = tr * T, r1 = std::min(rows(), r0 + T) - 1;
      const int pc = std::fabs(f.epsre[c1]) > std::fabs(f.epsre[c0]) ? c1 : c0;
      const int pr = std::fabs(f.epsim[r1]) > std::fabs(f.epsim[r0]) ? r1 : r0;
      float* x = &feats[((size_t)tr * ntc + tc) * SkipNet::kIn];
      x[0] = (float)lw;
      x[1] = wd > 0.0 ? (float)(f.epsre[pc] / wd) : 0.0f;
      x[2] = wd > 0.0 ? (float)(f.epsim[pr] / wd) : 0.0f;
      for (int k = 0; k < 128; k++) x[3 + k] = (float)taylor_feats_[k];
      x[131] = (float)std::log10((double)N);
    }
  return NM_OK;
}

// Indices where the prefix minimum of the log2 thresholds drops ("records"): a
// probe's first unstable index is one of them unless its |e|^2 lies within
// rounding of a threshold.  The network's target is the record's rank, scaled:
// t = (rank + 1) / (R + 1) for R records, and t = 1 for "never unstable"; its
// output snaps to the nearest rank.
void Mandelbrot::ensureRecords() {
  if (records_serial_ == ref_serial_) return;
  records_.clear();
  records_serial_ = ref_serial_;
  for (int i = 1; i < ref_.len; i++)
    if (ref_.PM[i] < (i > 1 ? ref_.PM[i - 1] : INFINITY)) records_.push_back(i);
}

// The network's output as a first unstable index: the record of the nearest
// rank, moved kSnapBack of the records earlier.  The tile search starts its scan
// at the prediction when every index below it is certainly stable (an exact hit
// costs one test), else at the first of six earlier points that is, else at 1:
// an early prediction costs the tests between it and the answer.
constexpr double kSnapBack = 0.02;
int Mandelbrot::snapIndex(double frac) const {
  const int R = (int)records_.size();
  const long rank = std::lround(frac * (R + 1)) - 1;
  if (rank >= R) return 0;  // no unstable index
  const long back = (long)std::ceil(kSnapBack * R);
  return records_[(size_t)std::max(0L, rank - back)];
}

double Mandelbrot::targetOf(int ua) const {
  const int R = (int)records_.size();
  if (ua <= 0) return 1.0;
  auto it = std::lower_bound(records_.begin(), records_.end(), ua);
  const long rank = it - records_.begin();  // a non-record index takes the next record's rank
  return (double)(rank + 1) / (double)(R + 1);
}

int Mandelbrot::tileSkips(int T, int mode, SkipNet* net, std::vector<int>& ua, std::vector<int>* how) {
  if (!have_ref_ || useHardware()) return NM_ERR_STATE;
  Engine* eng = nullptr;
  int st = Engine::instance(&eng);
  if (st != NM_OK) return st;
  if ((st = eng->set_reference(ref_, ref_serial_)) != NM_OK) return st;
  SaFrameHost f;
  if ((st = buildSaFrame(f)) != NM_OK) return st;
  RenderOpts o;
  applyLaunch(o);
  const int keep_mode = skip_mode, keep_T = skip_tile;
  SkipNet* keep_net = skipnet;
  skip_mode = mode == 0 ? SKIP_TILE_SCAN : (mode == 1 ? SKIP_TILE_TABLE : SKIP_TILE_LEARNED);
  skip_tile = T;
  skipnet = net;
  std::vector<const SaFrameHost*> ptr{&f};
  Engine::TileSearch ts;
  st = tileSkipsFor(eng, ptr, o, ua, &ts);
  skip_mode = keep_mode;
  skip_tile = keep_T;
  skipnet = keep_net;
  if (st == NM_OK && how) *how = ts.how;
  return st;
}

int Mandelbrot::tileSkipsFor(Engine* eng, const std::vector<const SaFrameHost*>& frames, const RenderOpts& o,
                             std::vector<int>& ua, Engine::TileSearch* out) {
  Engine::TileSearch ts;
  ts.T = skip_tile;
  tile_stats = TileStats();
  if (skip_mode == SKIP_TILE_LEARNED) {
    if (!skipnet || !skipnet->trained) return NM_ERR_STATE;
    if (frames.size() != 1) return NM_ERR_ARG;  // the features describe this object's view
    auto t0 = std::chrono::steady_clock::now();
    // the predictions are a function of the view and the network: kept until either changes
    const std::string key = snapshot() + "|" + std::to_string((uintptr_t)skipnet) + "|" +
                            std::to_string(skipnet->version) + "|" + std::to_string(skip_tile);
    if (key != pred_key_) {
      pred_key_.clear();
      std::vector<float> feats;
      int st = tileFeatures(skip_tile, feats, frames[0]);
      if (st != NM_OK) return st;
      const size_t nt = feats.size() / SkipNet::kIn;
      std::vector<float> y(nt);
      if ((st = skipnet->predict(feats.data(), nt, y.data())) != NM_OK) return NM_ERR_HIP;
      ensureRecords();
      pred_.resize(nt);
      for (size_t t = 0; t < nt; t++) pred_[t] = snapIndex((double)y[t]);
      pred_key_ = key;
    }
    ts.pred = pred_;
    tile_stats.ms_predict = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ts.mode = 2;
  } else {
    ts.mode = skip_mode == SKIP_TILE_TABLE ? 1 : 0;
  }
  int st = eng->tile_skip(frames, o, ts);
  if (st != NM_OK) return st;
  ua = ts.ua;
  tile_stats.ntiles = (int)ua.size();
  for (size_t t = 0; t < ua.size(); t++) {
    tile_stats.accepted += ts.how[t] == 2;
    tile_stats.rejected += ts.how[t] == 3;
    tile_stats.bounded += ts.how[t] == 4;
    tile_stats.evals += (uint64_t)ts.evals[t];
  }
  tile_stats.ms_search = ts.ms;
  if (out) *out = ts;
  return NM_OK;
}

int Mandelbrot::renderFrames(const std::vector<std::string>& widths, RenderOpts* opts, EscapeValue* out) {
  RenderOpts o = opts ? *opts : RenderOpts();
  applyLaunch(o);
  std::string key = snapshot() + "|frames";
  for (const auto& w : widths) key += "|" + w;
  if (key != cache_key_) {
    cache_key_.clear();
    cache_ = FrameSet();
    int st = buildFrames(widths, cache_);
    if (st != NM_OK) return st;
    cache_key_ = key;
    cache_serial_ = next_serial();
  }
  o.set_serial = cache_serial_;
  const int st = renderSet(cache_, o, out, 0, 1);
  dev_grid_ = false;
  return st;
}

// viewer.cpp:89-122 on the device.  A frame of the last renderFrames exists only
// on the device; the object's grid is coloured from the device copy of its last
// render when that copy is still the engine's and the view has not changed since
// (scaleUp / scaleDown change it), else from the host grid, as the viewer colours
// whatever the grid holds.
int Mandelbrot::colour(int frame, const uint8_t* pal_rgb, int npal, int smooth, int sc, uint8_t* out, double* ms) {
  if (sc <= 0 || !pal_rgb || !out || npal < N) return NM_ERR_ARG;
  Engine* eng = nullptr;
  int st = Engine::instance(&eng);
  if (st != NM_OK) return st;
  const bool on_device = dev_tag_ != 0 && dev_tag_ == eng->render_tag();
  if (frame >= 0) {
    if (dev_grid_ || !on_device || frame >= dev_frames_) return NM_ERR_STATE;
    return eng->colour(nullptr, dev_tag_, (size_t)frame * rows() * cols(), rows(), cols(), sc, N, pal_rgb, npal, smooth,
                       out, ms);
  }
  if (dev_grid_ && on_device && snap_ == snapshot())
    return eng->colour(nullptr, dev_tag_, 0, rows(), cols(), sc, N, pal_rgb, npal, smooth, out, ms);
  return eng->colour(grid.values.data(), 0, 0, rows(), cols(), sc, N, pal_rgb, npal, smooth, out, ms);
}

// viewer.cpp:186-253: a 1080 x 1920 x 3 x 3 grid with this object's N and centre
// and sz scaled by rows() / 3240 (a double factor, as the viewer computes it);
// rendered, then coloured with 3 x 3 colour averaging.  This object is unchanged,
// and so is the process-wide default precision (which the viewer leaves at the
// beauty grid's).
int Mandelbrot::beautyRender(const uint8_t* pal_rgb, int npal, int smooth, uint8_t* out, double* ms) const {
  if (!pal_rgb || !out) return NM_ERR_ARG;
  struct KeepBits {
    long bits = hp_default_bits();
    ~KeepBits() { hp_set_default_bits(bits); }
  } keep;
  const int sc = 3;
  Mandelbrot b(1080 * sc, 1920 * sc);
  // N, centre and the rescaled sz are copied (viewer.cpp:195-200); the error
  // tolerance and bailout are the new object's defaults, as there.  How the grid
  // is computed (delta arithmetic, launch shape, probe screening) follows this one.
  b.N = N;
  b.delta_mode = delta_mode;
  b.launch = launch;
  b.probe_mode = probe_mode;
  b.center = center;
  const double f = (double)rows() / b.rows();
  b.sz.re = hpf(sz.re.L);
  b.sz.im = hpf(sz.im.L);
  hp_mul_d(b.sz.re, sz.re, f);
  hp_mul_d(b.sz.im, sz.im, f);
  b.setPrecision();
  int st = b.precompute();
  if (st != NM_OK) return st;
  if ((st = b.render(nullptr, true)) != NM_OK) return st;
  return b.colour(-1, pal_rgb, npal, smooth, sc, out, ms);
}

// initAutoZoom (viewer.cpp:441-453): the full view of the set (sz = 4 / cols,
// 3 / rows of the current grid), back to one sample per pixel, then 3 x 3 samples
int Mandelbrot::zoomVideoStart(int sc) {
  if (sc <= 0 || rows() % sc || cols() % sc) return NM_ERR_ARG;
  const int L = hp_limbs_for_bits(hp_default_bits());
  sz.re = hpf(L);
  sz.im = hpf(L);
  hp_set_d(sz.re, 4.0 / cols());
  hp_set_d(sz.im, 3.0 / rows());
  int st = scaleDown(sc);
  if (st != NM_OK) return st;
  return scaleUp(3);
}

// update()'s zoom branch (viewer.cpp:270-286): render the 3 x 3-sampled grid,
// colour it 2 x 2-averaged into a keyframe 1.5 times the output size, make the
// in-between frames from the previous keyframe (none for the first), then zoom
// in 1.5 times about the unchanged centre.
int Mandelbrot::zoomVideoStep(const uint8_t* pal_rgb, int npal, int smooth, int rate, const uint8_t* prev_key,
                              uint8_t* key_out, uint8_t* frames_out, int* nframes, double* ms) {
  if (!pal_rgb || !key_out || rate <= 0 || rows() % 3 || cols() % 3) return NM_ERR_ARG;
  if (prev_key && !frames_out) return NM_ERR_ARG;
  int st = precompute();
  if (st != NM_OK) return st;
  if ((st = render(nullptr, true)) != NM_OK) return st;
  double t0 = 0.0, t1 = 0.0;
  if ((st = colour(-1, pal_rgb, npal, smooth, 2, key_out, &t0)) != NM_OK) return st;
  int n = 0;
  if (prev_key) {
    Engine* eng = nullptr;
    if ((st = Engine::instance(&eng)) != NM_OK) return st;
    st = eng->video(prev_key, key_out, rows() / 2, cols() / 2, rows() / 3, cols() / 3, rate, frames_out, &t1);
    if (st != NM_OK) return st;
    n = rate;
  }
  if (nframes) *nframes = n;
  if (ms) *ms = t0 + t1;
  zoom(1.5f);
  return NM_OK;
}

// The project's launch autotuner.  Candidates: sample order (row-major, 32x32,
// 64x64 and 128x128 tiles), persistent waves per SIMD (2, 4, 6, 8) and the block
// size (16, 32, 64; FP32X deltas 8, 16, 32); then, on the series path, the
// fastest of those with the active list interleaved in runs of 4 (chunk_run).
// Each candidate renders an evenly spaced subsample of at most 16 frames (so the
// set's depth mix is kept) once untimed and once timed; the fastest total kernel
// time wins.
int Mandelbrot::autotuneFrames(const std::vector<std::string>& widths, std::vector<Launch>* table) {
  FrameSet set;
  int st = buildFrames(widths, set);
  if (st != NM_OK) return st;
  const size_t n = set.hw ? set.h.size() : set.s.size();
  const size_t step = std::max<size_t>(1, (n + 15) / 16);
  const int tiles[4] = {0, 32, 64, 128};
  const int waves[4] = {2, 4, 6, 8};
  // series blocks: FP64 up to 64 steps, FP32X (renormalised at block ends) up to 32
  const int ks_sa[3] = {16, 32, 64}, ks_fx[3] = {8, 16, 32}, ks_hw[3] = {16, 32, 64};
  const int nks = 3;
  Launch best;
  best.ms = 1e300;
  if (table) table->clear();
  RenderStats keep = stats_;
  const int* ks = set.hw ? ks_hw : delta_mode == DELTA_FP32X ? ks_fx : ks_sa;
  const uint64_t subset_serial = next_serial();  // the subsample's tables stay on the device
  for (int tile : tiles)
    for (int w : waves)
      for (int ki = 0; ki < nks; ki++) {
        const int k = ks[ki];
        RenderOpts o;
        o.block_k = k;
        o.waves_per_simd = w;
        applyLaunch(o);
        o.tile = tile;
        o.graph = false;
        o.repeats = 1;
        o.chunk_run = 64;
        o.set_serial = subset_serial;
        if ((st = renderSet(set, o, nullptr, 0, step)) != NM_OK) return st;
        if ((st = renderSet(set, o, nullptr, 0, step)) != NM_OK) return st;
        Launch l;
        l.block_k = k;
        l.waves_per_simd = w;
        l.tile = tile;
        l.graph = launch.graph;
        l.pre_iter = launch.pre_iter;
        l.chunk_run = 0;
        l.ms = stats_.ms_total;
        if (table) table->push_back(l);
        if (l.ms < best.ms) best = l;
      }
  if (!set.hw) {  // the best setting with interleaved chunks
    RenderOpts o;
    o.block_k = best.block_k;
    o.waves_per_simd = best.waves_per_simd;
    applyLaunch(o);
    o.tile = best.tile;
    o.graph = false;
    o.repeats = 1;
    o.chunk_run = 4;
    o.set_serial = subset_serial;
    if ((st = renderSet(set, o, nullptr, 0, step)) != NM_OK) return st;
    if ((st = renderSet(set, o, nullptr, 0, step)) != NM_OK) return st;
    Launch l = best;
    l.chunk_run = 4;
    l.ms = stats_.ms_total;
    if (table) table->push_back(l);
    if (l.ms < best.ms) best = l;
  }
  stats_ = keep;
  launch = best;
  return NM_OK;
}

// mandelbrot.cpp:285-290
HPComplex Mandelbrot::pointAt(int r, int c, int sc) const {
  HPComplex pt;
  pixel_re(pt.re, center.re, sz.re, (long)(sc * c - cols() / 2));
  pixel_re(pt.im, center.im, sz.im, (long)(rows() / 2 - sc * r - 1));
  return pt;
}

// mandelbrot.cpp:292-295
void Mandelbrot::translate(int dr, int dc, int sc) {
  center.re = center.re - sz.re * (long)(dc * sc);
  center.im = center.im + sz.im * (long)(dr * sc);
}

// mandelbrot.cpp:297-302
void Mandelbrot::zoom(float scale) {
  sz.re *= (1.0 / scale);
  sz.im *= (1.0 / scale);
  setPrecision();
}

// mandelbrot.cpp:304-316
void Mandelbrot::zoomAt(float scale, int r, int c, int sc) {
  HPComplex pt;
  pixel_re(pt.re, center.re, sz.re, (long)(sc * c - cols() / 2));
  pixel_re(pt.im, center.im, sz.im, (long)(rows() / 2 - sc * r - 1));
  sz.re *= (1.0 / scale);
  sz.im *= (1.0 / scale);
  center.re = pt.re - sz.re * (long)(sc * c - cols() / 2);
  center.im = pt.im - sz.im * (long)(rows() / 2 - sc * r - 1);
  setPrecision();
}

// mandelbrot.cpp:318
const EscapeValue& Mandelbrot::at(int r, int c) { return grid.at(r, c); }

// mandelbrot.cpp:320-332 (float sum, r1 outer, c1 inner)
EscapeValue Mandelbrot::at(int r, int c, int sc) {
  float sum = 0.0f;
  for (int r1 = 0; r1 < sc; r1++)
    for (int c1 = 0; c1 < sc; c1++) {
      const EscapeValue& e = grid.at(sc * r + r1, sc * c + c1);
      sum += e.iterations + e.smoothing;
    }
  sum /= sc * sc;
  EscapeValue o;
  o.iterations = (int)sum;
  o.smoothing = sum - o.iterations;
  return o;
}

// mandelbrot.cpp:334-346
int Mandelbrot::scaleUp(int sc) {
  if (sc <= 0) return NM_ERR_ARG;
  RenderGrid scaled(rows() * sc, cols() * sc);
  for (int r = 0; r < scaled.nr; r++)
    for (int c = 0; c < scaled.nc; c++) scaled.at(r, c) = grid.at(r / sc, c / sc);
  grid = std::move(scaled);
  sz.re = sz.re / sc;
  sz.im = sz.im / sc;
  setPrecision();
  return NM_OK;
}

// mandelbrot.cpp:348-360, the averaging done by the device kernel
int Mandelbrot::scaleDown(int sc) {
  if (sc <= 0) return NM_ERR_ARG;
  RenderGrid scaled(rows() / sc, cols() / sc);
  if (scaled.nr > 0 && scaled.nc > 0) {
    Engine* eng = nullptr;
    int st = Engine::instance(&eng);
    if (st != NM_OK) return st;
    st = eng->downsample(grid.values.data(), rows(), cols(), sc, scaled.values.data());
    if (st != NM_OK) return st;
  }
  grid = std::move(scaled);
  sz.re = sz.re * (long)sc;
  sz.im = sz.im * (long)sc;
  setPrecision();
  return NM_OK;
}

}  // namespace nm
