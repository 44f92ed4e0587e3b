// kernels.hip - gfx950 kernels of the per-pixel deep-zoom iteration and the
// device engine that drives them.
//
// Layout in HBM (one frame set per render):
//   frames[]        DevFrame per frame (sample base, size, table offsets, output base)
//   col/row tables  plain path: truncated coordinate + double-double remainder;
//                   series path: epsilon = descend(Y0 - X[0]) per column / row
//   out[]           EscapeValue {int32, float} row-major per frame (grid.h layout)
//   active[]        u32 sample ids that still need iterating (compacted)
//   state[]         per active slot: delta (2 x f64), epsilon (2 x f64), it, m
//   Zd[]            reference orbit as f64 pairs, Zd[0] = 0, padded past its end
//
// Two kernels per path:
//   *_start   one thread per sample: cardioid/bulb test (plain path) or the
//             reference's per-sample series start (series path); resolves what it
//             can and appends the rest to active[] with wave ballots (one atomic
//             per wave).
//   *_iterate persistent waves: each lane holds one sample; K iterations run
//             without exits, escapes are found afterwards by a per-lane replay of
//             the block (escape is absorbing, so |z| at the block end is enough);
//             finished lanes are refilled from a wave-owned chunk of active[]
//             (ballot + popcount ranks, one atomic per 64 samples).
// This file is compiled with -ffp-contract=off: the plain path must reproduce the
// reference's double arithmetic bit for bit (complex.h:19-31, makefile:3 has no
// FMA); the perturbation path writes its FMAs explicitly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

namespace nm {

// ------------------------------------------------------------ device structs
struct DevFrame {
  uint32_t sample_base;  // first global sample id of this frame
  uint32_t nr, nc;
  uint32_t tab_c, tab_r;  // offsets into the column / row tables
  int32_t card_mode;
  uint64_t out_base;   // first output element
  uint64_t mask_base;  // first bit in the cardioid mask
};

enum CounterIdx {
  C_ACTIVE = 0,   // entries appended to active[] (u32 view of the low word)
  C_WORK = 1,     // chunk cursor of the iterate kernel
  C_CARDIOID = 2,
  C_SA_ESC = 3,
  C_EXECUTED = 4,
  C_SKIPPED = 5,
  C_REBASED = 6,
  C_CREDITED = 7,
  C_COUNT = 8
};

constexpr int kChunk = 64;  // samples a wave takes per atomic

struct PlainArgs {
  const DevFrame* frames;
  int nframes;
  uint32_t total;
  const double* cre;
  const double* crelo;
  const double* cim;
  const double* cimlo;
  EscapeValue* out;
  uint32_t* active;
  unsigned long long* cnt;
  int N;
  double b2, b2m, log_b;
};

struct SeriesArgs {
  const DevFrame* frames;
  int nframes;
  uint32_t total;
  const double* epsre;
  const double* epsim;
  const uint32_t* mask;
  EscapeValue* out;
  uint32_t* active;
  double4* state;    // dr, di, er, ei
  int2* state_i;     // it, m
  unsigned long long* cnt;
  // reference
  const double2* Zd;  // len + 2 (+ padding)
  const double2* Xlo; // len
  const double2* A;
  const double2* B;
  const double2* C;
  const double* PM;
  const double* BX;
  const double* BN;
  const double* CX;
  const double* CN;
  const int* SP;
  int len;
  int mend;  // len + 1
  int N;
  double tol, b2, b2m, log_b;
};

// ------------------------------------------------------------ device helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int find_frame(const DevFrame* f, int n, uint32_t gid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (f[mid].sample_base <= gid) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// mandelbrot.cpp:133-136 (log(bailout) comes from the host)
__device__ __forceinline__ float smooth_of(double r2, double log_b) {
  return (float)(1.0 - log2(0.5 * log(r2) / log_b));
}

// wave-aggregated u64 counter add
__device__ __forceinline__ void wave_add(unsigned long long* p, unsigned long long v) {
  // butterfly reduction over the 64 lanes
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane_id() == 0 && v) atomicAdd(p, v);
}

// Append `want` lanes to a list: returns the slot for wanting lanes.
__device__ __forceinline__ uint32_t wave_append(unsigned long long* counter, bool want) {
  unsigned long long m = __ballot(want);
  if (m == 0) return 0;
  int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = (uint32_t)atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  unsigned long long below = m & ((1ull << lane_id()) - 1ull);
  return base + (uint32_t)__popcll(below);
}

// ---- double-double arithmetic for the plain path's cardioid test -----------
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd two_sum(double a, double b) {
  double s = a + b;
  double bp = s - a;
  double e = (a - (s - bp)) + (b - bp);
  return {s, e};
}
__device__ __forceinline__ dd quick_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
__device__ __forceinline__ dd dd_mul(dd a, dd b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);  // exact low part of the product
  e = fma(a.hi, b.lo, e);
  e = fma(a.lo, b.hi, e);
  return quick_two_sum(p, e);
}
__device__ __forceinline__ dd dd_from(double hi, double lo) { return quick_two_sum(hi, lo); }
__device__ __forceinline__ bool dd_lt(dd a, dd b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// mandelbrot.cpp:63-71 on the double-double coordinate (hi = descend(Y0), lo =
// the HP remainder): q(q + x - 1/4) < y^2/4, then (x + 1)^2 + y^2 < 1/16.
__device__ bool dd_in_cardioid(double xh, double xl, double yh, double yl) {
  dd x = dd_from(xh, xl), y = dd_from(yh, yl);
  dd xmf = dd_add(x, dd{-0.25, 0.0});
  dd y2 = dd_mul(y, y);
  dd q = dd_add(dd_mul(xmf, xmf), y2);
  dd lhs = dd_mul(q, dd_add(q, xmf));
  dd rhs = {y2.hi * 0.25, y2.lo * 0.25};
  if (dd_lt(lhs, rhs)) return true;
  dd q2 = dd_add(x, dd{1.0, 0.0});
  dd lhs2 = dd_add(dd_mul(q2, q2), y2);
  return dd_lt(lhs2, dd{0.0625, 0.0});
}

// descend(X[k] + d) with X[k] = hi + lo: round-to-nearest sum, then one step
// toward zero when the exact residual has the opposite sign.
__device__ __forceinline__ double trunc_sum3(double hi, double lo, double d) {
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

// ================================================================ plain path
// getIterationsHW (mandelbrot.cpp:231-254), start kernel: cardioid/bulb test,
// compaction of the samples that iterate.
__global__ __launch_bounds__(256) void k_plain_start(PlainArgs a) {
  unsigned long long ncard = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < a.total; base += stride) {
    uint32_t gid = base + threadIdx.x;
    bool valid = gid < a.total;
    bool in = false;
    if (valid) {
      const DevFrame& f = a.frames[find_frame(a.frames, a.nframes, gid)];
      uint32_t local = gid - f.sample_base;
      uint32_t r = local / f.nc, c = local - r * f.nc;
      if (f.card_mode == CARD_ALL) in = true;
      else if (f.card_mode == CARD_TEST_DD)
        in = dd_in_cardioid(a.cre[f.tab_c + c], a.crelo[f.tab_c + c], a.cim[f.tab_r + r], a.cimlo[f.tab_r + r]);
      if (in) {
        a.out[f.out_base + local] = EscapeValue{a.N, 0.0f};
        ncard++;
      }
    }
    bool want = valid && !in;
    uint32_t slot = wave_append(&a.cnt[C_ACTIVE], want);
    if (want) a.active[slot] = gid;
  }
  wave_add(&a.cnt[C_CARDIOID], ncard);
}

// Iterate kernel: z <- sq(z) + z0 with the reference's operation order,
// count = index of the update that escapes (mandelbrot.cpp:243-246).
template <int K>
__global__ __launch_bounds__(256) void k_plain_iterate(PlainArgs a) {
  const uint32_t total = (uint32_t)a.cnt[C_ACTIVE];
  unsigned int* work = (unsigned int*)&a.cnt[C_WORK];
  uint32_t cpos = 0, cend = 0;
  bool exhausted = false;
  bool act = false;
  uint32_t pix = 0;
  double cx = 0.0, cy = 0.0, x = 0.0, y = 0.0, r2 = 0.0, i2 = 0.0;
  int k = 0;
  unsigned long long executed = 0;
  for (;;) {
    unsigned long long need = __ballot(!act);
    while (need && !exhausted) {
      if (cpos >= cend) {
        uint32_t b = 0;
        if (lane_id() == 0) b = atomicAdd(work, (unsigned int)kChunk);
        b = __shfl(b, 0);
        cpos = b;
        cend = min(b + (uint32_t)kChunk, total);
        if (cpos >= total) {
          exhausted = true;
          break;
        }
      }
      uint32_t avail = cend - cpos;
      uint32_t nneed = (uint32_t)__popcll(need);
      uint32_t take = min(nneed, avail);
      uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane_id()) - 1ull));
      if (!act && rank < take) {
        pix = a.active[cpos + rank];
        const DevFrame& f = a.frames[find_frame(a.frames, a.nframes, pix)];
        uint32_t local = pix - f.sample_base;
        uint32_t r = local / f.nc, c = local - r * f.nc;
        cx = a.cre[f.tab_c + c];
        cy = a.cim[f.tab_r + r];
        x = cx;
        y = cy;
        r2 = x * x;
        i2 = y * y;
        k = 0;
        act = true;
      }
      cpos += take;
      need = __ballot(!act);
    }
    if (__ballot(act) == 0) break;
    const bool fast = __ballot(act && k + K > a.N) == 0;
    const double x0 = x, y0 = y;
    const int k0 = k;
    bool replay = act;
    if (fast) {
#pragma unroll
      for (int j = 0; j < K; j++) {
        double t = x + x;
        y = t * y + cy;
        x = r2 - i2 + cx;
        r2 = x * x;
        i2 = y * y;
      }
      k += K;
      replay = act && !(r2 + i2 <= a.b2m);
      if (act && !replay) executed += K;
    }
    if (replay) {
      x = x0;
      y = y0;
      k = k0;
      r2 = x * x;
      i2 = y * y;
      bool done = false;
      int count = a.N;
      float sm = 0.0f;
      for (int j = 0; j < K; j++) {
        if (k >= a.N) {
          done = true;
          break;
        }
        double t = x + x;
        y = t * y + cy;
        x = r2 - i2 + cx;
        r2 = x * x;
        i2 = y * y;
        executed++;
        double mg = r2 + i2;
        if (mg > a.b2) {
          done = true;
          count = k;
          sm = smooth_of(mg, a.log_b);
          break;
        }
        k++;
      }
      if (!done && k >= a.N) done = true;
      if (done) {
        const DevFrame& f = a.frames[find_frame(a.frames, a.nframes, pix)];
        a.out[f.out_base + (pix - f.sample_base)] = EscapeValue{count, sm};
        act = false;
      }
    }
  }
  wave_add(&a.cnt[C_EXECUTED], executed);
}

// =============================================================== series path
// The reference's per-sample series start (mandelbrot.cpp:155-207).
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {  // complex.h:29-31
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// isUnstable(B[i] e^2, C[i] e^3) exactly as mandelbrot.cpp:138-142, 166-173
__device__ __forceinline__ bool unstable_at(const SeriesArgs& a, int i, double2 e2, double2 e3) {
  double2 b = cmul(a.B[i], e2);
  double2 c = cmul(a.C[i], e3);
  double bmag = b.x * b.x + b.y * b.y;
  double cmag = c.x * c.x + c.y * c.y;
  return bmag * a.tol < cmag;
}

// d[i] = A[i] e + B[i] e^2 + C[i] e^3 in the reference's order; d[0] = e
__device__ __forceinline__ double2 sa_term(const SeriesArgs& a, int i, double2 e, double2 e2, double2 e3) {
  if (i == 0) return e;
  double2 ta = cmul(a.A[i], e);
  double2 tb = cmul(a.B[i], e2);
  double2 tc = cmul(a.C[i], e3);
  double2 s = make_double2(ta.x + tb.x, ta.y + tb.y);
  return make_double2(s.x + tc.x, s.y + tc.y);
}

// first i in [lo, hi] with PM[i] (non-increasing) <= v, or hi + 1
__device__ __forceinline__ int first_le(const double* PM, int lo, int hi, double v) {
  int l = lo, h = hi + 1;
  while (l < h) {
    int mid = (l + h) >> 1;
    if (PM[mid] <= v) h = mid;
    else l = mid + 1;
  }
  return l;
}

// First unstable index of the reference's scan (0 if none).  The stop test is
// |e|^2 > T_i up to rounding; indices whose threshold is more than kMarg (in
// log2) away from log2|e|^2 are decided by the prefix minimum, the rest by the
// exact test, and the magnitude prefixes prove the skipped exact tests would not
// under- or overflow.  Otherwise the reference's linear scan runs.
constexpr double kMarg = 4.0e-9;
__device__ int sa_stop_index(const SeriesArgs& a, double2 e, double2 e2, double2 e3) {
  const int last = a.len - 1;
  if (last < 1) return 0;
  double mx = fmax(fabs(e.x), fabs(e.y));
  if (mx == 0.0) return 0;  // e = 0: b = c = 0 and the test never fires
  int ex = ilogb(mx);
  double sx = ldexp(e.x, -ex), sy = ldexp(e.y, -ex);
  double L = log2(sx * sx + sy * sy) + 2.0 * ex;
  int ia = first_le(a.PM, 1, last, L + kMarg);
  int safe_idx = ia > last ? last : ia - 1;
  bool safe = L > -600.0 && L < 600.0;
  if (safe && safe_idx >= 1) {
    safe = a.SP[safe_idx] == 0 && a.BX[safe_idx] + 2.0 * L < 960.0 && a.BN[safe_idx] + 2.0 * L > -960.0 &&
           a.CX[safe_idx] + 3.0 * L < 960.0 && a.CN[safe_idx] + 3.0 * L > -960.0 && a.BN[safe_idx] + 2.0 * L > -900.0;
  }
  int from = safe ? ia : 1;
  for (int i = from; i <= last; i++)
    if (unstable_at(a, i, e2, e3)) return i;
  return 0;
}

__global__ __launch_bounds__(256) void k_series_start(SeriesArgs a) {
  unsigned long long ncard = 0, nesc = 0, skipped = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < a.total; base += stride) {
    uint32_t gid = base + threadIdx.x;
    bool valid = gid < a.total;
    bool want = false;
    double2 dl = make_double2(0.0, 0.0), e = make_double2(0.0, 0.0);
    int it = 0, m = 0;
    if (valid) {
      const DevFrame& f = a.frames[find_frame(a.frames, a.nframes, gid)];
      uint32_t local = gid - f.sample_base;
      uint32_t r = local / f.nc, c = local - r * f.nc;
      bool in = f.card_mode == CARD_ALL;
      if (f.card_mode == CARD_MASK) {
        uint64_t bit = f.mask_base + local;
        in = (a.mask[bit >> 5] >> (bit & 31)) & 1u;
      }
      EscapeValue res{a.N, 0.0f};
      bool resolved = true;
      if (in) {
        ncard++;
      } else {
        e = make_double2(a.epsre[f.tab_c + c], a.epsim[f.tab_r + r]);
        // eps2 = sq(eps), eps3 = eps * eps2 (complex.h:19-21, mandelbrot.cpp:160-161)
        double2 e2 = make_double2(e.x * e.x - e.y * e.y, 2.0 * e.x * e.y);
        double2 e3 = cmul(e, e2);
        int ua = sa_stop_index(a, e, e2, e3);
        int dsize = ua ? max(ua - 3, 1) : a.len;
        int found = dsize - 1;
        double2 d = sa_term(a, found, e, e2, e3);
        double2 zf = a.Zd[found + 1], xl = a.Xlo[found];
        double yr = trunc_sum3(zf.x, xl.x, d.x), yi = trunc_sum3(zf.y, xl.y, d.y);
        if (yr * yr + yi * yi > a.b2) {
          // escape inside the skip: the reference's binary search (mandelbrot.cpp:186-206)
          int low = 0, high = dsize - 1, mid = dsize / 2;
          while (low <= high) {
            double2 dm = sa_term(a, mid, e, e2, e3);
            double2 zm = a.Zd[mid + 1], xm = a.Xlo[mid];
            double ym = trunc_sum3(zm.x, xm.x, dm.x), yn = trunc_sum3(zm.y, xm.y, dm.y);
            if (!(ym * ym + yn * yn > a.b2)) low = mid + 1;
            else {
              high = mid - 1;
              found = mid;
            }
            mid = (low + high) / 2;
          }
          d = sa_term(a, found, e, e2, e3);
          zf = a.Zd[found + 1];
          xl = a.Xlo[found];
          yr = trunc_sum3(zf.x, xl.x, d.x);
          yi = trunc_sum3(zf.y, xl.y, d.y);
          res = EscapeValue{found, smooth_of(yr * yr + yi * yi, a.log_b)};
          nesc++;
        } else if (found + 1 < a.N) {
          // first perturbation step from Y[found] = X[found] + d (its escape
          // test was the one above); rebase if |Y| < |d|
          double dr = d.x, di = d.y;
          m = found + 1;
          double zr = zf.x, zi = zf.y;
          double pr = zr + dr, pi = zi + di;
          double mag = fma(pr, pr, pi * pi);
          double dmag = fma(dr, dr, di * di);
          if (mag < dmag || m == a.mend) {
            dr = pr;
            di = pi;
            zr = 0.0;
            zi = 0.0;
            m = 0;
          }
          double wr = zr + pr, wi = zi + pi;
          double ndr = fma(dr, wr, fma(-di, wi, e.x));
          double ndi = fma(dr, wi, fma(di, wr, e.y));
          dl = make_double2(ndr, ndi);
          it = found + 1;
          m += 1;
          want = true;
          resolved = false;
          skipped += (unsigned long long)(found + 1);
        } else {
          nesc++;  // the skip reached N: (N, 0) as the empty HP loop returns
        }
      }
      if (resolved) a.out[f.out_base + local] = res;
    }
    uint32_t slot = wave_append(&a.cnt[C_ACTIVE], want);
    if (want) {
      a.active[slot] = gid;
      a.state[slot] = make_double4(dl.x, dl.y, e.x, e.y);
      a.state_i[slot] = make_int2(it, m);
    }
  }
  wave_add(&a.cnt[C_CARDIOID], ncard);
  wave_add(&a.cnt[C_SA_ESC], nesc);
  wave_add(&a.cnt[C_SKIPPED], skipped);
}

// FP64 perturbation with rebasing (replaces the HP continuation of
// mandelbrot.cpp:209-224): Y = Zd[m] + d, d' = d (Zd[m] + Y) + e, rebase to
// m = 0 when |Y| < |d| or at the end of the reference orbit.
template <int K>
__global__ __launch_bounds__(256) void k_series_iterate(SeriesArgs a) {
  const uint32_t total = (uint32_t)a.cnt[C_ACTIVE];
  unsigned int* work = (unsigned int*)&a.cnt[C_WORK];
  uint32_t cpos = 0, cend = 0;
  bool exhausted = false;
  bool act = false;
  uint32_t pix = 0;
  double dr = 0.0, di = 0.0, er = 0.0, ei = 0.0;
  int it = 0, m = 0;
  unsigned long long executed = 0, rebased = 0;
  const int N = a.N, mend = a.mend;
  for (;;) {
    unsigned long long need = __ballot(!act);
    while (need && !exhausted) {
      if (cpos >= cend) {
        uint32_t b = 0;
        if (lane_id() == 0) b = atomicAdd(work, (unsigned int)kChunk);
        b = __shfl(b, 0);
        cpos = b;
        cend = min(b + (uint32_t)kChunk, total);
        if (cpos >= total) {
          exhausted = true;
          break;
        }
      }
      uint32_t avail = cend - cpos;
      uint32_t nneed = (uint32_t)__popcll(need);
      uint32_t take = min(nneed, avail);
      uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane_id()) - 1ull));
      if (!act && rank < take) {
        uint32_t slot = cpos + rank;
        pix = a.active[slot];
        double4 s = a.state[slot];
        int2 si = a.state_i[slot];
        dr = s.x;
        di = s.y;
        er = s.z;
        ei = s.w;
        it = si.x;
        m = si.y;
        act = true;
      }
      cpos += take;
      need = __ballot(!act);
    }
    if (__ballot(act) == 0) break;
    if (!act) {  // idle lanes run the block on a harmless state
      dr = di = er = ei = 0.0;
      m = 0;
    }
    const bool fast = __ballot(act && (it + K > N || m + K > mend)) == 0;
    const double dr0 = dr, di0 = di;
    const int m0 = m, it0 = it;
    bool replay = act;
    if (fast) {
      double mag = 0.0;
      unsigned long long rb = 0;
#pragma unroll
      for (int j = 0; j < K; j++) {
        double2 z = a.Zd[m];
        double zr = z.x, zi = z.y;
        double yr = zr + dr, yi = zi + di;
        mag = fma(yr, yr, yi * yi);
        double dmag = fma(dr, dr, di * di);
        if (mag < dmag) {
          dr = yr;
          di = yi;
          zr = 0.0;
          zi = 0.0;
          m = 0;
          rb++;
        }
        double wr = zr + yr, wi = zi + yi;
        double ndr = fma(dr, wr, fma(-di, wi, er));
        double ndi = fma(dr, wi, fma(di, wr, ei));
        dr = ndr;
        di = ndi;
        m++;
      }
      it += K;
      replay = act && !(mag <= a.b2m);
      if (act && !replay) {
        executed += K;
        rebased += rb;
        if (it >= N) {  // tops up to N-1 did not escape
          const DevFrame& f = a.frames[find_frame(a.frames, a.nframes, pix)];
          a.out[f.out_base + (pix - f.sample_base)] = EscapeValue{N, 0.0f};
          act = false;
        }
      }
    }
    if (replay) {
      dr = dr0;
      di = di0;
      m = m0;
      it = it0;
      bool done = false;
      int count = N;
      float sm = 0.0f;
      for (int j = 0; j < K; j++) {
        double2 z = a.Zd[m];
        double zr = z.x, zi = z.y;
        double yr = zr + dr, yi = zi + di;
        double r2 = yr * yr + yi * yi;  // sqMag, as bailedOut does
        executed++;
        if (r2 > a.b2) {
          done = true;
          count = it;
          sm = smooth_of(r2, a.log_b);
          break;
        }
        if (it == N - 1) {
          done = true;
          break;
        }
        double mag = fma(yr, yr, yi * yi);
        double dmag = fma(dr, dr, di * di);
        if (mag < dmag || m == mend) {
          dr = yr;
          di = yi;
          zr = 0.0;
          zi = 0.0;
          m = 0;
          rebased++;
        }
        double wr = zr + yr, wi = zi + yi;
        double ndr = fma(dr, wr, fma(-di, wi, er));
        double ndi = fma(dr, wi, fma(di, wr, ei));
        dr = ndr;
        di = ndi;
        m++;
        it++;
      }
      if (done) {
        const DevFrame& f = a.frames[find_frame(a.frames, a.nframes, pix)];
        a.out[f.out_base + (pix - f.sample_base)] = EscapeValue{count, sm};
        act = false;
      }
    }
  }
  wave_add(&a.cnt[C_EXECUTED], executed);
  wave_add(&a.cnt[C_REBASED], rebased);
}

// ============================================================ grid utilities
// at(r, c, sc) over the fine grid (mandelbrot.cpp:320-332): float sum in r1-outer,
// c1-inner order, then the integer part and the fraction.
__global__ void k_downsample(const EscapeValue* fine, int nr, int nc, int sc, EscapeValue* out) {
  const int onr = nr / sc, onc = nc / sc;
  const long total = (long)onr * onc;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    int r = (int)(i / onc), c = (int)(i - (long)r * onc);
    float sum = 0.0f;
    for (int r1 = 0; r1 < sc; r1++)
      for (int c1 = 0; c1 < sc; c1++) {
        EscapeValue e = fine[(long)(sc * r + r1) * nc + (sc * c + c1)];
        sum += (float)e.iterations + e.smoothing;
      }
    sum /= (float)(sc * sc);
    int iv = (int)sum;
    out[i] = EscapeValue{iv, sum - (float)iv};
  }
}

// sum of reported iterations over an output range (credited work)
__global__ void k_sum_iterations(const EscapeValue* out, unsigned long long n, unsigned long long* dst) {
  unsigned long long s = 0;
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    s += (unsigned long long)out[i].iterations;
  wave_add(dst, s);
}

__global__ void k_zero(unsigned long long* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ================================================================== engine
#define NM_HIP(expr)                                   \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return fail(#expr, e_);      \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
};

class EngineImpl : public Engine {
 public:
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  int device = 0;
  int num_cu = 256;
  std::string err;
  // persistent buffers (grow-only)
  void* d_frames = nullptr; size_t n_frames_bytes = 0;
  void* d_tabs = nullptr; size_t n_tabs_bytes = 0;
  void* d_mask = nullptr; size_t n_mask_bytes = 0;
  void* d_out = nullptr; size_t n_out_bytes = 0;
  void* d_active = nullptr; size_t n_active_bytes = 0;
  void* d_state = nullptr; size_t n_state_bytes = 0;
  void* d_state_i = nullptr; size_t n_state_i_bytes = 0;
  void* d_ref = nullptr; size_t n_ref_bytes = 0;
  void* d_grid = nullptr; size_t n_grid_bytes = 0;  // downsample staging
  unsigned long long* d_cnt = nullptr;
  uint64_t ref_serial = 0;
  // reference layout (offsets into d_ref)
  int ref_len = 0;
  const double2 *rZd = nullptr, *rXlo = nullptr, *rA = nullptr, *rB = nullptr, *rC = nullptr;
  const double *rPM = nullptr, *rBX = nullptr, *rBN = nullptr, *rCX = nullptr, *rCN = nullptr;
  const int* rSP = nullptr;

  int fail(const char* what, hipError_t e) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return NM_ERR_HIP;
  }
  const char* last_error() const override { return err.c_str(); }

  int init() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
      err = "no HIP device";
      return NM_ERR_NODEVICE;
    }
    NM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    NM_HIP(hipGetDeviceProperties(&prop, device));
    num_cu = prop.multiProcessorCount;
    NM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& e : ev) NM_HIP(hipEventCreate(&e));
    NM_HIP(hipMalloc((void**)&d_cnt, sizeof(unsigned long long) * C_COUNT));
    return NM_OK;
  }

  int grow(void** p, size_t* have, size_t need) {
    if (need <= *have && *p) return NM_OK;
    if (*p) NM_HIP(hipFree(*p));
    *p = nullptr;
    *have = 0;
    size_t n = std::max(need, (size_t)256);
    NM_HIP(hipMalloc(p, n));
    *have = n;
    return NM_OK;
  }

  int set_reference(const SaReferenceHost& ref, uint64_t serial) override {
    if (serial == ref_serial && serial != 0) return NM_OK;
    const int len = ref.len;
    if (len < 1) return NM_ERR_ARG;
    // Zd padded with zeros so idle lanes can run a block from index 0
    const size_t nz = (size_t)len + 2 + 64;
    size_t off = 0;
    auto place = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    size_t oZ = place(nz * 16), oXlo = place((size_t)len * 16), oA = place((size_t)len * 16),
           oB = place((size_t)len * 16), oC = place((size_t)len * 16), oPM = place((size_t)len * 8),
           oBX = place((size_t)len * 8), oBN = place((size_t)len * 8), oCX = place((size_t)len * 8),
           oCN = place((size_t)len * 8), oSP = place((size_t)len * 4);
    std::vector<char> h(off, 0);
    std::memcpy(h.data() + oZ, ref.Zd.data(), ref.Zd.size() * 8);
    std::memcpy(h.data() + oXlo, ref.Xlo.data(), ref.Xlo.size() * 8);
    std::memcpy(h.data() + oA, ref.A.data(), ref.A.size() * 8);
    std::memcpy(h.data() + oB, ref.B.data(), ref.B.size() * 8);
    std::memcpy(h.data() + oC, ref.C.data(), ref.C.size() * 8);
    std::memcpy(h.data() + oPM, ref.PM.data(), ref.PM.size() * 8);
    std::memcpy(h.data() + oBX, ref.BX.data(), ref.BX.size() * 8);
    std::memcpy(h.data() + oBN, ref.BN.data(), ref.BN.size() * 8);
    std::memcpy(h.data() + oCX, ref.CX.data(), ref.CX.size() * 8);
    std::memcpy(h.data() + oCN, ref.CN.data(), ref.CN.size() * 8);
    std::memcpy(h.data() + oSP, ref.SP.data(), ref.SP.size() * 4);
    int st = grow(&d_ref, &n_ref_bytes, off);
    if (st) return st;
    NM_HIP(hipMemcpyAsync(d_ref, h.data(), off, hipMemcpyHostToDevice, stream));
    NM_HIP(hipStreamSynchronize(stream));
    char* b = (char*)d_ref;
    rZd = (const double2*)(b + oZ);
    rXlo = (const double2*)(b + oXlo);
    rA = (const double2*)(b + oA);
    rB = (const double2*)(b + oB);
    rC = (const double2*)(b + oC);
    rPM = (const double*)(b + oPM);
    rBX = (const double*)(b + oBX);
    rBN = (const double*)(b + oBN);
    rCX = (const double*)(b + oCX);
    rCN = (const double*)(b + oCN);
    rSP = (const int*)(b + oSP);
    ref_len = len;
    ref_serial = serial;
    return NM_OK;
  }

  int zero_counters() {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, d_cnt, (int)C_COUNT);
    NM_HIP(hipGetLastError());
    return NM_OK;
  }

  int read_counters(unsigned long long* h) {
    NM_HIP(hipMemcpyAsync(h, d_cnt, sizeof(unsigned long long) * C_COUNT, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    return NM_OK;
  }

  // frames + tables + output/active/state sizing shared by both paths
  struct Layout {
    std::vector<DevFrame> frames;
    uint64_t total = 0;
  };

  int start_grid() const { return num_cu * 16; }

  template <class Args, class StartFn, class IterFn>
  int run(Args& args, StartFn start, IterFn iter, int waves, const RenderOpts& opt, RenderStats* stats,
          uint64_t nout, EscapeValue* host_out) {
    const int repeats = std::max(1, opt.repeats);
    float ms_start = 0, ms_iter = 0, ms_total = 0;
    unsigned long long h[C_COUNT] = {};
    for (int rep = 0; rep < repeats; rep++) {
      int st = zero_counters();
      if (st) return st;
      NM_HIP(hipEventRecord(ev[0], stream));
      hipLaunchKernelGGL(start, dim3(start_grid()), dim3(256), 0, stream, args);
      NM_HIP(hipGetLastError());
      NM_HIP(hipEventRecord(ev[1], stream));
      hipLaunchKernelGGL(iter, dim3(num_cu * waves), dim3(256), 0, stream, args);
      NM_HIP(hipGetLastError());
      NM_HIP(hipEventRecord(ev[2], stream));
      NM_HIP(hipEventSynchronize(ev[2]));
      float t1 = 0, t2 = 0;
      NM_HIP(hipEventElapsedTime(&t1, ev[0], ev[1]));
      NM_HIP(hipEventElapsedTime(&t2, ev[1], ev[2]));
      ms_start = t1;
      ms_iter += t2;
      ms_total += t1 + t2;
    }
    hipLaunchKernelGGL(k_sum_iterations, dim3(num_cu * 4), dim3(256), 0, stream, (const EscapeValue*)d_out,
                       (unsigned long long)nout, d_cnt + C_CREDITED);
    NM_HIP(hipGetLastError());
    int st = read_counters(h);
    if (st) return st;
    if (host_out)
      NM_HIP(hipMemcpyAsync(host_out, d_out, nout * sizeof(EscapeValue), hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    if (stats) {
      stats->samples = nout;
      stats->n_cardioid = h[C_CARDIOID];
      stats->n_sa_resolved = h[C_SA_ESC];
      stats->n_iterated = h[C_ACTIVE];
      stats->executed = h[C_EXECUTED];
      stats->skipped = h[C_SKIPPED];
      stats->n_rebased = h[C_REBASED];
      stats->credited = h[C_CREDITED] - h[C_CARDIOID] * (unsigned long long)opt.N;
      stats->ms_start = ms_start;
      stats->ms_iterate = ms_iter / repeats;
      stats->ms_total = ms_total / repeats;
    }
    return NM_OK;
  }

  int render_hw(const std::vector<const HwFrameHost*>& frames, const RenderOpts& opt, EscapeValue* host_out,
                RenderStats* stats) override {
    if (frames.empty()) return NM_ERR_ARG;
    std::vector<DevFrame> df;
    std::vector<double> tabs;
    uint64_t total = 0;
    // table layout: [cre | crelo | cim | cimlo] concatenated per kind
    size_t ncol = 0, nrow = 0;
    for (auto* f : frames) {
      if (!f || f->nr <= 0 || f->nc <= 0) return NM_ERR_ARG;
      ncol += f->nc;
      nrow += f->nr;
    }
    std::vector<double> cre, crelo, cim, cimlo;
    cre.reserve(ncol); crelo.reserve(ncol); cim.reserve(nrow); cimlo.reserve(nrow);
    for (auto* f : frames) {
      DevFrame d{};
      d.sample_base = (uint32_t)total;
      d.nr = f->nr;
      d.nc = f->nc;
      d.tab_c = (uint32_t)cre.size();
      d.tab_r = (uint32_t)cim.size();
      d.card_mode = f->card_mode;
      d.out_base = total;
      d.mask_base = 0;
      cre.insert(cre.end(), f->cre.begin(), f->cre.end());
      crelo.insert(crelo.end(), f->crelo.begin(), f->crelo.end());
      cim.insert(cim.end(), f->cim.begin(), f->cim.end());
      cimlo.insert(cimlo.end(), f->cimlo.begin(), f->cimlo.end());
      total += (uint64_t)f->nr * f->nc;
      df.push_back(d);
    }
    if (total >= 0xFFFFFFF0ull) return NM_ERR_RANGE;
    int st;
    if ((st = grow(&d_frames, &n_frames_bytes, df.size() * sizeof(DevFrame)))) return st;
    size_t tb = (2 * ncol + 2 * nrow) * sizeof(double);
    if ((st = grow(&d_tabs, &n_tabs_bytes, tb))) return st;
    if ((st = grow(&d_out, &n_out_bytes, total * sizeof(EscapeValue)))) return st;
    if ((st = grow(&d_active, &n_active_bytes, total * sizeof(uint32_t)))) return st;
    double* t = (double*)d_tabs;
    NM_HIP(hipMemcpyAsync(d_frames, df.data(), df.size() * sizeof(DevFrame), hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(t, cre.data(), ncol * 8, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(t + ncol, crelo.data(), ncol * 8, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(t + 2 * ncol, cim.data(), nrow * 8, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(t + 2 * ncol + nrow, cimlo.data(), nrow * 8, hipMemcpyHostToDevice, stream));
    PlainArgs a;
    a.frames = (const DevFrame*)d_frames;
    a.nframes = (int)df.size();
    a.total = (uint32_t)total;
    a.cre = t;
    a.crelo = t + ncol;
    a.cim = t + 2 * ncol;
    a.cimlo = t + 2 * ncol + nrow;
    a.out = (EscapeValue*)d_out;
    a.active = (uint32_t*)d_active;
    a.cnt = d_cnt;
    a.N = opt.N;
    a.b2 = opt.bailout * opt.bailout;
    a.b2m = a.b2 * (1.0 - 1.0 / 1099511627776.0);
    a.log_b = std::log(opt.bailout);
    const int K = opt.block_k > 0 ? opt.block_k : 32;
    const int waves = opt.waves_per_simd > 0 ? opt.waves_per_simd : 4;
    if (stats) {
      stats->block_k = K;
      stats->waves_per_simd = waves;
    }
    // 256-thread blocks = 4 waves = one per SIMD: num_cu * waves blocks
    switch (K) {
      case 8: return run(a, k_plain_start, k_plain_iterate<8>, waves, opt, stats, total, host_out);
      case 16: return run(a, k_plain_start, k_plain_iterate<16>, waves, opt, stats, total, host_out);
      case 32: return run(a, k_plain_start, k_plain_iterate<32>, waves, opt, stats, total, host_out);
      case 64: return run(a, k_plain_start, k_plain_iterate<64>, waves, opt, stats, total, host_out);
      default: return NM_ERR_ARG;
    }
  }

  int render_sa(const std::vector<const SaFrameHost*>& frames, const RenderOpts& opt, EscapeValue* host_out,
                RenderStats* stats) override {
    if (frames.empty() || ref_len < 1) return NM_ERR_STATE;
    std::vector<DevFrame> df;
    size_t ncol = 0, nrow = 0, nmask = 0;
    for (auto* f : frames) {
      if (!f || f->nr <= 0 || f->nc <= 0) return NM_ERR_ARG;
      ncol += f->nc;
      nrow += f->nr;
      if (f->card_mode == CARD_MASK) nmask += f->mask.size();
    }
    std::vector<double> er, ei;
    std::vector<uint32_t> mask;
    er.reserve(ncol); ei.reserve(nrow); mask.reserve(nmask);
    uint64_t total = 0;
    for (auto* f : frames) {
      DevFrame d{};
      d.sample_base = (uint32_t)total;
      d.nr = f->nr;
      d.nc = f->nc;
      d.tab_c = (uint32_t)er.size();
      d.tab_r = (uint32_t)ei.size();
      d.card_mode = f->card_mode;
      d.out_base = total;
      d.mask_base = (uint64_t)mask.size() * 32;
      if (f->card_mode == CARD_MASK) {
        if (f->mask.size() * 32 < (size_t)f->nr * f->nc) return NM_ERR_ARG;
        mask.insert(mask.end(), f->mask.begin(), f->mask.end());
      }
      er.insert(er.end(), f->epsre.begin(), f->epsre.end());
      ei.insert(ei.end(), f->epsim.begin(), f->epsim.end());
      total += (uint64_t)f->nr * f->nc;
      df.push_back(d);
    }
    if (total >= 0xFFFFFFF0ull) return NM_ERR_RANGE;
    int st;
    if ((st = grow(&d_frames, &n_frames_bytes, df.size() * sizeof(DevFrame)))) return st;
    if ((st = grow(&d_tabs, &n_tabs_bytes, (ncol + nrow) * sizeof(double)))) return st;
    if ((st = grow(&d_mask, &n_mask_bytes, std::max<size_t>(mask.size(), 1) * 4))) return st;
    if ((st = grow(&d_out, &n_out_bytes, total * sizeof(EscapeValue)))) return st;
    if ((st = grow(&d_active, &n_active_bytes, total * sizeof(uint32_t)))) return st;
    if ((st = grow(&d_state, &n_state_bytes, total * sizeof(double4)))) return st;
    if ((st = grow(&d_state_i, &n_state_i_bytes, total * sizeof(int2)))) return st;
    double* t = (double*)d_tabs;
    NM_HIP(hipMemcpyAsync(d_frames, df.data(), df.size() * sizeof(DevFrame), hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(t, er.data(), ncol * 8, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(t + ncol, ei.data(), nrow * 8, hipMemcpyHostToDevice, stream));
    if (!mask.empty())
      NM_HIP(hipMemcpyAsync(d_mask, mask.data(), mask.size() * 4, hipMemcpyHostToDevice, stream));
    SeriesArgs a;
    a.frames = (const DevFrame*)d_frames;
    a.nframes = (int)df.size();
    a.total = (uint32_t)total;
    a.epsre = t;
    a.epsim = t + ncol;
    a.mask = (const uint32_t*)d_mask;
    a.out = (EscapeValue*)d_out;
    a.active = (uint32_t*)d_active;
    a.state = (double4*)d_state;
    a.state_i = (int2*)d_state_i;
    a.cnt = d_cnt;
    a.Zd = rZd;
    a.Xlo = rXlo;
    a.A = rA;
    a.B = rB;
    a.C = rC;
    a.PM = rPM;
    a.BX = rBX;
    a.BN = rBN;
    a.CX = rCX;
    a.CN = rCN;
    a.SP = rSP;
    a.len = ref_len;
    a.mend = ref_len + 1;
    a.N = opt.N;
    a.tol = opt.tol;
    a.b2 = opt.bailout * opt.bailout;
    a.b2m = a.b2 * (1.0 - 1.0 / 1099511627776.0);
    a.log_b = std::log(opt.bailout);
    const int K = opt.block_k > 0 ? opt.block_k : 16;
    const int waves = opt.waves_per_simd > 0 ? opt.waves_per_simd : 4;
    if (stats) {
      stats->block_k = K;
      stats->waves_per_simd = waves;
    }
    switch (K) {
      case 8: return run(a, k_series_start, k_series_iterate<8>, waves, opt, stats, total, host_out);
      case 16: return run(a, k_series_start, k_series_iterate<16>, waves, opt, stats, total, host_out);
      case 32: return run(a, k_series_start, k_series_iterate<32>, waves, opt, stats, total, host_out);
      default: return NM_ERR_ARG;
    }
  }

  int downsample(const EscapeValue* fine, int nr, int nc, int sc, EscapeValue* out) override {
    if (sc <= 0 || nr <= 0 || nc <= 0) return NM_ERR_ARG;
    size_t nin = (size_t)nr * nc, nout = (size_t)(nr / sc) * (nc / sc);
    int st;
    if ((st = grow(&d_grid, &n_grid_bytes, (nin + nout) * sizeof(EscapeValue)))) return st;
    EscapeValue* din = (EscapeValue*)d_grid;
    EscapeValue* dout = din + nin;
    NM_HIP(hipMemcpyAsync(din, fine, nin * sizeof(EscapeValue), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_downsample, dim3(num_cu * 4), dim3(256), 0, stream, (const EscapeValue*)din, nr, nc, sc,
                       dout);
    NM_HIP(hipGetLastError());
    NM_HIP(hipMemcpyAsync(out, dout, nout * sizeof(EscapeValue), hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    return NM_OK;
  }
};

static std::mutex g_engine_mu;
static EngineImpl* g_engine = nullptr;
static int g_engine_status = 1;

int Engine::instance(Engine** out) {
  std::lock_guard<std::mutex> lk(g_engine_mu);
  if (g_engine_status == 1) {
    EngineImpl* e = new EngineImpl();
    int st = e->init();
    if (st != NM_OK) {
      g_engine_status = st;
      static std::string msg;
      msg = e->err;
      delete e;
      *out = nullptr;
      return st;
    }
    g_engine = e;
    g_engine_status = NM_OK;
  }
  *out = g_engine;
  return g_engine_status;
}

}  // namespace nm
