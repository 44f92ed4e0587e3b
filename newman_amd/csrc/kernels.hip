// kernels.hip - gfx950 kernels of the per-pixel deep-zoom iteration and the
// device engine that drives them.
//
// Layout in HBM (one frame set per render):
//   frames[]        DevFrame per frame (sample base, size, sample order, table
//                   offsets, output base)
//   col/row tables  plain path: truncated coordinate + double-double remainder;
//                   series path: epsilon = descend(Y0 - X[0]) per column / row
//   out[]           EscapeValue {int32, float} per sample, in sample order (the
//                   grid.h layout itself for row-major orders; k_emit scatters
//                   tiled orders into the row-major grids)
//   hand[]          u64 per 64 samples: bit l = sample handed to the iteration
//                   kernel (one word per wave and sample slot of the start kernels)
//   state[]         per sample: delta (or z) and epsilon (or c), 4 x f64; it, m
//   r2[]            |Y|^2 at the escape of iterated samples (smoothing input)
//   active[]        u32 sample ids that iterate, compacted in sample order
//   reference block Zd (the orbit as f64 pairs, Zd[0] = 0, padded past its end),
//                   Xlo, A, B, C and the stop-search tables
//
// Pipeline per render pass, with no contended device-scope atomic on the hot
// path (one contended counter serialises at ~10 ns per atomic on this part; the
// start kernels add one count per wave and tile, four adds per counter):
//   k_zero      clears the per-workgroup counter slots and the tile counts (a
//               kernel, so a captured HIP graph holds no memset node)
//   *_start     tiles of 1024 samples: cardioid/bulb test and up to pre_iter
//               updates (plain path), or the reference's per-sample series start
//               (series path: stop index, series skip, escape inside the skip,
//               first step); resolves what it can, marks the rest (one ballot
//               word per wave), counts per tile
//   scan        exclusive scan of the tile counts (range totals, then each
//               range's scan from the earlier ranges' sum; up to 64 workgroups)
//   scatter     compacts the marked sample ids into active[] in sample order
//   *_iterate   persistent waves, one sample per lane; blocks of K steps; the
//               first escape of a lane in a block records its step and |z|^2
//               (FP32X: and poisons the lane with NaN); lanes are refilled (ballot +
//               popcount ranks) from 64-sample chunks of active[], which waves
//               claim from 8 claim ranges, with the next chunk prefetched (plain:
//               its states; series: its ids); the series kernel keeps the orbit's head and
//               tail windows in LDS
//   finalize    smoothing of the iterated samples from r2
//   emit        tiled sample orders only: results into the row-major grids
// Counters go to per-workgroup partial slots reduced on the device.
// This file is compiled with -ffp-contract=off: the plain path must reproduce the
// reference's double arithmetic bit for bit (complex.h:19-31, makefile:3 has no
// FMA); the perturbation paths write their FMAs explicitly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "colour.hpp"
#include "engine.hpp"
#include "sample_index.hpp"
#include "video.hpp"

namespace nm {

// ------------------------------------------------------------ device structs
// (DevFrame, the sample orders and the start tiling: sample_index.hpp)

enum CounterIdx {
  C_CARDIOID = 0,
  C_SA_RES = 1,
  C_EXECUTED = 2,
  C_SKIPPED = 3,
  C_REBASED = 4,
  C_CREDITED = 5,
  C_EXEC_START = 6,
  C_COUNT = 7
};

constexpr int kHead = 64;               // orbit padding past its end (>= max K)
// orbit windows each iterate workgroup keeps in LDS: the head (where rebased
// lanes live) and the tail (where samples that never rebased run out the
// reference); lanes whose block fits a window read it from LDS, the others
// gather from global memory, whose cache tag lookups bound the gathers
constexpr int kLdsHead = 1024, kLdsTail = 1024;
// The series iterate kernel runs 8 waves per workgroup: the 32 KB of LDS orbit
// windows is shared by 8 waves, so LDS does not cap the waves a CU holds below
// what the registers allow (8 per SIMD at 64 VGPRs: four workgroups, 128 KB).
constexpr int kSeriesIterThreads = 512;
static_assert(kLdsTail >= 64 && kLdsHead >= 64, "a block of up to 64 steps reads inside one window");
constexpr int kPartSlots = 10240;       // per-workgroup partial counter slots

// partial-slot ranges of the kernels of one render
constexpr int kSlotStart = 0;
constexpr int kSlotIter = 4096;
constexpr int kSlotFinal = 6144;

struct Common {
  const DevFrame* frames;
  int nframes;
  uint32_t frame_size;  // samples per frame (all frames of a set have one size)
  double inv_frame;     // 1.0 / frame_size
  uint32_t total;
  EscapeValue* out;
  double* r2;  // |z|^2 at the escape of iterated samples (per sample)
  double4* state;  // per sample: plain (x, y, cx, cy); series: (dr, di) as double2[]
  int2* state_i;   // per sample, series path: (it, m)
  uint64_t* hand;  // hand-on bits (hand_word)
  uint32_t* tile_count;
  uint32_t* active;
  const uint32_t* nactive;  // written by the scan
  unsigned long long* part;
  int N;
  double b2, b2m, log_b;
  double smooth_k;  // 0.5 ln 2 / ln(bailout), or 0 for a bailout below 2
};

struct PlainArgs {
  Common c;
  int pre_iter;  // updates tried in the start kernel before a sample is handed on
  int k_start;   // updates every handed-on sample has done: pre_iter rounded up to the start kernel's blocks of 4
  const double* cre;
  const double* crelo;
  const double* cim;
  const double* cimlo;
};

struct SeriesArgs {
  Common c;
  const double* epsre;
  const double* epsim;
  const uint32_t* mask;
  // reference
  const double2* Zd;   // len + 2 (+ padding)
  const double2* Xlo;  // len
  const double2* A;
  const double2* B;
  const double2* C;
  const double* PM;
  const double* BX;
  const double* BN;
  const double* CX;
  const double* CN;
  const int* SP;
  const double* ST;   // range-minimum sparse table of the thresholds (st_levels x len)
  const double* BNs;  // suffix minima of log2 |B_j|^2, |C_j|^2
  const double* CNs;
  const int* PMT;  // stop-search lookup over the finite PM range (pmt_n + 1 entries)
  double pmt_base, pmt_inv_h;
  int pmt_n;
  int st_levels;
  int len;
  int mend;  // len + 1
  int zd_count;  // entries allocated in Zd (past len + 1: zero padding)
  double tol;
  double ltol;  // log2 tol
  // per-tile series skip (the first unstable index of each tile's probe); null:
  // every sample searches its own (the reference's per-pixel scan)
  const int* tile_ua;
  uint32_t skip_T, skip_ntc, skip_tpf;  // tile side, tiles per frame row, tiles per frame
  // floatexp series (fxs != 0): A, B, C as mantissas (double2) and exponents
  // (int2) per index; the thresholds above are then built from them
  const double2 *Am, *Bm, *Cm;
  const int2 *Ae, *Be, *Ce;
  int fxs;
  uint32_t run_shift;  // active-list chunks (ChunkMap): runs of 2^run_shift entries
};

// ------------------------------------------------------------ device helpers
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }


// log2 q for a positive normal q, within 3e-15 absolute (measured against a
// long-double log2 on 5e7 points of [2^-20, 2^20]): q = m 2^E with m in
// [1/sqrt 2, sqrt 2), log2 m = (2 / ln 2) atanh(s), s = (m - 1) / (m + 1),
// |s| < 0.1716, by the odd series through s^17 (the next term is below 4e-16).
// The stop search compares the result with thresholds through a margin of
// 4e-9, so this bound, not correct rounding, is what it needs; the smoothing
// stores its result as a float.
__device__ __forceinline__ double log2_near(double q) {
  int E;
  double m = frexp(q, &E);  // [0.5, 1)
  if (m < 0.70710678118654752) {
    m *= 2.0;
    E -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 17.0;
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  p = fma(p, s2, 1.0);
  return fma(2.8853900817779268147 * s, p, (double)E);  // 2 / ln 2
}

// log2 q for a positive normal q by a table: q = m 2^E with m in [1, 2); the top
// seven fraction bits of m pick r_i, a double near 1 / c_i for the centre c_i of
// m's 1/128-wide interval, and l_i = -log2 r_i (rounded from a long double
// evaluation on the host, d_log2tab); then f = m r_i - 1 (one rounding by the
// fma, |f| <= 2^-8) and log2 m = l_i + log2(1 + f), the series through f^7 (the
// next term is below 2^-66).  The error is that of l_i and of the final
// additions: below 2^-52 + ulp(E + l_i), 3.6e-15 over [2^-20, 2^20] (measured
// against a long-double log2 on 2e7 points).  About 18 VALU instructions where
// log2_near's division costs about 30.
constexpr int kLog2Tab = 128;
__constant__ double2 d_log2tab[kLog2Tab];  // (r_i, l_i), filled at engine start
template <class Tab>
__device__ __forceinline__ double log2_tab(double q, Tab tab) {
  const long long b = __double_as_longlong(q);
  const int E = (int)((b >> 52) & 0x7ff) - 1023;
  const double m = __longlong_as_double((b & 0x000fffffffffffffLL) | 0x3ff0000000000000LL);
  const double2 t = tab[(int)((b >> 45) & (kLog2Tab - 1))];
  const double f = fma(m, t.x, -1.0);
  // (-1)^(k+1) / (k ln 2), k = 7 .. 1
  double p = 0.2060992915555662;
  p = fma(p, f, -0.2404491734814939);
  p = fma(p, f, 0.28853900817779266);
  p = fma(p, f, -0.36067376022224085);
  p = fma(p, f, 0.4808983469629878);
  p = fma(p, f, -0.7213475204444817);
  p = fma(p, f, 1.4426950408889634);
  return ((double)E + t.y) + f * p;
}

// mandelbrot.cpp:133-136: 1 - log2(0.5 ln r2 / ln b), with 0.5 ln r2 / ln b =
// log2(r2) * (0.5 ln 2 / ln b) (smooth_k, from the host) and both base-2
// logarithms by log2_tab: within 1e-15 of the formula evaluated in long double
// (measured on 2e7 radii at bailout 1024; the stored float differed from the
// correctly rounded one once), far below the float the value is stored in.  An
// infinite r2, or a bailout below 2 (smooth_k = 0), takes the library logarithms.
// The table is d_log2tab itself or a workgroup's LDS copy of it (load_log2tab):
// a wave waiting for an LDS read does not also wait for its outstanding stores,
// as a wait for a global load does (one in-order counter covers both).
template <class Tab = const double2*>
__device__ __forceinline__ float smooth_of(double r2, const Common& c, Tab tab = d_log2tab) {
  if (c.smooth_k == 0.0 || !(r2 < INFINITY)) return (float)(1.0 - log2(0.5 * log(r2) / c.log_b));
  return (float)(1.0 - log2_tab(log2_tab(r2, tab) * c.smooth_k, tab));
}

// the workgroup's LDS copy of d_log2tab (all threads call it; it ends with a barrier)
__device__ __forceinline__ void load_log2tab(double2* lds) {
  for (int i = threadIdx.x; i < kLog2Tab; i += blockDim.x) lds[i] = d_log2tab[i];
  __syncthreads();
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Adds v (per thread) into the workgroup's partial slot: wave reduction, LDS
// accumulation, one plain store per workgroup.  Called by all threads.
__device__ __forceinline__ void block_partial(unsigned long long* part, int slot, const unsigned long long* v,
                                              const int* idx, int n) {
  __shared__ unsigned long long acc[C_COUNT];
  if (threadIdx.x < C_COUNT) acc[threadIdx.x] = 0ull;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    unsigned long long s = wave_sum(v[k]);
    if (lane_id() == 0 && s) atomicAdd(&acc[idx[k]], s);
  }
  __syncthreads();
  if (threadIdx.x < C_COUNT) part[(size_t)slot * C_COUNT + threadIdx.x] += acc[threadIdx.x];
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
// getIterationsHW (mandelbrot.cpp:231-254), start kernel: the cardioid/bulb
// test, then up to kPreIter updates per sample with the exact escape test (most
// samples of a shallow view escape here, converged); the rest are marked for
// the persistent kernel with their state (z, c, updates done).
constexpr int kPreIterDefault = 16;

// One update z <- sq(z) + z0 (complex.h:19-27) with the reference's rounding:
// x' = (r2 - i2) + cx, y' = (2x) y + cy, r2 = x'^2, i2 = y'^2.  FMA2 computes y'
// as fma(x y, 2, cy): 2x is exact and fl((2x) y) = 2 fl(x y) whenever x y is 0
// or a normal double, so the single rounding of fma(p, 2, cy) = fl(2p + cy) is
// the reference's fl(fl((2x) y) + cy), bit for bit.  x y is 0 or normal before
// escape when every coordinate of the frame is 0 or at least 2^-200 in
// magnitude (a nonzero sum of two doubles is at least 2^-54 times the smaller
// one's binade, so |y| >= 2^-254 unless y = 0 and |x| >= 2^-560 unless x = 0)
// and the bailout is at most 2^400; the host checks both (fma_ok) and runs the
// three-operation form otherwise.  Seven FP64 instructions instead of eight for
// the same eight algorithmic flops.
template <bool FMA2>
__device__ __forceinline__ void plain_update(double& x, double& y, double& r2, double& i2, const double cx,
                                             const double cy) {
  if (FMA2) {
    const double p = x * y;
    y = fma(p, 2.0, cy);
  } else {
    const double t = x + x;
    y = t * y + cy;
  }
  x = r2 - i2 + cx;
  r2 = x * x;
  i2 = y * y;
}

// The main cardioid and the period-2 disk lie inside x in [-1.25, 0.375],
// |y| <= 0.65: the cardioid e^{it}/2 - e^{2it}/4 reaches Re 3/8 at t = pi/3 and
// |Im| 3 sqrt(3) / 8 = 0.6495 at t = 2 pi/3, the disk spans [-1.25, -0.75].
// Samples outside the box below skip the test (the HP remainders, below
// 2^-52 |x|, do not reach the margins).
__device__ __forceinline__ bool card_box(double x, double y) { return x > -1.3 && x < 0.42 && fabs(y) < 0.7; }

// mandelbrot.cpp:63-71 in double inside that box, with fixed error bounds: +1 in,
// -1 out, 0 when double-double must decide.  In the box |x - 1/4| <= 1.55,
// q <= 2.9, |q (q + x - 1/4)| + y^2 / 4 <= 13.2, so the computed cardioid
// polynomial errs by at most 8 u 13.2 < 1.2e-14, and the HP remainders of the
// coordinates (below 2.2e-16 for x, 1.1e-16 for y) move it by at most
// 25.6 |xl| + 10.6 |yl| < 7e-15; the disk polynomial errs by at most 2e-15 in
// all.  The bounds used, 1e-12 and 1e-13, leave a margin of 50.
constexpr double kCardBound = 1e-12, kDiskBound = 1e-13;
__device__ __forceinline__ int cardioid_box_filter(double x, double y) {
  const double xm = x - 0.25, y2 = y * y;
  const double q = xm * xm + y2;
  const double f = q * (q + xm) - 0.25 * y2;
  if (f < -kCardBound) return 1;
  if (!(f > kCardBound)) return 0;
  const double xp = x + 1.0;
  const double g = xp * xp + y2 - 0.0625;
  if (g < -kDiskBound) return 1;
  if (g > kDiskBound) return -1;
  return 0;
}

// Pre-iteration of the start kernel: a lane's state, its own flag and the
// wave's scalar mask of lanes still iterating.  A lane that does not iterate
// (cardioid, invalid) holds z = c = 0, which stays 0; a lane that escaped
// records its escape and keeps updating unrecorded (its values overflow to inf
// or NaN), so the escape test is masked by the wave's mask of lanes still
// iterating (a scalar AND, instead of parking the lane's six doubles).
struct PreLane {
  double &x, &y, &r2, &i2, &cx, &cy;
  unsigned long long& live;
  bool& alive;
  int& kesc;      // update whose test first held (-1: none yet)
  double& mgesc;  // |z|^2 there
};

__device__ __forceinline__ void park(double& x, double& y, double& r2, double& i2, double& cx, double& cy) {
  x = y = r2 = i2 = cx = cy = 0.0;
}

// The escape test of a group of four updates.  For bailout radius R, 2 <= R <=
// 2^30, and every |c| <= 0.999 (R^2 - R): when the test of an update holds
// (computed |z_j|^2 > R^2), that of the next update holds too.  |z_j|^2 >
// R^2 (1 - 2^-51) exactly, so the computed |z_{j+1}| is at least
// |z_j|^2 (1 - 2^-49) - |c| (1 + 2^-49) >= R + 0.001 R (R - 1) - 2^-48 R^2 >=
// R + 0.0004 R^2, and its computed square exceeds R^2.  The test of the update
// before the group did not hold, so |z| <= 2 R^2, 5 R^4, 26 R^8, 677 R^16 over
// the group's four updates and every value stays below 2^980 (no overflow, no
// NaN).  So a group tests only its last update's |z|^2 and, on an escape there,
// finds the first escaping update from the earlier updates' r2 + i2, the
// reference's sqMag of each bit for bit (their r2 and i2 stay in registers):
// three adds and three compare-and-branches less per four updates.  A lane that
// escaped inside the group runs the group's remaining updates unrecorded.
inline bool grouped_test_ok(double radius, double c_bound) {
  return radius >= 2.0 && radius <= 0x1p30 && c_bound <= 0.999 * (radius * radius - radius);
}

// G updates and their escape test: G = 1 tests every update as the reference
// does; G = 2 or 4 tests only the group's last |z|^2 (grouped_test_ok).  When
// the test holds for some lane of the wave mask `live`, the wave calls
// esc(bool e, int j, double mg) once: e is the lane's test, j (0..G-1) its
// group's first escaping update and mg that update's |z|^2, r2 + i2 as the
// reference's sqMag (mandelbrot.cpp:245), kept in registers for the earlier
// updates (valid where e holds on a lane of `live`).
template <int G, bool FMA2, class Esc>
__device__ __forceinline__ void update_group(double& x, double& y, double& r2, double& i2, const double cx,
                                             const double cy, const double b2, const unsigned long long live,
                                             Esc&& esc) {
  static_assert(G == 1 || G == 2 || G == 4, "groups of 1, 2 or 4 updates");
  double r2s[G], i2s[G];
#pragma unroll
  for (int t = 0; t < G; t++) {
    plain_update<FMA2>(x, y, r2, i2, cx, cy);
    r2s[t] = r2;
    i2s[t] = i2;
  }
  const double md = r2 + i2;
  const bool e = md > b2;
  if (__builtin_expect((__ballot(e) & live) != 0, 0)) {
    int j = G - 1;
    double mg = md;
#pragma unroll
    for (int t = G - 2; t >= 0; t--) {
      const double m = r2s[t] + i2s[t];
      if (m > b2) {
        j = t;
        mg = m;
      }
    }
    esc(e, j, mg);
  }
}

// Updates k0 + J .. k0 + 3 of a lockstep block of four in groups of G
// (unrolled by recursion).
template <int J, int G, bool FMA2>
__device__ __forceinline__ void pre_block(PreLane& L, const double b2, const int k0) {
  if constexpr (J < 4) {
    update_group<G, FMA2>(L.x, L.y, L.r2, L.i2, L.cx, L.cy, b2, L.live, [&](bool e, int j, double mg) {
      const bool ne = e && L.alive;
      L.live &= ~__ballot(ne);
      if (ne) {
        L.kesc = k0 + J + j;
        L.mgesc = mg;
        L.alive = false;
      }
    });
    pre_block<J + G, G, FMA2>(L, b2, k0);
  }
}

// The plain iterate kernel's default persistent waves per SIMD.
constexpr int kPlainWaves = 4;

// Group sizes of the grouped escape test: the start kernel's pre-iteration,
// where most lanes escape within a few updates, and the iteration kernel.
constexpr int kStartGroup = 4;
constexpr int kIterGroup = 4;

// The frame record at a wave-uniform index, through the constant address space
// (the records are written before the launch and only read by it), so the
// fields load once into scalar registers.
__device__ __forceinline__ DevFrame uniform_frame(const DevFrame* frames, uint32_t idx) {
#if __HIP_DEVICE_COMPILE__
  typedef __attribute__((address_space(4))) const uint32_t ConstWord;
  const ConstWord* w = (const ConstWord*)(frames + __builtin_amdgcn_readfirstlane(idx));
  uint32_t v[sizeof(DevFrame) / 4];
#pragma unroll
  for (unsigned i = 0; i < sizeof(DevFrame) / 4; i++) v[i] = w[i];
  DevFrame f;
  __builtin_memcpy(&f, v, sizeof f);
  return f;
#else
  return frames[idx];
#endif
}

// Per-lane counters of the start kernel, in 32 bits: a sample adds at most
// pre_iter + 3 <= 1027 updates or credited iterations (the kernel resolves only
// samples whose count is decided within its lockstep updates), and a lane takes
// at most 4 samples of each of its workgroup's tiles, at most 2^22 / 16 tiles
// (the grid is at least 16 workgroups when there are 16 tiles), so a lane's sums
// stay below 2^20 * 1027 < 2^31.
struct StartCounts {
  uint32_t ncard = 0, nres = 0, executed = 0, credited = 0;
};

// One sample of the start kernel (getIterationsHW, mandelbrot.cpp:231-254): the
// lane's sample is frame-local index gid - f.sample_base of frame f when valid.
// Every lane of the wave runs this (no early return), so the pre-iteration is
// wave-uniform control flow.  Returns whether the sample is handed on.
template <bool FMA2, bool GROUPED>
__device__ __forceinline__ bool start_sample(const PlainArgs& a, const DevFrame& f, const uint32_t gid, bool valid,
                                             const int pre, const double2* ltab, StartCounts& n) {
  const Common& c = a.c;
  const int N = c.N;
  bool in = false;
  double cx = 0.0, cy = 0.0;
  if (valid) {
    const Loc lc = locate(f, gid - f.sample_base);
    valid = lc.valid;
    if (valid) {
      cx = a.cre[f.tab_c + lc.c];
      cy = a.cim[f.tab_r + lc.r];
      if (f.card_mode == CARD_ALL) in = true;
      else if (f.card_mode == CARD_TEST_DD && card_box(cx, cy)) {
        int fc = cardioid_box_filter(cx, cy);
        if (fc == 0) {
          const double xl = a.crelo[f.tab_c + lc.c], yl = a.cimlo[f.tab_r + lc.r];
          fc = dd_in_cardioid(cx, xl, cy, yl) ? 1 : -1;
        }
        in = fc > 0;
      }
    }
  }
  // z <- sq(z) + z0 in the reference's operation order (complex.h:19-27), in
  // lockstep blocks of four updates while some lane of the wave iterates
  int kesc = -1;
  bool alive = valid && !in;
  double zx = alive ? cx : 0.0, zy = alive ? cy : 0.0;  // lanes not iterating hold z = c = 0
  double x = zx, y = zy, r2 = x * x, i2 = y * y, mgesc = 0.0;
  unsigned long long live = __ballot(alive);
  int kend = 0;
  PreLane L{x, y, r2, i2, zx, zy, live, alive, kesc, mgesc};
  while (live != 0 && kend < pre) {
    pre_block<0, GROUPED ? kStartGroup : 1, FMA2>(L, c.b2, kend);
    kend += 4;
  }
  bool want = false;
  if (valid && in) {
    c.out[gid] = EscapeValue{N, 0.0f};
    n.ncard++;
  } else if (valid) {
    // the reference tests k >= N before each update: an escape counts
    // only below N; a lane still iterating after kend updates is handed on
    if (kesc >= 0 && kesc < N) {
      n.executed += (uint32_t)(kesc + 1);
      c.out[gid] = EscapeValue{kesc, smooth_of(mgesc, c, ltab)};
      n.nres++;
      n.credited += (uint32_t)kesc;
    } else if (kesc >= N || kend >= N) {
      n.executed += (uint32_t)N;
      c.out[gid] = EscapeValue{N, 0.0f};
      n.nres++;
      n.credited += (uint32_t)N;
    } else {
      n.executed += (uint32_t)kend;
      // kend == k_start for every lane handed on (the lockstep loop ran to
      // the first multiple of 4 at or past pre_iter)
      c.state[gid] = make_double4(x, y, cx, cy);
      want = true;
    }
  }
  return want;
}

// The wave's hand-on word of a sample slot (every lane calls it); returns its
// count of marked samples (wave-uniform).
__device__ __forceinline__ uint32_t store_hand(uint64_t* hand, uint32_t gid, bool want) {
  const unsigned long long w = __ballot(want);
  if (lane_id() == 0) hand[hand_word(gid)] = w;
  return (uint32_t)__popcll(w);
}

// A wave's count of marked samples of a tile into tile_count[tile] (cleared by
// k_zero): one atomic add per wave and tile, no barrier between the waves.
__device__ __forceinline__ void tile_count_add(uint32_t* tile_count, uint32_t tile, uint32_t wave_count) {
  if (lane_id() == 0 && wave_count) atomicAdd(&tile_count[tile], wave_count);
}

template <bool FMA2, bool GROUPED>
__global__ __launch_bounds__(kThreads) void k_plain_start(PlainArgs a) {
  const Common& c = a.c;
  __shared__ double2 ltab[kLog2Tab];
  load_log2tab(ltab);
  StartCounts n;
  const uint32_t ntiles = start_tiles(c.total);
  const int pre = min(a.pre_iter, c.N);
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t mine = 0;
    uint32_t f0, f1;
    tile_frame_span(tile, c.total, c.inv_frame, &f0, &f1);
    if (f0 == f1) {  // the tile lies in one frame: its record in scalar registers
      const DevFrame f = uniform_frame(c.frames, f0);
      for (int j = 0; j < kPerThread; j++) {
        const uint32_t gid = start_gid(tile, j, threadIdx.x);
        const bool want = start_sample<FMA2, GROUPED>(a, f, gid, gid < c.total, pre, ltab, n);
        mine += store_hand(c.hand, gid, want);
      }
    } else {
      for (int j = 0; j < kPerThread; j++) {
        const uint32_t gid = start_gid(tile, j, threadIdx.x);
        const bool valid = gid < c.total;
        const DevFrame& f = c.frames[valid ? frame_of(gid, c.inv_frame) : 0u];
        const bool want = start_sample<FMA2, GROUPED>(a, f, gid, valid, pre, ltab, n);
        mine += store_hand(c.hand, gid, want);
      }
    }
    tile_count_add(c.tile_count, tile, mine);
  }
  const unsigned long long v[4] = {n.ncard, n.nres, n.executed, n.credited};
  const int idx[4] = {C_CARDIOID, C_SA_RES, C_EXEC_START, C_CREDITED};
  block_partial(c.part, kSlotStart + blockIdx.x, v, idx, 4);
}

// ============================================================ compaction
// exclusive scan of tile_count (in place), total -> *nactive, in two launches
// over the groups of scan_split: k_tile_sums writes each group's total of
// counts, then k_scan_tiles starts each group at the sum of the earlier groups'
// totals. A workgroup takes a round of 16384 consecutive counts at a time: 16
// per thread (four 16-byte loads), wave scans, a carry between rounds; the next
// round's loads are issued before this round's scan, so one memory latency is
// paid per round at most. (One workgroup over all of a 512-frame 640x480 set's
// 153,600 counts took 34 us, ten rounds back to back.)
constexpr int kScanPer = 16;
static_assert(kScanRound == 1024 * kScanPer, "a scan round is 16 counts for each of 1024 threads");
__device__ __forceinline__ void scan_load(const uint32_t* tc, uint32_t i0, uint32_t hi, uint32_t v[kScanPer]) {
  if (i0 + kScanPer <= hi) {
    const uint4* q = (const uint4*)(tc + i0);
#pragma unroll
    for (int k = 0; k < kScanPer / 4; k++) {
      const uint4 u = q[k];
      v[4 * k] = u.x;
      v[4 * k + 1] = u.y;
      v[4 * k + 2] = u.z;
      v[4 * k + 3] = u.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanPer; k++) v[k] = i0 + k < hi ? tc[i0 + k] : 0u;
  }
}

// group blockIdx.x's total of counts -> gsum[blockIdx.x]
__global__ __launch_bounds__(1024) void k_tile_sums(const uint32_t* tile_count, uint32_t ntiles, uint32_t per,
                                                     uint32_t* gsum) {
  __shared__ uint32_t wsum[16];
  uint32_t lo, hi;
  scan_group_range(ntiles, per, blockIdx.x, lo, hi);  // the group's counts [lo, hi)
  uint32_t s = 0;
  for (uint32_t base = lo; base < hi; base += kScanRound) {
    uint32_t v[kScanPer];
    scan_load(tile_count, base + kScanPer * threadIdx.x, hi, v);
#pragma unroll
    for (int k = 0; k < kScanPer; k++) s += v[k];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 16; w++) t += wsum[w];
    gsum[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t* tile_count, uint32_t ntiles, uint32_t per,
                                                      const uint32_t* gsum, uint32_t* nactive) {
  __shared__ uint32_t wsum[16];
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  uint32_t lo, hi;
  scan_group_range(ntiles, per, blockIdx.x, lo, hi);  // the group's counts [lo, hi)
  uint32_t carry = 0;
  for (uint32_t g = 0; g < blockIdx.x; g++) carry += gsum[g];
  uint32_t nxt[kScanPer];
  scan_load(tile_count, lo + kScanPer * threadIdx.x, hi, nxt);
  for (uint32_t base = lo; base < hi; base += kScanRound) {
    const uint32_t i0 = base + kScanPer * threadIdx.x;
    uint32_t v[kScanPer];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
      v[k] = nxt[k];
      mine += v[k];
    }
    if (base + kScanRound < hi) scan_load(tile_count, i0 + kScanRound, hi, nxt);
    uint32_t x = mine;  // inclusive wave scan of the thread totals
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < 16; w++) {
      const uint32_t t = wsum[w];
      before += w < wave ? t : 0u;
      all += t;
    }
    uint32_t run = carry + before + x - mine;
    if (i0 + kScanPer <= hi) {
      uint4* q = (uint4*)(tile_count + i0);
#pragma unroll
      for (int k = 0; k < kScanPer / 4; k++) {
        uint4 u;
        u.x = run;
        run += v[4 * k];
        u.y = run;
        run += v[4 * k + 1];
        u.z = run;
        run += v[4 * k + 2];
        u.w = run;
        run += v[4 * k + 3];
        q[k] = u;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanPer; k++) {
        if (i0 + k < hi) tile_count[i0 + k] = run;
        run += v[k];
      }
    }
    carry += all;
    __syncthreads();
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *nactive = carry;
}

// marked sample ids -> active[] in sample order: one wave per tile, lanes
// 0..15 holding its 16 hand-on words (one load each) and their exclusive prefix
// of set bits; then for each word (scalar) every lane writes its own sample's id
// at hand_position (consecutive lanes, consecutive positions).  A wave's next
// tile (words and offset) is loaded before it writes the current one, so the
// loads' latency overlaps the stores.
static_assert(kHandWordsPerTile == 16, "k_scatter: 16 words per tile, one per lane of lanes 0..15");
__device__ __forceinline__ void scatter_load(const uint64_t* hand, const uint32_t* tile_off, uint32_t tile,
                                             uint32_t ntiles, uint32_t lane, uint64_t& w, uint32_t& off) {
  w = 0ull;
  off = 0u;
  if (tile < ntiles) {  // wave-uniform
    if (lane < (uint32_t)kHandWordsPerTile) w = hand[tile * kHandWordsPerTile + lane];
    off = tile_off[tile];
  }
}
__global__ __launch_bounds__(kThreads) void k_scatter(const uint64_t* hand, const uint32_t* tile_off, uint32_t total,
                                                      uint32_t* active) {
  const uint32_t ntiles = start_tiles(total);
  const uint32_t lane = (uint32_t)lane_id();
  const uint32_t waves = gridDim.x * (kThreads / 64);
  uint32_t tile = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  uint64_t wnext;
  uint32_t offnext;
  scatter_load(hand, tile_off, tile, ntiles, lane, wnext, offnext);
  for (; tile < ntiles; tile += waves) {
    const uint64_t w = wnext;
    const uint32_t off = __builtin_amdgcn_readfirstlane(offnext);
    scatter_load(hand, tile_off, tile + waves, ntiles, lane, wnext, offnext);
    const uint32_t n = (uint32_t)__popcll(w);
    uint32_t x = n;  // inclusive scan over lanes 0..15
    for (int o = 1; o < kHandWordsPerTile; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    const uint32_t before = x - n;
    const uint32_t m = __builtin_amdgcn_readlane(x, kHandWordsPerTile - 1);  // the tile's marked samples
    if (m <= kScatterSparse) {  // wave-uniform: one store per 64 marked samples
      const uint32_t wlo = (uint32_t)w, whi = (uint32_t)(w >> 32);
      for (uint32_t p0 = 0; p0 < m; p0 += 64) {
        const uint32_t p = p0 + lane;
        int k = 0;  // the last word whose prefix is <= p (lanes 0..15 hold the prefixes)
#pragma unroll
        for (int step = kHandWordsPerTile / 2; step >= 1; step >>= 1)
          if ((uint32_t)__shfl((int)before, k + step) <= p) k += step;
        const uint32_t bk = (uint32_t)__shfl((int)before, k);
        const uint64_t wk = ((uint64_t)(uint32_t)__shfl((int)whi, k) << 32) | (uint32_t)__shfl((int)wlo, k);
        if (p < m) active[off + p] = tile * kTile + (uint32_t)k * 64 + select_bit(wk, p - bk);
      }
      continue;
    }
#pragma unroll 4
    for (int k = 0; k < kHandWordsPerTile; k++) {
      const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)w, k), hi = __builtin_amdgcn_readlane((uint32_t)(w >> 32), k);
      const uint64_t wk = ((uint64_t)hi << 32) | lo;
      if (wk == 0) continue;  // wave-uniform
      const uint32_t bk = __builtin_amdgcn_readlane(before, k);
      if ((wk >> lane) & 1ull) active[hand_position(off, bk, wk, lane)] = tile * kTile + (uint32_t)k * 64 + lane;
    }
  }
}

// ---- work pipeline of a persistent wave -------------------------------------
// A wave takes chunks of 64 entries of active[] (ChunkMap) by claims
// (ChunkClaim).  Each lane holds one entry of the chunks in flight: CUR
// (decoded: handed to lanes that need work through cross-lane shuffles), NEXT
// (its decode loads in flight) and, in the plain kernel, the chunk after (its
// sample-id load in flight), so a refill reads registers only.

// Chunk claims: persistent waves take chunks of the active list dynamically.
// The chunks are split into kQueues equal ranges with one claim counter each
// (each counter on its own 256-byte line; zeroed with the partial counters at
// the start of a pass).  A wave claims from its home range (its workgroup's
// index mod kQueues: consecutive workgroups land on different XCDs) and, once
// that is empty, from the others in turn.  Waves on a SIMD do not progress at
// one rate (the issue arbiter prefers older waves), so with a static chunk list
// per wave the waves finished anywhere between 30% and 100% of the kernel's
// time and the last ones ran nearly alone on their SIMDs; claiming lets fast
// waves take more.  One atomic per 64 samples; counters sharing one cache line
// had serialised the claims (~100 ns each).
constexpr int kQueueStride = 64;  // u32 between counters: 256 bytes
// d_part (u64 words): the partial counter slots, the claim counters, then the
// tile scan's group totals (kScanGroups u32)
constexpr int kPartWords = kPartSlots * C_COUNT + kQueues * kQueueStride / 2 + kScanGroups / 2;
__host__ __device__ inline uint32_t* scan_group_sums(unsigned long long* part) {
  return (uint32_t*)(part + (size_t)kPartSlots * C_COUNT + kQueues * kQueueStride / 2);
}
constexpr uint32_t kNoChunk = 0xffffffffu;
struct ChunkClaim {
  uint32_t* ctr;  // kQueues claim counters, kQueueStride apart
  uint32_t nch;   // chunks
  uint32_t home;  // wave-uniform: the range claimed from
  uint32_t done;  // wave-uniform: ranges found empty
  __device__ void init(unsigned long long* part, uint32_t chunks) {
    ctr = (uint32_t*)(part + (size_t)kPartSlots * C_COUNT);
    nch = chunks;
    home = blockIdx.x & (kQueues - 1);
    done = 0;
  }
  // the next chunk, or kNoChunk when every range is empty
  __device__ __forceinline__ uint32_t next() {
    for (int t = 0; t < kQueues; t++) {
      const uint32_t h = (home + (uint32_t)t) & (kQueues - 1);
      if (done & (1u << h)) continue;
      const uint32_t lo = claim_lo(nch, h), hi = claim_lo(nch, h + 1);
      uint32_t v = 0;
      if (lane_id() == 0) v = atomicAdd(&ctr[h * kQueueStride], 1u);
      v = __builtin_amdgcn_readfirstlane(v);
      if (v < hi - lo) {
        home = h;
        return lo + v;
      }
      done |= 1u << h;
    }
    return kNoChunk;
  }
};

// Chunks that interleave the active list: lane l of chunk c takes entry
// R ((l / R) nch + c) + l % R, R = 2^shift (64 / R runs of R consecutive
// entries, nch chunks apart), so every chunk is a systematic sample of the whole
// list and the per-sample work (from a few updates to N) averages out within a
// chunk.  The plain iterate kernel uses R = 1 (on 512 frames of 640x480 at
// N 1024: 0.96 ms against 1.05 ms with runs of 4 and 1.15 ms with 64
// consecutive entries, though each lane's 32-byte state load then touches its
// own line);
// the series kernel R = 64 (64 consecutive entries: neighbouring samples stay on
// the orbit's LDS windows together) unless the launch asks for interleaving
// (deep views whose long samples cluster).
struct ChunkMap {
  uint32_t n;      // entries of the active list
  uint32_t nch;    // chunks: ceil(n / 64)
  uint32_t shift;  // log2 R
  __device__ __forceinline__ void init(uint32_t total, uint32_t run_shift = 0) {
    n = total;
    nch = chunk_count(total);
    shift = run_shift;
  }
  // below R (64 / R nch) <= n + 63 < 2^32 for a real chunk (prepare_common caps
  // the sample count); kNoChunk is tested first
  __device__ __forceinline__ uint32_t pos(uint32_t c) const { return chunk_pos_run((uint32_t)lane_id(), c, nch, shift); }
  // the chunk's entries are the lanes below the returned count (pos grows with
  // the lane)
  __device__ __forceinline__ uint32_t items(uint32_t c) const {
    const bool ok = c < nch && pos(c) < n;
    return (uint32_t)__popcll(__ballot(ok));
  }
  __device__ __forceinline__ uint32_t pix(const uint32_t* active, uint32_t c, uint32_t cnt) const {
    return (uint32_t)lane_id() < cnt ? active[pos(c)] : 0u;
  }
};

// A lane's iteration state inside a block of the plain iterate kernel; `live`
// is the wave's scalar mask of lanes still iterating, `act` the lane's own flag.
// A lane that is not iterating a sample holds z = c = 0 (parked), so the escape
// test of a step needs no mask (one compare and one branch on its result).
struct PlainLane {
  double &x, &y, &r2, &i2, &cx, &cy;
  bool& act;
  int& jesc;      // update of the block at which the lane escaped (-1: none)
  double& mgesc;  // |z|^2 there
};

// Updates J .. K - 1 of a block in groups of G (unrolled by recursion).  The
// handler of an escape records the update and |z|^2 and parks the lane (the
// block's end writes them); the no-escape case falls through.
template <int J, int K, int G, bool FMA2>
__device__ __forceinline__ void plain_block(PlainLane& L, const double b2) {
  static_assert(K % G == 0, "a block holds whole groups");
  if constexpr (J < K) {
    update_group<G, FMA2>(L.x, L.y, L.r2, L.i2, L.cx, L.cy, b2, ~0ull, [&](bool e, int j, double mg) {
      if (e) {
        L.jesc = J + j;
        L.mgesc = mg;
        L.act = false;
        park(L.x, L.y, L.r2, L.i2, L.cx, L.cy);
      }
    });
    plain_block<J + G, K, G, FMA2>(L, b2);
  }
}

template <int K, bool FMA2, bool GROUPED>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_plain_iterate(
    PlainArgs a) {
  const Common& c = a.c;
  const uint32_t total = *c.nactive;
  ChunkMap cm;
  cm.init(total);
  ChunkClaim q;
  q.init(c.part, cm.nch);
  const uint32_t c0 = q.next(), c1 = q.next(), c2 = q.next();
  uint32_t ncur = cm.items(c0), nnext = cm.items(c1), nfar = cm.items(c2);
  uint32_t cur_pix = cm.pix(c.active, c0, ncur), nxt_pix = cm.pix(c.active, c1, nnext);
  uint32_t far_pix = cm.pix(c.active, c2, nfar);
  double4 cur_s = make_double4(0.0, 0.0, 0.0, 0.0), nxt_s = cur_s;
  if ((uint32_t)lane_id() < ncur) cur_s = c.state[cur_pix];
  if ((uint32_t)lane_id() < nnext) nxt_s = c.state[nxt_pix];
  uint32_t pos = 0;
  bool act = false;
  uint32_t pix = 0;
  double cx = 0.0, cy = 0.0, x = 0.0, y = 0.0, r2 = 0.0, i2 = 0.0;
  int k = 0;
  unsigned long long executed = 0;
  const int N = c.N;
  for (;;) {
    unsigned long long need = __ballot(!act);
    while (need && ncur > 0) {
      uint32_t take = min((uint32_t)__popcll(need), ncur - pos);
      uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane_id()) - 1ull));
      int src = (int)min(pos + rank, ncur - 1);
      uint32_t g_pix = __shfl(cur_pix, src);
      double g0 = __shfl(cur_s.x, src), g1 = __shfl(cur_s.y, src), g2 = __shfl(cur_s.z, src),
             g3 = __shfl(cur_s.w, src);
      if (!act && rank < take) {
        pix = g_pix;
        x = g0;
        y = g1;
        cx = g2;
        cy = g3;
        k = a.k_start;
        r2 = x * x;
        i2 = y * y;
        act = true;
      }
      pos += take;
      if (pos == ncur) {  // rotate: NEXT becomes CUR, the far chunk is decoded
        ncur = nnext;
        cur_pix = nxt_pix;
        cur_s = nxt_s;
        nnext = nfar;
        nxt_pix = far_pix;
        if ((uint32_t)lane_id() < nnext) nxt_s = c.state[nxt_pix];
        const uint32_t cf = q.next();
        nfar = cm.items(cf);
        far_pix = cm.pix(c.active, cf, nfar);
        pos = 0;
      }
      need = __ballot(!act);
    }
    unsigned long long live = __ballot(act);  // wave-uniform: the lanes iterating a sample
    if (live == 0) break;
    const int k0 = k;
    const bool was = act;
    int jesc = -1;
    double mgesc = 0.0;
    PlainLane L{x, y, r2, i2, cx, cy, act, jesc, mgesc};
    plain_block<0, K, GROUPED ? kIterGroup : 1, FMA2>(L, c.b2);
    if (was) {
      // the reference tests k >= N before each update: an escape at k0 + j
      // counts only below N, and a sample that reaches N in the block ends there
      const int ran = jesc >= 0 ? jesc + 1 : K;
      if (k0 + ran > N || (jesc < 0 && k0 + K == N)) {
        c.out[pix] = EscapeValue{N, 0.0f};
        act = false;
        park(x, y, r2, i2, cx, cy);
      } else if (jesc >= 0) {
        c.out[pix] = EscapeValue{k0 + jesc, 0.0f};
        c.r2[pix] = mgesc;  // smoothing: k_finalize
      }
      executed += (unsigned long long)min(ran, N - k0);
      k = k0 + K;
    }
  }
  const int idx[1] = {C_EXECUTED};
  block_partial(c.part, kSlotIter + blockIdx.x, &executed, idx, 1);
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

// log2 |e|^2 without under/overflow (to about 1e-13 absolute); -inf for e = 0
__device__ __forceinline__ double log2_sqmag(double2 e) {
  double mx = fmax(fabs(e.x), fabs(e.y));
  if (mx == 0.0) return -INFINITY;
  int ex = ilogb(mx);
  double sx = ldexp(e.x, -ex), sy = ldexp(e.y, -ex);
  return log2_near(sx * sx + sy * sy) + 2.0 * ex;
}

// [lo, hi] holding first_le(PM, 1, last, v) (or hi + 1 = the answer's lower
// bound when v lies below the lookup's range), from the lookup cell of v widened
// by one cell on each side against the rounding of the cell index
__device__ __forceinline__ int2 stop_bracket(const SeriesArgs& a, double v) {
  const int last = a.len - 1;
  if (a.pmt_n == 0 || !(v == v)) return make_int2(1, last);
  const double x = (v - a.pmt_base) * a.pmt_inv_h;
  const int k = x < 0.0 ? 0 : x >= (double)(a.pmt_n - 1) ? a.pmt_n - 1 : (int)x;
  const int lo = a.PMT[min(k + 2, a.pmt_n)];
  const int hi = a.PMT[max(k - 1, 0)];
  return make_int2(max(lo, 1), min(hi, last));
}

// last i in [1, last] with the non-decreasing CX[i] < v, or 0
__device__ __forceinline__ int last_lt(const double* CX, int last, double v) {
  int l = 1, h = last + 1;  // first index with CX >= v
  while (l < h) {
    const int mid = (l + h) >> 1;
    if (CX[mid] >= v) h = mid;
    else l = mid + 1;
  }
  return l - 1;
}

// When a threshold certifies the exact test stable.  T_j > L + kMarg means
// bmag tol >= cmag (1 + 2.7e-9) in exact arithmetic.  The computed test
// (bmag tol < cmag, mandelbrot.cpp:138-142) is then stable when bmag tol >=
// 2^-1030: its own rounding is then under 2^-44 relative; the computed cmag is
// off by under 2^-37 relative (for |e|^2 > 2^-690, e^3 lies at most 21 bits in
// the subnormals) plus 2^-1072 absolute from underflow, which is below 2.7e-9 of
// bmag tol; no product overflows while bmag and cmag stay below 2^1000 (and
// |e|^2 below 2^600).  No lower bound on cmag is needed: a tiny cmag only makes
// the test more stable.  bmin: least log2 |B_j|^2, bmax/cmax: greatest
// log2 |B_j|^2, |C_j|^2 over the indices concerned.
constexpr double kLogLo = -1030.0, kLogHi = 1000.0, kLMin = -690.0, kLMax = 600.0;
__device__ __forceinline__ bool rounding_certified(const SeriesArgs& a, double L, double bmin, double bmax,
                                                   double cmax) {
  return L > kLMin && L < kLMax && bmin + 2.0 * L + a.ltol > kLogLo && bmax + 2.0 * L < kLogHi &&
         cmax + 3.0 * L < kLogHi;
}

// first i >= lo with T_i <= v (last + 1 if none), by the sparse table
__device__ __forceinline__ int first_le_from(const SeriesArgs& a, int lo, double v) {
  const int len = a.len;
  int pos = lo;
  for (int k = a.st_levels - 1; k >= 0; k--)
    if (pos + (1 << k) <= len && a.ST[(size_t)k * len + pos] > v) pos += 1 << k;
  return pos;
}

// Where the reference's exact scan must start when the prefix certificate of
// sa_stop_index fails (its terms under- or overflow somewhere before the stop).
// Indices whose |C_j e^3|^2 lies below 2^-1080 are stable whatever their
// threshold says: cmag rounds to 0 and bmag tol < 0 is false; that holds for a
// whole prefix [1, ic] (CX is a prefix maximum, no special index inside).  From
// ic + 1 on, the threshold table applies where every exact-test intermediate of
// the range [ic + 1, ia) is certified by rounding_certified (suffix minima from
// ic + 1, prefix maxima up to ia - 1).
__device__ int deep_start(const SeriesArgs& a, double L) {
  const int last = a.len - 1;
  int ic = last_lt(a.CX, last, -1080.0 - 3.0 * L);
  if (ic > 0 && a.SP[ic] != 0) ic = 0;
  const int lo = ic + 1;
  if (lo > last || !(L > kLMin && L < kLMax)) return lo;
  const int ia = first_le_from(a, lo, L + kMarg);
  const int sidx = min(ia, last + 1) - 1;  // the last index the table certifies
  if (sidx < lo) return lo;
  const bool ok = a.SP[sidx] - a.SP[lo - 1] == 0 &&
                  rounding_certified(a, L, a.BNs[lo], a.BX[sidx], a.CX[sidx]);
  return ok ? ia : lo;
}

__device__ int sa_stop_index(const SeriesArgs& a, double L, int2 br, double2 e2, double2 e3) {
  const int last = a.len - 1;
  if (last < 1) return 0;
  if (L == -INFINITY) return 0;  // e = 0: b = c = 0 and the test never fires
  int ia = first_le(a.PM, br.x, br.y, L + kMarg);
  if (ia > br.y && br.y < last) ia = first_le(a.PM, br.y + 1, last, L + kMarg);
  int safe_idx = ia > last ? last : ia - 1;
  bool safe = L > kLMin && L < kLMax;
  if (safe && safe_idx >= 1)
    safe = a.SP[safe_idx] == 0 &&
           rounding_certified(a, L, a.BN[safe_idx], a.BX[safe_idx], a.CX[safe_idx]);
  int from = safe ? ia : deep_start(a, L);
  for (int i = from; i <= last; i++)
    if (unstable_at(a, i, e2, e3)) return i;
  return 0;
}

// ------------------------------------------------ floatexp series (extension)
// The reference's series test and skip (mandelbrot.cpp:138-142, 160-178) in the
// same operation order, each real held as m * 2^e with m = 0 (e = kFxZeroE) or
// 1/2 <= |m| < 1.  A product or sum rounds the mantissa once, to nearest, as the
// double operation on the scaled operands does, so wherever the reference's
// doubles are normal numbers the result is the same number; below widths of
// about 1e-90, where its |B e^2|^2 and |C e^3|^2 leave the double range and the
// test degenerates, nothing here under- or overflows.
constexpr int kFxZeroE = -0x40000000;
struct fxr {
  double m;
  int e;
};
struct fxc {
  fxr re, im;
};
__device__ __forceinline__ fxr fx_make(double d, int e) {
  int k;
  const double m = frexp(d, &k);
  return d == 0.0 ? fxr{0.0, kFxZeroE} : fxr{m, e + k};
}
__device__ __forceinline__ fxr fx_mul(fxr a, fxr b) { return fx_make(a.m * b.m, a.e + b.e); }
// the operand with the larger exponent keeps its scale (a zero has the least
// exponent, and the other operand comes through unchanged)
__device__ __forceinline__ fxr fx_add(fxr a, fxr b) {
  const bool ab = a.e >= b.e;
  const fxr h = ab ? a : b, l = ab ? b : a;
  return fx_make(h.m + ldexp(l.m, max(l.e - h.e, -2100)), h.e);
}
__device__ __forceinline__ fxr fx_neg(fxr a) { return fxr{-a.m, a.e}; }
// a < b for a, b >= 0
__device__ __forceinline__ bool fx_lt(fxr a, fxr b) { return a.e < b.e || (a.e == b.e && a.m < b.m); }
__device__ __forceinline__ double fx_get(fxr a) { return ldexp(a.m, max(a.e, -2100)); }
__device__ __forceinline__ fxc fxc_of(double2 z) { return fxc{fx_make(z.x, 0), fx_make(z.y, 0)}; }
__device__ __forceinline__ fxc fxc_mul(fxc a, fxc b) {  // complex.h:29-31
  return fxc{fx_add(fx_mul(a.re, b.re), fx_neg(fx_mul(a.im, b.im))), fx_add(fx_mul(a.re, b.im), fx_mul(a.im, b.re))};
}
__device__ __forceinline__ fxc fxc_sq(fxc a) {  // complex.h:19-21: (re^2 - im^2, (2 re) im)
  return fxc{fx_add(fx_mul(a.re, a.re), fx_neg(fx_mul(a.im, a.im))), fx_mul(fxr{a.re.m, a.re.e + 1}, a.im)};
}
__device__ __forceinline__ fxr fxc_sqmag(fxc a) { return fx_add(fx_mul(a.re, a.re), fx_mul(a.im, a.im)); }
__device__ __forceinline__ fxc fx_coef(const double2* m, const int2* e, int i) {
  const double2 mm = m[i];
  const int2 ee = e[i];
  return fxc{fxr{mm.x, ee.x}, fxr{mm.y, ee.y}};
}

// e, e^2 and e^3 of a sample in floatexp (mandelbrot.cpp:160-161)
struct EpsFx {
  fxc e, e2, e3;
};
__device__ __forceinline__ EpsFx eps_fx(double2 e) {
  EpsFx p;
  p.e = fxc_of(e);
  p.e2 = fxc_sq(p.e);
  p.e3 = fxc_mul(p.e, p.e2);
  return p;
}

__device__ __forceinline__ bool unstable_fx(const SeriesArgs& a, int i, const EpsFx& p) {
  const fxc b = fxc_mul(fx_coef(a.Bm, a.Be, i), p.e2);
  const fxc c = fxc_mul(fx_coef(a.Cm, a.Ce, i), p.e3);
  return fx_lt(fx_mul(fxc_sqmag(b), fx_make(a.tol, 0)), fxc_sqmag(c));
}

__device__ __forceinline__ double2 sa_term_fx(const SeriesArgs& a, int i, double2 e, const EpsFx& p) {
  if (i == 0) return e;
  const fxc ta = fxc_mul(fx_coef(a.Am, a.Ae, i), p.e);
  const fxc tb = fxc_mul(fx_coef(a.Bm, a.Be, i), p.e2);
  const fxc tc = fxc_mul(fx_coef(a.Cm, a.Ce, i), p.e3);
  return make_double2(fx_get(fx_add(fx_add(ta.re, tb.re), tc.re)), fx_get(fx_add(fx_add(ta.im, tb.im), tc.im)));
}

// The first unstable index with the floatexp test.  Its rounding is a few units
// in the last place of |B e^2|^2 tol and |C e^3|^2 with no range limit, so an
// index whose threshold T_j = log2(tol |B_j|^2 / |C_j|^2) lies more than kMarg
// above log2 |e|^2 is stable (T_j = -inf where B_j = 0 and C_j != 0: unstable;
// +inf where C_j = 0: stable); only the indices at or below that line are tested,
// in order, found on the prefix minimum and then on the sparse table.
__device__ int sa_stop_index_fx(const SeriesArgs& a, double L, int2 br, const EpsFx& p) {
  const int last = a.len - 1;
  if (last < 1 || L == -INFINITY) return 0;  // e = 0: the test never fires
  const double v = L + kMarg;
  int i = first_le(a.PM, br.x, br.y, v);
  if (i > br.y && br.y < last) i = first_le(a.PM, br.y + 1, last, v);
  while (i <= last) {
    if (unstable_fx(a, i, p)) return i;
    i = first_le_from(a, i + 1, v);
  }
  return 0;
}

// Per-tile series skip: the first index at which the reference's stability test
// fails at each tile's probe (its sample farthest from the reference point).
// One wave per tile.  mode 0: the reference's scan over i = 1, 2, ... (64 indices
// per round, ballot for the first unstable one); mode 1: the threshold-table
// search of the per-sample path; mode 2: validate pred[t] and fall back to the
// scan.  The check of a start s: every i < s is certainly stable by the prefix
// minimum of the thresholds (PM[s-1] > log2|e|^2 + margin, with the magnitude
// checks that make the exact test safe; for the prediction p = 0, "never
// unstable", the whole prefix).  The scan runs from the first of p, p - p/64,
// p - p/16, p - p/4, p/2, p/4 and p/16 that passes the check (p itself is the
// answer when the exact test fails there), else from 1.  The result is the
// scan's in every case.
// how[t]: 0 scanned, 1 table, 2 prediction exact, 3 prediction rejected (scanned
// from 1), 4 prediction a certified lower bound (scanned from it); evals[t]:
// exact tests evaluated.
__global__ __launch_bounds__(64) void k_tile_skip(SeriesArgs a, const int2* probe, const int* pred, int mode,
                                                  int ntiles, int* ua_out, int* how, int* evals) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int lane = lane_id();
  const int2 pp = probe[t];
  const double2 e = make_double2(a.epsre[pp.x], a.epsim[pp.y]);
  const double2 e2 = make_double2(e.x * e.x - e.y * e.y, 2.0 * e.x * e.y);
  const double2 e3 = cmul(e, e2);
  const double L = log2_sqmag(e);
  const EpsFx pf = a.fxs ? eps_fx(e) : EpsFx{};
  const int last = a.len - 1;
  int ua = -1, h = 0, ev = 0, from = 1;
  if (last < 1 || L == -INFINITY) {  // e = 0: the test never fires
    ua = 0;
    h = 1;
  } else if (mode == 1) {
    ua = a.fxs ? sa_stop_index_fx(a, L, stop_bracket(a, L + kMarg), pf)
               : sa_stop_index(a, L, stop_bracket(a, L + kMarg), e2, e3);
    h = 1;
  } else if (mode == 2) {
    const int p = pred[t];
    // the floatexp test has no range limits: the prefix minimum alone certifies
    auto stable_below = [&](int s) {
      const int below = s - 1;
      return below >= 1 && a.PM[below] > L + kMarg &&
             (a.fxs || (a.SP[below] == 0 && rounding_certified(a, L, a.BN[below], a.BX[below], a.CX[below])));
    };
    h = 3;
    if (p == 0 && stable_below(last + 1)) {
      ua = 0;
      h = 2;
    } else if (p == 0 || (p > 1 && p <= last)) {
      // "never unstable" that fails its check backs off from the last index
      const int q = p == 0 ? last : p;
      const int cand[7] = {q, q - max(1, q >> 6), q - (q >> 4), q - (q >> 2), q >> 1, q >> 2, q >> 4};
      for (int k = 0; k < 7; k++)
        if (cand[k] > 1 && stable_below(cand[k])) {
          from = cand[k];
          h = 4;
          break;
        }
    }
  }
  if (ua < 0) {  // the reference's scan, from `from`
    ua = 0;
    for (int base = from; base <= last; base += 64) {
      const int i = base + lane;
      const bool u = i <= last && (a.fxs ? unstable_fx(a, i, pf) : unstable_at(a, i, e2, e3));
      const unsigned long long m = __ballot(u);
      if (m) {
        const int k = __ffsll((unsigned long long)m);  // 1 + the first unstable lane
        ua = base + k - 1;
        ev += k;
        if (h == 4 && ua == from) h = 2;  // the prediction itself
        break;
      }
      ev += min(64, last - base + 1);
    }
  }
  if (lane == 0) {
    ua_out[t] = ua;
    how[t] = h;
    evals[t] = ev;
  }
}

// The series start kernel: per sample the reference's series start
// (mandelbrot.cpp:144-206) and the first perturbation step; samples that need
// more are flagged for the iterate kernel.  Held to 80 VGPRs (6 waves per SIMD):
// 5% less time than at its natural 89 (5 waves) where the stop search dominates.
// FXS: the series test and skip in floatexp.
template <bool FXS>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_series_start(SeriesArgs a) {
  const Common& c = a.c;
  unsigned long long ncard = 0, nres = 0, skipped = 0, credited = 0;
  const uint32_t ntiles = start_tiles(c.total);
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t mine = 0;
    for (int j = 0; j < kPerThread; j++) {
      uint32_t gid = start_gid(tile, j, threadIdx.x);
      bool valid = gid < c.total;
      // phase 1: locate the sample, cardioid class, epsilon and log2 |e|^2
      bool in = false;
      double L = -INFINITY;
      double2 e = make_double2(0.0, 0.0);
      if (valid) {
        const DevFrame& f = c.frames[frame_of(gid, c.inv_frame)];
        const Loc lc = locate(f, gid - f.sample_base);
        valid = lc.valid;
        const uint32_t r = lc.r, col = lc.c;
        in = f.card_mode == CARD_ALL;
        if (valid && f.card_mode == CARD_MASK) {
          uint64_t bit = f.mask_base + (uint64_t)r * f.nc + col;
          in = (a.mask[bit >> 5] >> (bit & 31)) & 1u;
        }
        if (valid && !in) {
          e = make_double2(a.epsre[f.tab_c + col], a.epsim[f.tab_r + r]);
          L = log2_sqmag(e);
        }
      }
      // phase 2 (all lanes): the wave's bracket of the stop-index search
      const int2 br = a.tile_ua ? make_int2(1, a.len - 1) : stop_bracket(a, L + kMarg);
      uint32_t skip_tile = 0;
      if (a.tile_ua && valid) {
        const uint32_t fi = frame_of(gid, c.inv_frame);
        const DevFrame& f = c.frames[fi];
        const Loc lc = locate(f, gid - f.sample_base);
        skip_tile = fi * a.skip_tpf + (lc.r / a.skip_T) * a.skip_ntc + lc.c / a.skip_T;
      }
      // phase 3 (valid lanes): the reference's series start for this sample
      bool want = false;
      if (valid) {
        EscapeValue res{c.N, 0.0f};
        if (in) {
          ncard++;
        } else {
          // eps2 = sq(eps), eps3 = eps * eps2 (complex.h:19-21, mandelbrot.cpp:160-161)
          double2 e2, e3;
          EpsFx pf;
          if constexpr (FXS) {
            pf = eps_fx(e);
          } else {
            e2 = make_double2(e.x * e.x - e.y * e.y, 2.0 * e.x * e.y);
            e3 = cmul(e, e2);
          }
          auto term = [&](int i) {
            if constexpr (FXS) return sa_term_fx(a, i, e, pf);
            else return sa_term(a, i, e, e2, e3);
          };
          int ua;
          if (a.tile_ua) ua = a.tile_ua[skip_tile];
          else if constexpr (FXS) ua = sa_stop_index_fx(a, L, br, pf);
          else ua = sa_stop_index(a, L, br, e2, e3);
          int dsize = ua ? max(ua - 3, 1) : a.len;
          int found = dsize - 1;
          double2 d = term(found);
          double2 zf = a.Zd[found + 1], xl = a.Xlo[found];
          double yr = trunc_sum3(zf.x, xl.x, d.x), yi = trunc_sum3(zf.y, xl.y, d.y);
          if (yr * yr + yi * yi > c.b2) {
            // escape inside the skip: the reference's binary search (mandelbrot.cpp:186-206)
            int low = 0, high = dsize - 1, mid = dsize / 2;
            while (low <= high) {
              double2 dm = term(mid);
              double2 zm = a.Zd[mid + 1], xm = a.Xlo[mid];
              double ym = trunc_sum3(zm.x, xm.x, dm.x), yn = trunc_sum3(zm.y, xm.y, dm.y);
              if (!(ym * ym + yn * yn > c.b2)) low = mid + 1;
              else {
                high = mid - 1;
                found = mid;
              }
              mid = (low + high) / 2;
            }
            d = term(found);
            zf = a.Zd[found + 1];
            xl = a.Xlo[found];
            yr = trunc_sum3(zf.x, xl.x, d.x);
            yi = trunc_sum3(zf.y, xl.y, d.y);
            res = EscapeValue{found, smooth_of(yr * yr + yi * yi, c)};
            nres++;
          } else if (found + 1 < c.N) {
            // first perturbation step from Y[found] = X[found] + d (its escape
            // test was the one above); rebase if |Y| < |d|
            double dr = d.x, di = d.y;
            int m = found + 1;
            double zr = zf.x, zi = zf.y;
            double pr = zr + dr, pi = zi + di;
            double r2 = pr * pr + pi * pi;
            double dmag = fma(dr, dr, di * di);
            double wr = zr + pr, wi = zi + pi;
            if (r2 < dmag || m == a.mend) {
              dr = pr;
              di = pi;
              wr = pr;
              wi = pi;
              m = 0;
            }
            double ndr = fma(dr, wr, fma(-di, wi, e.x));
            double ndi = fma(dr, wi, fma(di, wr, e.y));
            // the delta only: the iterate kernel reads epsilon from the tables
            ((double2*)c.state)[gid] = make_double2(ndr, ndi);
            c.state_i[gid] = make_int2(found + 1, m + 1);
            want = true;
            skipped += (unsigned long long)(found + 1);
          } else {
            nres++;  // the skip reached N: (N, 0) as the empty HP loop returns
          }
        }
        if (!want) c.out[gid] = res;
        if (!want && !in) credited += (unsigned long long)res.iterations;
      }
      mine += store_hand(c.hand, gid, want);
    }
    tile_count_add(c.tile_count, tile, mine);
  }
  const unsigned long long v[4] = {ncard, nres, skipped, credited};
  const int idx[4] = {C_CARDIOID, C_SA_RES, C_SKIPPED, C_CREDITED};
  block_partial(c.part, kSlotStart + blockIdx.x, v, idx, 4);
}

// epsilon of sample gid (its pixel's offset from the reference point), from the
// frame's delta tables, as the series start read it
__device__ __forceinline__ double2 sample_eps(const SeriesArgs& a, uint32_t gid) {
  const Common& c = a.c;
  const DevFrame& f = c.frames[frame_of(gid, c.inv_frame)];
  const Loc lc = locate(f, gid - f.sample_base);
  return make_double2(a.epsre[f.tab_c + lc.c], a.epsim[f.tab_r + lc.r]);
}

// an orbit entry as one 16-byte vector, and pointers to the orbit: in LDS (the
// whole wave), or per lane in LDS or HBM (generic addresses, flat loads)
typedef double ld_double2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const ld_double2* lds_orbit;
typedef const ld_double2* gen_orbit;

// p - j orbit entries, as an address (the result may point before the array; it
// is only read at offsets j' > j)
template <class P>
__device__ __forceinline__ P orbit_back(P head, int j) {
  return (P)((uintptr_t)head - (uintptr_t)(16 * j));
}
template <class P>
__device__ __forceinline__ int orbit_index(P p, P head) {
  return (int)(((intptr_t)p - (intptr_t)head) >> 4);
}

// One block of K perturbation steps for a lane (Y_it = Zd[m] + d):
//   d' = d (Zd[m] + Y) + e, rebase to the orbit start (m = 0, d = Y) when
//   |Y| < |d| (or, in END blocks, when m reaches the end of the reference).
// Step j reads its orbit entry at zp[j], zp = the lane's Zd[m0]: in LDS when
// every lane's block lies in an LDS window, else per lane in a window or in HBM
// (a flat load: only the lanes reading HBM cost the L1 a line lookup, which
// limits the kernel when a wave's lanes are spread over the orbit); a rebase in
// step j moves zp to head - j (the LDS head window), so step j' > j reads
// Zd[j' - j].  Each entry is loaded one step ahead, as soon as that step's
// rebase is decided.
// The first escape (exact sqMag > bailout^2, mandelbrot.cpp:60) records its
// step and |Y|^2: the step whose test holds while no earlier step's of the block
// did (a flag in a scalar mask: escape need not be absorbing, |Y| can fall back
// under a bailout below 2 + |c|, and overflow turns into NaN, which fails the
// test).
// A block may end after K/2 or 3K/4 steps when at least kExitDone lanes of the
// wave have nothing left to do in it (their sample escaped, or the lane was idle)
// and the wave can refill them; it returns the steps it ran.
constexpr int kExitDone = 12;

template <int K, bool END, class P>
__device__ __forceinline__ int pert_block(double& dr, double& di, const double er, const double ei, int& m,
                                          const int it, P zp, const P head, const int endj, const int mend,
                                          const double b2, const bool idle, const bool can_refill, int& esc,
                                          double& esc_r2, unsigned int& rb) {
  int kdone = K;
  bool moved = false;
  int escj = -1;  // step of the block's first escape
  unsigned long long seen = 0;  // wave-uniform: lanes with an escape earlier in the block
  ld_double2 z = zp[0];
#pragma unroll
  for (int j = 0; j < K; j++) {
    const double zr = z.x;
    const double zi = z.y;
    const double yr = zr + dr, yi = zi + di;
    const double r2 = yr * yr + yi * yi;  // sqMag (mandelbrot.cpp:60)
    const double dmag = fma(dr, dr, di * di);
    bool reb = r2 < dmag;
    // the forced rebase at the orbit's end, m == mend
    if (END) reb = reb || (moved ? orbit_index(zp, head) + j == mend : j == endj);
    double wr = zr + yr, wi = zi + yi;
    double d0r = dr, d0i = di;
    // a rebase (d = Y against Zd[0] = 0, so W = Y) moves four doubles and the
    // cursor under the lane mask; about half the wave-steps have no rebasing
    // lane and skip those moves
    if (__builtin_expect(__ballot(reb) != 0, 0)) {
      if (reb) {
        d0r = yr;
        d0i = yi;
        wr = yr;
        wi = yi;
        zp = orbit_back(head, j);
        moved = true;
      }
    }
    if (j + 1 < K) z = zp[j + 1];
    // the first escape's record, only on wave-steps where some lane escapes
    // for the first time in the block
    const unsigned long long first = __ballot(r2 > b2) & ~seen;
    if (__builtin_expect(first != 0, 0)) {
      seen |= first;
      if ((first >> lane_id()) & 1ull) {
        escj = j;
        esc_r2 = r2;
      }
    }
    dr = fma(d0r, wr, fma(-d0i, wi, er));
    di = fma(d0r, wi, fma(d0i, wr, ei));
    // the escape record is complete at the end of each step (no chain of
    // per-step values kept live to the end of the block)
    __asm__ volatile("" ::"v"(escj), "v"(esc_r2));
    if (K >= 8 && (j + 1 == K / 2 || j + 1 == 3 * K / 4) && can_refill &&
        __popcll(__ballot(idle || escj >= 0)) >= kExitDone) {
      kdone = j + 1;
      break;
    }
  }
  esc = escj >= 0 ? it + escj : esc;
  m = moved ? orbit_index(zp, head) + kdone : m + kdone;
  rb += moved ? 1u : 0u;
  return kdone;
}

// The FP32X form of pert_block: the delta is w * 2^ex with an FP32 mantissa
// pair w and an int32 exponent ex shared by re and im; es = e * 2^-ex.  Per step,
// in FP32: d = w * 2^ex (0 below the FP32 range), Y = Z + d, the rebase test
// |Y|^2 < |d|^2, W = Z + Y and w' = w W + es.  In FP64: the escape test when the
// FP32 |Y|^2 comes within 2^-10 of the bailout (Y = Zd + w * 2^ex, the reference's
// sqMag), and every rebase (d = Y from the FP64 Zd and delta, renormalised).
__device__ __forceinline__ void fx_normalise(double yr, double yi, double er, double ei, float& wr, float& wi,
                                             int& ex, float& esr, float& esi) {
  const double mx = fmax(fabs(yr), fabs(yi));
  if (mx > 0.0) ex = ilogb(mx);
  wr = (float)ldexp(yr, -ex);
  wi = (float)ldexp(yi, -ex);
  esr = (float)ldexp(er, -ex);
  esi = (float)ldexp(ei, -ex);
}

template <int K, bool END, class P>
__device__ __forceinline__ void pert_block_fx(float& wr, float& wi, int& ex, float& esr, float& esi, const double er,
                                              const double ei, int& m, const int it, P zp, const P head,
                                              const int endj, const int mend, const double b2, const float b2g,
                                              int& esc, double& esc_r2, unsigned int& rb) {
  const float qnanf = __int_as_float(0x7fc00000);
  bool moved = false;
  ld_double2 z = zp[0];
#pragma unroll
  for (int j = 0; j < K; j++) {
    const double zrd = z.x;
    const double zid = z.y;
    const float zr = (float)zrd, zi = (float)zid;
    const float dr = ldexpf(wr, ex), di = ldexpf(wi, ex);
    const float yr = zr + dr, yi = zi + di;
    const float r2 = fmaf(yr, yr, yi * yi);
    // near the bailout: the reference's test in FP64 (only on wave-steps where
    // some lane comes near it)
    if (__builtin_expect(__ballot(r2 > b2g) != 0, 0)) {
      if (r2 > b2g) {
        const double ydr = zrd + ldexp((double)wr, ex), ydi = zid + ldexp((double)wi, ex);
        const double r2d = ydr * ydr + ydi * ydi;
        if (r2d > b2) {  // first escape: NaN in wr reaches every later step
          esc = it + j;
          esc_r2 = r2d;
          wr = qnanf;
        }
      }
    }
    // the rebase test |Y|^2 < |d|^2 in the mantissa's scale: |Z 2^-ex + w|^2 <
    // |w|^2 (d = w 2^ex underflows FP32 long before the deltas this form is for;
    // Z 2^-ex is formed in FP64, then rounded; an overflow to inf means |Y| >> |d|)
    const float ysr = (float)ldexp(zrd, -ex) + wr, ysi = (float)ldexp(zid, -ex) + wi;
    bool reb = fmaf(ysr, ysr, ysi * ysi) < fmaf(wr, wr, wi * wi);
    if (END) reb = reb || (moved ? orbit_index(zp, head) + j == mend : j == endj);
    float Wr = zr + yr, Wi = zi + yi;
    if (__builtin_expect(__ballot(reb) != 0, 0)) {
      if (reb) {  // rebase in FP64: d = Y against Zd[0] = 0, so W = Y
        const double ydr = zrd + ldexp((double)wr, ex), ydi = zid + ldexp((double)wi, ex);
        fx_normalise(ydr, ydi, er, ei, wr, wi, ex, esr, esi);
        Wr = (float)ydr;
        Wi = (float)ydi;
        zp = orbit_back(head, j);
        moved = true;
      }
    }
    if (j + 1 < K) z = zp[j + 1];
    const float nwr = fmaf(wr, Wr, fmaf(-wi, Wi, esr));
    const float nwi = fmaf(wr, Wi, fmaf(wi, Wr, esi));
    wr = nwr;
    wi = nwi;
  }
  m = moved ? orbit_index(zp, head) + K : m + K;
  rb += moved ? 1u : 0u;
}

// FP64 (FX = false) or FP32X (FX = true) perturbation with rebasing (replaces the
// HP continuation of mandelbrot.cpp:209-224).  Persistent waves, lanes refilled
// from 64-sample chunks; blocks of K steps, each lane reading the orbit through
// its own cursor into the workgroup's LDS windows or HBM.  The FP64 kernel is
// held to 64 VGPRs (8 waves per SIMD, four 512-thread workgroups per CU): only
// the next chunk's sample ids are prefetched, its states are loaded when it
// becomes current (a wave waits one load per 64 refills, while the other waves
// of its SIMD run).  The chunk map and claims are inlined (as calls they kept
// their objects in scratch and spilled the LDS head pointer, reloaded in the
// step loop's rebase branch): no scratch at all.
template <int K, bool FX>
__global__ __launch_bounds__(kSeriesIterThreads) __attribute__((amdgpu_waves_per_eu(FX || K < 16 ? 1 : 8, 8)))
void k_series_iterate(SeriesArgs a) {
  const Common& c = a.c;
  const uint32_t total = *c.nactive;
  uint32_t executed = 0, rebased = 0;  // per lane (< 2^32 steps per lane per launch)
  const int N = c.N, mend = a.mend;
  // the orbit's tail and head windows (zlds[t] = Zd[tail_lo + t],
  // zlds[kLdsTail + h] = Zd[h] for h < kLdsHead); entries past Zd's allocation
  // read 0.  The head sits above 16 K positions so that a rebased lane's base
  // address (head - 16 j) stays non-negative.
  __shared__ double2 zlds[kLdsTail + kLdsHead];
  const int tail_lo = max(kLdsHead, min(a.zd_count, a.len + 2 + K) - kLdsTail);
  const int tail_hi = tail_lo + kLdsTail;
  for (int i = threadIdx.x; i < kLdsHead + kLdsTail; i += blockDim.x) {
    const int src = i < kLdsTail ? tail_lo + i : i - kLdsTail;
    zlds[i] = src < a.zd_count ? a.Zd[src] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double2* zhead = zlds + kLdsTail;  // rebased lanes read the orbit from its start
  ChunkMap cm;
  cm.init(total, a.run_shift);
  ChunkClaim q;
  q.init(c.part, cm.nch);
  const uint32_t c0 = q.next(), c1 = q.next();
  uint32_t ncur = cm.items(c0), nnext = cm.items(c1);
  uint32_t cur_pix = cm.pix(c.active, c0, ncur), nxt_pix = cm.pix(c.active, c1, nnext);
  const double2* dstate = (const double2*)c.state;  // the series start's deltas
  double2 cur_s = make_double2(0.0, 0.0), cur_e = cur_s;
  int2 cur_i = make_int2(0, 0);
  if ((uint32_t)lane_id() < ncur) {
    cur_s = dstate[cur_pix];
    cur_i = c.state_i[cur_pix];
    cur_e = sample_eps(a, cur_pix);
  }
  uint32_t pos = 0;
  bool act = false;
  uint32_t pix = 0;
  double dr = 0.0, di = 0.0, er = 0.0, ei = 0.0;
  float wr = 0.0f, wi = 0.0f, esr = 0.0f, esi = 0.0f;  // FP32X delta and scaled epsilon
  int ex = 0;
  int it = 0, m = 0;
  for (;;) {
    unsigned long long need = __ballot(!act);
    while (need && ncur > 0) {
      uint32_t take = min((uint32_t)__popcll(need), ncur - pos);
      uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane_id()) - 1ull));
      int src = (int)min(pos + rank, ncur - 1);
      uint32_t g_pix = __shfl(cur_pix, src);
      double g0 = __shfl(cur_s.x, src), g1 = __shfl(cur_s.y, src), g2 = __shfl(cur_e.x, src),
             g3 = __shfl(cur_e.y, src);
      int g4 = __shfl(cur_i.x, src), g5 = __shfl(cur_i.y, src);
      if (!act && rank < take) {
        pix = g_pix;
        dr = g0;
        di = g1;
        er = g2;
        ei = g3;
        it = g4;
        m = g5;
        act = true;
        if (FX) {
          ex = 0;
          const double mx = fmax(fabs(dr), fabs(di));
          const double me = fmax(fabs(er), fabs(ei));
          if (mx == 0.0 && me > 0.0) ex = ilogb(me);
          fx_normalise(dr, di, er, ei, wr, wi, ex, esr, esi);
        }
      }
      pos += take;
      if (pos == ncur) {  // rotate: NEXT becomes CUR, its states loaded now
        ncur = nnext;
        cur_pix = nxt_pix;
        if ((uint32_t)lane_id() < ncur) {
          cur_s = dstate[cur_pix];
          cur_i = c.state_i[cur_pix];
          cur_e = sample_eps(a, cur_pix);
        }
        const uint32_t cn = q.next();
        nnext = cm.items(cn);
        nxt_pix = cm.pix(c.active, cn, nnext);
        pos = 0;
      }
      need = __ballot(!act);
    }
    if (__ballot(act) == 0) break;
    if (!act) {  // idle lanes run the block on a harmless state
      dr = di = er = ei = 0.0;
      wr = wi = esr = esi = 0.0f;
      ex = 0;
      m = 0;
      it = 0;
    }
    // lanes whose block crosses the reference's end rebase there (END variant)
    const bool end_zone = __ballot(act && m + K > mend) != 0 || mend <= K;
    const bool in_head = m + K <= kLdsHead;
    const bool in_tail = !in_head && m >= tail_lo && m + K <= tail_hi;
    int esc = -1;
    double esc_r2 = 0.0;
    unsigned int rb = 0;
    int kdone = K;
    const int it0 = it;
    const bool can_refill = pos < ncur || nnext > 0;  // wave-uniform
    if (FX) {
      const float b2g = (float)(c.b2 * (1.0 - 1.0 / 1024.0));
      if (__ballot(act && !in_head && !in_tail) == 0) {
        const lds_orbit lhead = (lds_orbit)(const void*)zhead;
        const lds_orbit zp = in_head ? lhead + m : (lds_orbit)(const void*)zlds + (m - tail_lo);
        if (end_zone)
          pert_block_fx<K, true>(wr, wi, ex, esr, esi, er, ei, m, it, zp, lhead, mend - m, mend, c.b2, b2g, esc,
                                 esc_r2, rb);
        else
          pert_block_fx<K, false>(wr, wi, ex, esr, esi, er, ei, m, it, zp, lhead, 0, mend, c.b2, b2g, esc, esc_r2,
                                  rb);
      } else {
        const gen_orbit fhead = (gen_orbit)(const void*)zhead;
        const gen_orbit fz = in_head ? fhead + m
                                     : in_tail ? (gen_orbit)(const void*)zlds + (m - tail_lo)
                                               : (gen_orbit)(const void*)a.Zd + m;
        if (end_zone)
          pert_block_fx<K, true>(wr, wi, ex, esr, esi, er, ei, m, it, fz, fhead, mend - m, mend, c.b2, b2g, esc,
                                 esc_r2, rb);
        else
          pert_block_fx<K, false>(wr, wi, ex, esr, esi, er, ei, m, it, fz, fhead, 0, mend, c.b2, b2g, esc, esc_r2,
                                  rb);
      }
      // keep the mantissa near 1: |w| grows by at most |W| per step
      const float mx = fmaxf(fabsf(wr), fabsf(wi));
      if (mx > 0.0f && mx < INFINITY) {
        const int k = ilogbf(mx);
        if (k > 8 || k < -8) {
          wr = ldexpf(wr, -k);
          wi = ldexpf(wi, -k);
          ex += k;
          esr = (float)ldexp(er, -ex);
          esi = (float)ldexp(ei, -ex);
        }
      }
    } else if (__ballot(act && !in_head && !in_tail) == 0) {
      // every lane's block lies in an LDS window: the whole block reads LDS
      const lds_orbit lhead = (lds_orbit)(const void*)zhead;
      const lds_orbit zp = in_head ? lhead + m : (lds_orbit)(const void*)zlds + (m - tail_lo);
      if (end_zone)
        kdone = pert_block<K, true>(dr, di, er, ei, m, it, zp, lhead, mend - m, mend, c.b2, !act, can_refill, esc,
                                    esc_r2, rb);
      else
        kdone = pert_block<K, false>(dr, di, er, ei, m, it, zp, lhead, 0, mend, c.b2, !act, can_refill, esc, esc_r2,
                                     rb);
    } else {
      // per lane: an LDS window where the lane's block lies in one, else HBM
      // (generic addresses: one flat load per step serves both spaces, and only
      // the lanes reading HBM cost the L1 a line lookup)
      const gen_orbit fhead = (gen_orbit)(const void*)zhead;
      const gen_orbit fz = in_head ? fhead + m
                                   : in_tail ? (gen_orbit)(const void*)zlds + (m - tail_lo)
                                             : (gen_orbit)(const void*)a.Zd + m;
      if (end_zone)
        kdone = pert_block<K, true>(dr, di, er, ei, m, it, fz, fhead, mend - m, mend, c.b2, !act, can_refill, esc,
                                    esc_r2, rb);
      else
        kdone = pert_block<K, false>(dr, di, er, ei, m, it, fz, fhead, 0, mend, c.b2, !act, can_refill, esc,
                                     esc_r2, rb);
    }
    if (act) {
      it = it0 + kdone;
      const bool escaped = esc >= 0 && esc <= N - 1;
      executed += escaped ? (unsigned)(esc - it0 + 1) : (unsigned)min(kdone, N - it0);
      rebased += rb;
      if (escaped || it >= N) {  // tops up to N-1 decided
        c.out[pix] = EscapeValue{escaped ? esc : N, 0.0f};
        if (escaped) c.r2[pix] = esc_r2;  // smoothing: k_finalize
        act = false;
      }
    }
  }
  const unsigned long long v[2] = {executed, rebased};
  const int idx[2] = {C_EXECUTED, C_REBASED};
  block_partial(c.part, kSlotIter + blockIdx.x, v, idx, 2);
}

// smoothing of the iterated samples (mandelbrot.cpp:133-136), converged: one
// thread per entry of active[]
// tiled sample order: results (sample order) -> the row-major grids
__global__ __launch_bounds__(kThreads) void k_emit(Common c, EscapeValue* grid) {
  // 64-bit loop index: a total near 2^32 must not wrap the stride
  for (uint64_t g = blockIdx.x * blockDim.x + threadIdx.x; g < c.total; g += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t gid = (uint32_t)g;
    const DevFrame& f = c.frames[frame_of(gid, c.inv_frame)];
    const Loc lc = locate(f, gid - f.sample_base);
    if (lc.valid) grid[f.out_base + (uint64_t)lc.r * f.nc + lc.c] = c.out[gid];
  }
}

constexpr int kFinalizePer = 8;
__global__ __launch_bounds__(kThreads) void k_finalize(Common c) {
  const uint32_t total = *c.nactive;
  const uint32_t stride = gridDim.x * blockDim.x;
  unsigned long long credited = 0;  // the iterated samples' counts (the start kernel adds the rest)
  // kFinalizePer samples per thread per round, their loads issued together
  // (64-bit entry indices: a total near 2^32 must not wrap the stride)
  for (uint64_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)kFinalizePer * stride) {
    uint32_t pix[kFinalizePer];
    int it[kFinalizePer];
    double r2[kFinalizePer];
#pragma unroll
    for (int k = 0; k < kFinalizePer; k++) {
      const uint64_t e = i + (uint64_t)k * stride;
      pix[k] = e < total ? c.active[e] : c.total;
    }
#pragma unroll
    for (int k = 0; k < kFinalizePer; k++) {
      it[k] = pix[k] < c.total ? c.out[pix[k]].iterations : c.N;
      r2[k] = pix[k] < c.total ? c.r2[pix[k]] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kFinalizePer; k++) {
      if (pix[k] >= c.total) continue;
      credited += (unsigned long long)it[k];
      if (it[k] < c.N) c.out[pix[k]].smoothing = smooth_of(r2[k], c);
    }
  }
  const int idx[1] = {C_CREDITED};
  block_partial(c.part, kSlotFinal + blockIdx.x, &credited, idx, 1);
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

// sum of reported iterations over the output (credited work)
// partial slots -> one total per counter (one workgroup)
// one workgroup of 1024: each thread sums every counter of its slots (one
// contiguous C_COUNT-entry record per slot), then one wave reduction per counter
constexpr int kReduceThreads = 1024;
__global__ __launch_bounds__(kReduceThreads) void k_reduce_partials(const unsigned long long* part, int nslots,
                                                                    unsigned long long* totals) {
  __shared__ unsigned long long acc[kReduceThreads / 64][C_COUNT];
  unsigned long long v[C_COUNT] = {};
  for (int sl = threadIdx.x; sl < nslots; sl += blockDim.x)
#pragma unroll
    for (int k = 0; k < C_COUNT; k++) v[k] += part[(size_t)sl * C_COUNT + k];
#pragma unroll
  for (int k = 0; k < C_COUNT; k++) {
    const unsigned long long sk = wave_sum(v[k]);
    if (lane_id() == 0) acc[threadIdx.x >> 6][k] = sk;
  }
  __syncthreads();
  if (threadIdx.x < C_COUNT) {
    unsigned long long t = 0;
    for (int w = 0; w < kReduceThreads / 64; w++) t += acc[w][threadIdx.x];
    totals[threadIdx.x] = t;
  }
}

// clears the counter slots p[0, n) and the tile counts q[0, m)
__global__ void k_zero(unsigned long long* p, int n, uint32_t* q, uint32_t m) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) q[i] = 0u;
}

// ================================================================== engine
#define NM_HIP(expr)                              \
  do {                                            \
    hipError_t e_ = (expr);                       \
    if (e_ != hipSuccess) return fail(#expr, e_); \
  } while (0)

// threads per workgroup of a path's iterate kernel
template <class A>
constexpr int iter_threads() { return kThreads; }
template <>
constexpr int iter_threads<SeriesArgs>() { return kSeriesIterThreads; }

// workgroups of the tile kernels (start, scatter, finalize, emit)
static int tile_grid(uint64_t total, int num_cu) {
  return start_grid(total, num_cu, (uint64_t)(kSlotIter - kSlotStart));
}

// The kernels of one pass on `stream`: zero the counters (part: kPartWords
// words), start, scan, scatter, iterate, finalize (and emit for tiled sample
// orders).  evs != nullptr records an event before each stage (eager timing);
// under stream capture it is null.  Returns the first launch error.
template <class Args, class StartFn, class IterFn>
static hipError_t launch_pass_on(hipStream_t stream, int num_cu, unsigned long long* part, uint32_t* nactive,
                                 Args& args, StartFn start, IterFn iter, int waves, EscapeValue* emit_to,
                                 hipEvent_t* evs) {
  const Common& c = args.c;
  const uint32_t ntiles = (uint32_t)((c.total + kTile - 1) / kTile);
  const int gtile = tile_grid(c.total, num_cu);
  // waves = persistent waves per SIMD: 4 per CU for each 256 threads
  const int ti = iter_threads<Args>();
  const int giter = std::max(1, std::min(num_cu * waves * 256 / ti, kSlotFinal - kSlotIter));
  hipError_t e;
#define NM_LAUNCH_OK(expr)                   \
  do {                                       \
    expr;                                    \
    if ((e = hipGetLastError()) != hipSuccess) return e; \
  } while (0)
  NM_LAUNCH_OK(hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, stream, part, kPartWords, c.tile_count, ntiles));
  if (evs && (e = hipEventRecord(evs[0], stream)) != hipSuccess) return e;
  NM_LAUNCH_OK(hipLaunchKernelGGL(start, dim3(gtile), dim3(kThreads), 0, stream, args));
  if (evs && (e = hipEventRecord(evs[1], stream)) != hipSuccess) return e;
  const ScanSplit sp = scan_split(ntiles);
  uint32_t* gsum = scan_group_sums(part);
  NM_LAUNCH_OK(hipLaunchKernelGGL(k_tile_sums, dim3(sp.groups), dim3(1024), 0, stream, (const uint32_t*)c.tile_count,
                                  ntiles, sp.per, gsum));
  NM_LAUNCH_OK(hipLaunchKernelGGL(k_scan_tiles, dim3(sp.groups), dim3(1024), 0, stream, c.tile_count, ntiles, sp.per,
                                  (const uint32_t*)gsum, nactive));
  NM_LAUNCH_OK(hipLaunchKernelGGL(k_scatter, dim3(gtile), dim3(kThreads), 0, stream, (const uint64_t*)c.hand,
                                  (const uint32_t*)c.tile_count, c.total, c.active));
  if (evs && (e = hipEventRecord(evs[2], stream)) != hipSuccess) return e;
  NM_LAUNCH_OK(hipLaunchKernelGGL(iter, dim3(giter), dim3(ti), 0, stream, args));
  if (evs && (e = hipEventRecord(evs[3], stream)) != hipSuccess) return e;
  static_assert(kPartSlots - kSlotFinal >= kSlotIter - kSlotStart, "finalize runs on the start grid");
  NM_LAUNCH_OK(hipLaunchKernelGGL(k_finalize, dim3(gtile), dim3(kThreads), 0, stream, c));
  if (emit_to) NM_LAUNCH_OK(hipLaunchKernelGGL(k_emit, dim3(gtile), dim3(kThreads), 0, stream, c, emit_to));
  if (evs && (e = hipEventRecord(evs[4], stream)) != hipSuccess) return e;
#undef NM_LAUNCH_OK
  return hipSuccess;
}

// log2_tab's table: r_i = nearest(1 / c_i), l_i = -log2 r_i in long double
static hipError_t fill_log2_table() {
  double2 tab[kLog2Tab];
  for (int i = 0; i < kLog2Tab; i++) {
    const long double ci = 1.0L + ((long double)i + 0.5L) / kLog2Tab;
    const double r = (double)(1.0L / ci);
    tab[i] = make_double2(r, (double)(-log2l((long double)r)));
  }
  return hipMemcpyToSymbol(HIP_SYMBOL(d_log2tab), tab, sizeof tab);
}

class EngineImpl : public Engine {
 public:
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> evpool;  // stage boundaries of timed passes (grow-only)
  unsigned long long* d_totals = nullptr;  // C_COUNT counter totals
  int device = 0;
  int num_cu = 256;
  std::string err;
  // persistent buffers (grow-only)
  void* d_frames = nullptr; size_t n_frames_bytes = 0;
  void* d_tabs = nullptr; size_t n_tabs_bytes = 0;
  void* d_mask = nullptr; size_t n_mask_bytes = 0;
  void* d_out = nullptr; size_t n_out_bytes = 0;
  void* d_hand = nullptr; size_t n_hand_bytes = 0;
  void* d_r2 = nullptr; size_t n_r2_bytes = 0;
  void* d_tiles = nullptr; size_t n_tiles_bytes = 0;
  void* d_active = nullptr; size_t n_active_bytes = 0;
  void* d_state = nullptr; size_t n_state_bytes = 0;
  void* d_state_i = nullptr; size_t n_state_i_bytes = 0;
  void* d_ref = nullptr; size_t n_ref_bytes = 0;
  void* d_grid = nullptr; size_t n_grid_bytes = 0;  // downsample staging
  void* d_pal = nullptr; size_t n_pal_bytes = 0;    // colouring: palette (RGBA) and image
  void* d_img = nullptr; size_t n_img_bytes = 0;
  void* d_keys = nullptr; size_t n_keys_bytes = 0;   // video: two keyframes and the schedule
  void* d_vframes = nullptr; size_t n_vframes_bytes = 0;
  uint64_t tag_ = 0;                                 // render count (render_tag)
  const EscapeValue* last_result = nullptr;          // the last render's row-major grids
  uint64_t last_nout = 0;
  void* d_emit = nullptr; size_t n_emit_bytes = 0;  // row-major grids of a tiled sample order
  void* d_ttabs = nullptr; size_t n_ttabs_bytes = 0;  // tile search: epsilon tables
  void* d_tio = nullptr; size_t n_tio_bytes = 0;      // tile search: probes, predictions, results
  void* d_tile_ua = nullptr; size_t n_tile_ua_bytes = 0;  // per-tile skips of a render
  unsigned long long* d_part = nullptr;               // kPartSlots x C_COUNT
  uint32_t* d_nactive = nullptr;
  uint64_t ref_serial = 0;
  // reference layout (offsets into d_ref)
  int ref_len = 0;
  const double2 *rZd = nullptr, *rXlo = nullptr, *rA = nullptr, *rB = nullptr, *rC = nullptr;
  const double2 *rAm = nullptr, *rBm = nullptr, *rCm = nullptr;
  const int2 *rAe = nullptr, *rBe = nullptr, *rCe = nullptr;
  int rfxs = 0;
  const double *rPM = nullptr, *rBX = nullptr, *rBN = nullptr, *rCX = nullptr, *rCN = nullptr;
  const int* rSP = nullptr;
  const double *rST = nullptr, *rBNs = nullptr, *rCNs = nullptr;
  int rst_levels = 0;
  const int* rPMT = nullptr;
  int rpmt_n = 0;
  double rpmt_base = 0.0, rpmt_inv_h = 0.0;
  int ref_zd_count = 0;

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
    NM_HIP(fill_log2_table());
    NM_HIP(hipMalloc((void**)&d_totals, sizeof(unsigned long long) * C_COUNT));
    NM_HIP(hipMalloc((void**)&d_part, sizeof(unsigned long long) * kPartWords));  // slots + claim counters
    NM_HIP(hipMalloc((void**)&d_nactive, 256));
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
    // Zd padded with zeros so idle lanes can run a block from index 0 and the
    // LDS orbit head (kHead + 1 entries) always exists
    const size_t nz = (size_t)len + 2 + kHead + 64;
    ref_zd_count = (int)nz;
    size_t off = 0;
    auto place = [&](size_t bytes) {
      size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    size_t oZ = place(nz * 16), oXlo = place((size_t)len * 16), oA = place((size_t)len * 16),
           oB = place((size_t)len * 16), oC = place((size_t)len * 16), oPM = place((size_t)len * 8),
           oBX = place((size_t)len * 8), oBN = place((size_t)len * 8), oCX = place((size_t)len * 8),
           oCN = place((size_t)len * 8), oSP = place((size_t)len * 4), oST = place(ref.ST.size() * 8),
           oBNs = place(ref.BNs.size() * 8), oCNs = place(ref.CNs.size() * 8), oPMT = place(ref.PMT.size() * 4);
    const bool fxs = ref.fxs && ref.Am.size() == 2 * (size_t)len;
    const size_t nfm = fxs ? (size_t)len * 16 : 0, nfe = fxs ? (size_t)len * 8 : 0;
    size_t oAm = place(nfm), oBm = place(nfm), oCm = place(nfm), oAe = place(nfe), oBe = place(nfe), oCe = place(nfe);
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
    std::memcpy(h.data() + oST, ref.ST.data(), ref.ST.size() * 8);
    std::memcpy(h.data() + oBNs, ref.BNs.data(), ref.BNs.size() * 8);
    std::memcpy(h.data() + oCNs, ref.CNs.data(), ref.CNs.size() * 8);
    if (!ref.PMT.empty()) std::memcpy(h.data() + oPMT, ref.PMT.data(), ref.PMT.size() * 4);
    if (fxs) {
      std::memcpy(h.data() + oAm, ref.Am.data(), nfm);
      std::memcpy(h.data() + oBm, ref.Bm.data(), nfm);
      std::memcpy(h.data() + oCm, ref.Cm.data(), nfm);
      std::memcpy(h.data() + oAe, ref.Ae.data(), nfe);
      std::memcpy(h.data() + oBe, ref.Be.data(), nfe);
      std::memcpy(h.data() + oCe, ref.Ce.data(), nfe);
    }
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
    rST = (const double*)(b + oST);
    rBNs = (const double*)(b + oBNs);
    rCNs = (const double*)(b + oCNs);
    rst_levels = ref.st_levels;
    rPMT = (const int*)(b + oPMT);
    rpmt_n = ref.PMT.empty() ? 0 : ref.pmt_n;
    rpmt_base = ref.pmt_base;
    rpmt_inv_h = ref.pmt_inv_h;
    rfxs = fxs ? 1 : 0;
    rAm = fxs ? (const double2*)(b + oAm) : nullptr;
    rBm = fxs ? (const double2*)(b + oBm) : nullptr;
    rCm = fxs ? (const double2*)(b + oCm) : nullptr;
    rAe = fxs ? (const int2*)(b + oAe) : nullptr;
    rBe = fxs ? (const int2*)(b + oBe) : nullptr;
    rCe = fxs ? (const int2*)(b + oCe) : nullptr;
    ref_len = len;
    ref_serial = serial;
    return NM_OK;
  }

  // frame tables on the device belong to (set serial, sample order); a render
  // with the same pair and unmoved buffers reuses them
  uint64_t tab_serial = 0;
  int tab_tile = -1;
  bool tables_stale(const RenderOpts& opt, bool same_buffers) const {
    return opt.set_serial == 0 || opt.set_serial != tab_serial || opt.tile != tab_tile || !same_buffers;
  }
  void tables_uploaded(const RenderOpts& opt) {
    tab_serial = opt.set_serial;
    tab_tile = opt.tile;
  }

  // buffers and the common kernel arguments for `total` samples
  int prepare_common(Common& c, uint64_t total, const RenderOpts& opt) {
    // sample ids and the plain iterate kernel's chunk positions (ChunkMap) stay below 2^32
    if (total == 0 || total >= 0xFFFFF000ull) return NM_ERR_RANGE;
    uint64_t ntiles = (total + kTile - 1) / kTile;
    int st;
    if ((st = grow(&d_out, &n_out_bytes, total * sizeof(EscapeValue)))) return st;
    if ((st = grow(&d_hand, &n_hand_bytes, (uint64_t)hand_words((uint32_t)total) * 8))) return st;
    if ((st = grow(&d_r2, &n_r2_bytes, total * 8))) return st;
    if ((st = grow(&d_state, &n_state_bytes, total * sizeof(double4)))) return st;
    if ((st = grow(&d_state_i, &n_state_i_bytes, total * sizeof(int2)))) return st;
    if ((st = grow(&d_tiles, &n_tiles_bytes, ntiles * 4))) return st;
    if ((st = grow(&d_active, &n_active_bytes, total * 4))) return st;
    c.frames = (const DevFrame*)d_frames;
    c.total = (uint32_t)total;
    c.out = (EscapeValue*)d_out;
    c.r2 = (double*)d_r2;
    c.state = (double4*)d_state;
    c.state_i = (int2*)d_state_i;
    c.hand = (uint64_t*)d_hand;
    c.tile_count = (uint32_t*)d_tiles;
    c.active = (uint32_t*)d_active;
    c.nactive = d_nactive;
    c.part = d_part;
    c.N = opt.N;
    c.b2 = opt.bailout * opt.bailout;
    c.b2m = c.b2 * (1.0 - 1.0 / 1099511627776.0);
    c.log_b = std::log(opt.bailout);
    c.smooth_k = opt.bailout >= 2.0 ? 0.5 * std::log(2.0) / c.log_b : 0.0;
    return NM_OK;
  }

  template <class Args, class StartFn, class IterFn>
  int launch_pass(Args& args, StartFn start, IterFn iter, int waves, EscapeValue* emit_to, hipEvent_t* evs) {
    NM_HIP(launch_pass_on(stream, num_cu, d_part, d_nactive, args, start, iter, waves, emit_to, evs));
    return NM_OK;
  }

  int ensure_events(size_t n) {
    while (evpool.size() < n) {
      hipEvent_t e = nullptr;
      NM_HIP(hipEventCreate(&e));
      evpool.push_back(e);
    }
    return NM_OK;
  }

  // The pass `repeats` times back to back (one host synchronisation at the end);
  // then counters and output.  Without opt.graph every pass runs eagerly with
  // per-stage events; with it the first pass runs eagerly, the pass is captured
  // once into a HIP graph and the other repeats replay it (timed as a whole).
  template <class Args, class StartFn, class IterFn>
  int run(Args& args, StartFn start, IterFn iter, int waves, const RenderOpts& opt, RenderStats* stats,
          uint64_t nout, EscapeValue* host_out, EscapeValue* emit_to) {
    const int repeats = std::max(1, opt.repeats);
    const bool use_graph = opt.graph && repeats > 1;
    int st;
    if ((st = ensure_events(use_graph ? 5 + 2 * (size_t)repeats : 5 * (size_t)repeats))) return st;
    hipEvent_t* ev = evpool.data();
    const int eager = use_graph ? 1 : repeats;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
      // captured and instantiated before any pass runs, so the host work of the
      // instantiation does not overlap the timed eager pass
      NM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      st = launch_pass(args, start, iter, waves, emit_to, nullptr);
      hipError_t ce = hipStreamEndCapture(stream, &graph);
      if (st) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      if (ce != hipSuccess) return fail("hipStreamEndCapture", ce);
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail("hipGraphInstantiate", ie);
      }
    }
    for (int rep = 0; rep < eager && st == NM_OK; rep++)
      st = launch_pass(args, start, iter, waves, emit_to, ev + 5 * rep);
    if (use_graph) {
      for (int rep = 1; rep < repeats && st == NM_OK; rep++) {
        hipEvent_t* e = ev + 5 + 2 * (rep - 1);
        hipError_t e1 = hipEventRecord(e[0], stream);
        hipError_t e2 = e1 == hipSuccess ? hipGraphLaunch(exec, stream) : e1;
        hipError_t e3 = e2 == hipSuccess ? hipEventRecord(e[1], stream) : e2;
        if (e3 != hipSuccess) st = fail("graph replay", e3);
      }
    }
    hipError_t se = hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (st) return st;
    if (se != hipSuccess) return fail("hipStreamSynchronize", se);
    double ms[4] = {};
    for (int rep = 0; rep < eager; rep++)
      for (int k = 0; k < 4; k++) {
        float t = 0;
        NM_HIP(hipEventElapsedTime(&t, ev[5 * rep + k], ev[5 * rep + k + 1]));
        ms[k] += t / eager;
      }
    double ms_graph = 0.0;
    for (int rep = 1; use_graph && rep < repeats; rep++) {
      float t = 0;
      NM_HIP(hipEventElapsedTime(&t, ev[5 + 2 * (rep - 1)], ev[5 + 2 * (rep - 1) + 1]));
      ms_graph += t / (repeats - 1);
    }
    const EscapeValue* result = emit_to ? emit_to : (const EscapeValue*)d_out;
    tag_++;
    last_result = result;
    last_nout = nout;
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kReduceThreads), 0, stream, (const unsigned long long*)d_part,
                       kPartSlots, d_totals);
    NM_HIP(hipGetLastError());
    uint32_t nact = 0;
    unsigned long long h[C_COUNT] = {};
    NM_HIP(hipMemcpyAsync(h, d_totals, sizeof h, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipMemcpyAsync(&nact, d_nactive, 4, hipMemcpyDeviceToHost, stream));
    if (host_out)
      NM_HIP(hipMemcpyAsync(host_out, result, nout * sizeof(EscapeValue), hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    if (stats) {
      stats->samples = nout;
      stats->n_cardioid = h[C_CARDIOID];
      stats->n_sa_resolved = h[C_SA_RES];
      stats->n_iterated = nact;
      stats->executed = h[C_EXECUTED] + h[C_EXEC_START];
      stats->executed_start = h[C_EXEC_START];
      stats->skipped = h[C_SKIPPED];
      stats->n_rebased = h[C_REBASED];
      stats->credited = h[C_CREDITED];
      stats->ms_start = ms[0];
      stats->ms_compact = ms[1];
      stats->ms_iterate = ms[2];
      stats->ms_finalize = ms[3];
      stats->ms_total = use_graph ? ms_graph : ms[0] + ms[1] + ms[2] + ms[3];
      stats->graph = use_graph ? 1 : 0;
      stats->tile = emit_to ? opt.tile : 0;
    }
    return NM_OK;
  }

  int render_hw(const std::vector<const HwFrameHost*>& frames, const RenderOpts& opt, EscapeValue* host_out,
                RenderStats* stats) override {
    if (frames.empty()) return NM_ERR_ARG;
    std::vector<DevFrame> df;
    uint64_t total = 0;
    size_t ncol = 0, nrow = 0;
    for (auto* f : frames) {
      if (!f || f->nr <= 0 || f->nc <= 0) return NM_ERR_ARG;
      if (f->nr != frames[0]->nr || f->nc != frames[0]->nc) return NM_ERR_ARG;  // one frame size per set
      ncol += f->nc;
      nrow += f->nr;
    }
    const uint64_t fs = frame_samples(frames[0]->nr, frames[0]->nc, opt.tile);
    if (fs == 0) return NM_ERR_ARG;
    uint64_t nout = 0;
    size_t tc = 0, tr = 0;
    for (auto* f : frames) {
      DevFrame d = make_frame(df.size(), f->nr, f->nc, opt.tile, (uint32_t)fs);
      d.tab_c = (uint32_t)tc;
      d.tab_r = (uint32_t)tr;
      d.card_mode = f->card_mode;
      tc += f->cre.size();
      tr += f->cim.size();
      total += fs;
      nout += (uint64_t)f->nr * f->nc;
      df.push_back(d);
    }
    int st;
    const void *pf = d_frames, *pt = d_tabs;
    if ((st = grow(&d_frames, &n_frames_bytes, df.size() * sizeof(DevFrame)))) return st;
    if ((st = grow(&d_tabs, &n_tabs_bytes, (2 * ncol + 2 * nrow) * sizeof(double)))) return st;
    const bool upload = tables_stale(opt, pf == d_frames && pt == d_tabs);
    PlainArgs a;
    if ((st = prepare_common(a.c, total, opt))) return st;
    EscapeValue* emit_to = nullptr;
    if (opt.tile) {
      if ((st = grow(&d_emit, &n_emit_bytes, nout * sizeof(EscapeValue)))) return st;
      emit_to = (EscapeValue*)d_emit;
    }
    a.c.nframes = (int)df.size();
    a.c.frame_size = (uint32_t)fs;
    a.c.inv_frame = 1.0 / (double)a.c.frame_size;
    double* t = (double*)d_tabs;
    if (upload) {
      // the coordinate tables of every frame, concatenated (built only when the
      // device copy is stale: a repeated frame set reuses it)
      std::vector<double> cre, crelo, cim, cimlo;
      cre.reserve(ncol); crelo.reserve(ncol); cim.reserve(nrow); cimlo.reserve(nrow);
      for (auto* f : frames) {
        cre.insert(cre.end(), f->cre.begin(), f->cre.end());
        crelo.insert(crelo.end(), f->crelo.begin(), f->crelo.end());
        cim.insert(cim.end(), f->cim.begin(), f->cim.end());
        cimlo.insert(cimlo.end(), f->cimlo.begin(), f->cimlo.end());
      }
      NM_HIP(hipMemcpyAsync(d_frames, df.data(), df.size() * sizeof(DevFrame), hipMemcpyHostToDevice, stream));
      NM_HIP(hipMemcpyAsync(t, cre.data(), ncol * 8, hipMemcpyHostToDevice, stream));
      NM_HIP(hipMemcpyAsync(t + ncol, crelo.data(), ncol * 8, hipMemcpyHostToDevice, stream));
      NM_HIP(hipMemcpyAsync(t + 2 * ncol, cim.data(), nrow * 8, hipMemcpyHostToDevice, stream));
      NM_HIP(hipMemcpyAsync(t + 2 * ncol + nrow, cimlo.data(), nrow * 8, hipMemcpyHostToDevice, stream));
      NM_HIP(hipStreamSynchronize(stream));  // the host vectors end with this scope
      tables_uploaded(opt);
    }
    a.pre_iter = opt.pre_iter > 0 ? opt.pre_iter : kPreIterDefault;
    a.k_start = (std::min(a.pre_iter, opt.N) + 3) / 4 * 4;
    a.cre = t;
    a.crelo = t + ncol;
    a.cim = t + 2 * ncol;
    a.cimlo = t + 2 * ncol + nrow;
    const int K = opt.block_k > 0 ? opt.block_k : 64;
    // 4 persistent waves per SIMD: the iterate kernel's tail is shorter than
    // with 8 (512 frames of 640x480 at N 1024: 0.93 against 0.95 ms; 2 waves
    // hide too little latency, 1.03 ms)
    const int waves = opt.waves_per_simd > 0 ? opt.waves_per_simd : kPlainWaves;
    if (stats) {
      stats->block_k = K;
      stats->waves_per_simd = waves;
      stats->chunk_run = 1;
    }
    // 256-thread workgroups = 4 waves = one per SIMD: num_cu * waves workgroups
    // the folded update (plain_update) needs every frame's coordinates clear of
    // tiny nonzero values and a bailout of at most 2^400
    bool fma2 = opt.bailout <= 0x1p400, grouped = true;
    for (auto* f : frames) {
      fma2 = fma2 && f->fma_ok;
      grouped = grouped && grouped_test_ok(opt.bailout, f->c_bound);
    }
    if (stats) {
      stats->fma2 = fma2 ? 1 : 0;
      stats->grouped = grouped ? 1 : 0;
    }
#define NM_PLAIN(KK)                                                                                           \
  return fma2 && grouped ? run(a, k_plain_start<true, true>, k_plain_iterate<KK, true, true>, waves, opt, stats,  \
                            nout, host_out, emit_to)                                                           \
         : fma2      ? run(a, k_plain_start<true, false>, k_plain_iterate<KK, true, false>, waves, opt, stats, \
                           nout, host_out, emit_to)                                                            \
         : grouped   ? run(a, k_plain_start<false, true>, k_plain_iterate<KK, false, true>, waves, opt, stats, \
                           nout, host_out, emit_to)                                                            \
                     : run(a, k_plain_start<false, false>, k_plain_iterate<KK, false, false>, waves, opt,      \
                           stats, nout, host_out, emit_to)
    switch (K) {
      case 8: NM_PLAIN(8);
      case 16: NM_PLAIN(16);
      case 32: NM_PLAIN(32);
      case 64: NM_PLAIN(64);
      default: return NM_ERR_ARG;
    }
#undef NM_PLAIN
  }

  int render_sa(const std::vector<const SaFrameHost*>& frames, const RenderOpts& opt, EscapeValue* host_out,
                RenderStats* stats) override {
    if (frames.empty() || ref_len < 1) return NM_ERR_STATE;
    std::vector<DevFrame> df;
    size_t ncol = 0, nrow = 0, nmask = 0;
    for (auto* f : frames) {
      if (!f || f->nr <= 0 || f->nc <= 0) return NM_ERR_ARG;
      if (f->nr != frames[0]->nr || f->nc != frames[0]->nc) return NM_ERR_ARG;  // one frame size per set
      ncol += f->nc;
      nrow += f->nr;
      if (f->card_mode == CARD_MASK) nmask += f->mask.size();
    }
    uint64_t total = 0, nout = 0;
    const uint64_t fs = frame_samples(frames[0]->nr, frames[0]->nc, opt.tile);
    if (fs == 0) return NM_ERR_ARG;
    size_t tc = 0, tr = 0, tm = 0;
    for (auto* f : frames) {
      DevFrame d = make_frame(df.size(), f->nr, f->nc, opt.tile, (uint32_t)fs);
      d.tab_c = (uint32_t)tc;
      d.tab_r = (uint32_t)tr;
      d.card_mode = f->card_mode;
      d.mask_base = (uint64_t)tm * 32;
      if (f->card_mode == CARD_MASK) {
        if (f->mask.size() * 32 < (size_t)f->nr * f->nc) return NM_ERR_ARG;
        tm += f->mask.size();
      }
      tc += f->epsre.size();
      tr += f->epsim.size();
      total += fs;
      nout += (uint64_t)f->nr * f->nc;
      df.push_back(d);
    }
    int st;
    const void *pf = d_frames, *pt = d_tabs, *pm = d_mask;
    if ((st = grow(&d_frames, &n_frames_bytes, df.size() * sizeof(DevFrame)))) return st;
    if ((st = grow(&d_tabs, &n_tabs_bytes, (ncol + nrow) * sizeof(double)))) return st;
    if ((st = grow(&d_mask, &n_mask_bytes, std::max<size_t>(nmask, 1) * 4))) return st;
    const bool upload = tables_stale(opt, pf == d_frames && pt == d_tabs && pm == d_mask);
    SeriesArgs a;
    if ((st = prepare_common(a.c, total, opt))) return st;
    EscapeValue* emit_to = nullptr;
    if (opt.tile) {
      if ((st = grow(&d_emit, &n_emit_bytes, nout * sizeof(EscapeValue)))) return st;
      emit_to = (EscapeValue*)d_emit;
    }
    a.c.nframes = (int)df.size();
    a.c.frame_size = (uint32_t)fs;
    a.c.inv_frame = 1.0 / (double)a.c.frame_size;
    double* t = (double*)d_tabs;
    if (upload) {
      // every frame's delta tables and cardioid masks, concatenated (built only
      // when the device copy is stale: a repeated frame set reuses it)
      std::vector<double> er, ei;
      std::vector<uint32_t> mask;
      er.reserve(ncol); ei.reserve(nrow); mask.reserve(nmask);
      for (auto* f : frames) {
        if (f->card_mode == CARD_MASK) mask.insert(mask.end(), f->mask.begin(), f->mask.end());
        er.insert(er.end(), f->epsre.begin(), f->epsre.end());
        ei.insert(ei.end(), f->epsim.begin(), f->epsim.end());
      }
      NM_HIP(hipMemcpyAsync(d_frames, df.data(), df.size() * sizeof(DevFrame), hipMemcpyHostToDevice, stream));
      NM_HIP(hipMemcpyAsync(t, er.data(), ncol * 8, hipMemcpyHostToDevice, stream));
      NM_HIP(hipMemcpyAsync(t + ncol, ei.data(), nrow * 8, hipMemcpyHostToDevice, stream));
      if (!mask.empty())
        NM_HIP(hipMemcpyAsync(d_mask, mask.data(), mask.size() * 4, hipMemcpyHostToDevice, stream));
      NM_HIP(hipStreamSynchronize(stream));  // the host vectors end with this scope
      tables_uploaded(opt);
    }
    a.epsre = t;
    a.epsim = t + ncol;
    a.mask = (const uint32_t*)d_mask;
    fill_reference(a, opt);
    a.tile_ua = nullptr;
    a.skip_T = a.skip_ntc = a.skip_tpf = 0;
    // default chunk order: runs of 4 when the deltas follow a floatexp series
    // (the 1e-100 and 1e-200 views: FP64 iterate kernel 15% and 8% faster, FP32X
    // 8% and 11%; their long samples cluster), 64 consecutive samples otherwise
    // (the 1e-30 and 1e-40 views: 8% and 0-6% slower with runs of 4)
    const int run_shift = chunk_run_shift(opt.chunk_run == 0 && a.fxs ? 4 : opt.chunk_run);
    if (run_shift < 0) return NM_ERR_ARG;
    a.run_shift = (uint32_t)run_shift;
    if (opt.tile_ua) {
      if (opt.skip_T < 8) return NM_ERR_ARG;
      const uint32_t T = (uint32_t)opt.skip_T, ntc = (frames[0]->nc + T - 1) / T, ntr = (frames[0]->nr + T - 1) / T;
      const size_t nt = (size_t)ntc * ntr * frames.size();
      if ((st = grow(&d_tile_ua, &n_tile_ua_bytes, nt * 4))) return st;
      NM_HIP(hipMemcpyAsync(d_tile_ua, opt.tile_ua, nt * 4, hipMemcpyHostToDevice, stream));
      a.tile_ua = (const int*)d_tile_ua;
      a.skip_T = T;
      a.skip_ntc = ntc;
      a.skip_tpf = ntc * ntr;
    }
    const bool fx = opt.delta_mode == DELTA_FP32X;
    // FP64 blocks hold one orbit entry per lane whatever K is (64 VGPRs), so the
    // longest block amortises its set-up best; FP32X blocks renormalise their
    // mantissa at block ends, so they stay short
    const int K = opt.block_k > 0 ? opt.block_k : fx ? 8 : 64;
    // FP64: 4 persistent waves per SIMD (2 of the 512-thread workgroups per CU;
    // against 8: the 1e-30 deep-zoom frame within 1%, the 1e-100 and 1e-200
    // views 8% and 11% faster, their waves stalling less on the HBM path);
    // an even count keeps the workgroups on the CUs evenly
    const int waves = opt.waves_per_simd > 0 ? opt.waves_per_simd : fx ? 8 : 4;
    if (stats) {
      stats->block_k = K;
      stats->waves_per_simd = waves;
      stats->chunk_run = 1 << a.run_shift;
    }
    auto go = [&](auto iter) {
      return a.fxs ? run(a, k_series_start<true>, iter, waves, opt, stats, nout, host_out, emit_to)
                   : run(a, k_series_start<false>, iter, waves, opt, stats, nout, host_out, emit_to);
    };
    switch (K) {
      case 8: return fx ? go(k_series_iterate<8, true>) : go(k_series_iterate<8, false>);
      case 12: return fx ? go(k_series_iterate<12, true>) : go(k_series_iterate<12, false>);
      case 16: return fx ? go(k_series_iterate<16, true>) : go(k_series_iterate<16, false>);
      case 32: return fx ? go(k_series_iterate<32, true>) : go(k_series_iterate<32, false>);
      case 64:  // FP64 only: an FP32X mantissa is renormalised at block ends
        if (fx) return NM_ERR_ARG;
        return go(k_series_iterate<64, false>);
      default: return NM_ERR_ARG;
    }
  }

  void fill_reference(SeriesArgs& a, const RenderOpts& opt) const {
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
    a.ST = rST;
    a.BNs = rBNs;
    a.CNs = rCNs;
    a.st_levels = rst_levels;
    a.PMT = rPMT;
    a.pmt_n = rpmt_n;
    a.pmt_base = rpmt_base;
    a.pmt_inv_h = rpmt_inv_h;
    a.len = ref_len;
    a.mend = ref_len + 1;
    a.zd_count = ref_zd_count;
    a.tol = opt.tol;
    a.ltol = std::log2(opt.tol);
    a.fxs = rfxs;
    a.Am = rAm;
    a.Bm = rBm;
    a.Cm = rCm;
    a.Ae = rAe;
    a.Be = rBe;
    a.Ce = rCe;
  }

  int tile_skip(const std::vector<const SaFrameHost*>& frames, const RenderOpts& opt, TileSearch& ts) override {
    if (frames.empty() || ref_len < 1) return NM_ERR_STATE;
    if (ts.T < 8 || ts.mode < 0 || ts.mode > 2) return NM_ERR_ARG;
    const int nr = frames[0]->nr, nc = frames[0]->nc, T = ts.T;
    const int ntc = (nc + T - 1) / T, ntr = (nr + T - 1) / T;
    const size_t nt = (size_t)ntc * ntr * frames.size();
    if (ts.mode == 2 && ts.pred.size() != nt) return NM_ERR_ARG;
    std::vector<double> eps;
    std::vector<int2> probe(nt);
    size_t ncol = 0, nrow = 0;
    for (auto* f : frames) {
      if (!f || f->nr != nr || f->nc != nc) return NM_ERR_ARG;
      ncol += nc;
      nrow += nr;
    }
    eps.reserve(ncol + nrow);
    for (auto* f : frames) eps.insert(eps.end(), f->epsre.begin(), f->epsre.end());
    for (auto* f : frames) eps.insert(eps.end(), f->epsim.begin(), f->epsim.end());
    for (size_t k = 0; k < frames.size(); k++) {
      const SaFrameHost& f = *frames[k];
      for (int tr = 0; tr < ntr; tr++)
        for (int tc = 0; tc < ntc; tc++) {
          // the probe: the tile's sample farthest from the reference point; the
          // epsilon tables are separable and monotone, so it is a corner
          const int c0 = tc * T, c1 = std::min(nc, c0 + T) - 1, r0 = tr * T, r1 = std::min(nr, r0 + T) - 1;
          const int pc = std::fabs(f.epsre[c1]) > std::fabs(f.epsre[c0]) ? c1 : c0;
          const int pr = std::fabs(f.epsim[r1]) > std::fabs(f.epsim[r0]) ? r1 : r0;
          probe[k * (size_t)ntc * ntr + (size_t)tr * ntc + tc] = make_int2((int)(k * nc + pc), (int)(k * nr + pr));
        }
    }
    int st;
    if ((st = grow(&d_ttabs, &n_ttabs_bytes, eps.size() * 8))) return st;
    if ((st = grow(&d_tio, &n_tio_bytes, nt * 20))) return st;
    char* io = (char*)d_tio;
    int2* dprobe = (int2*)io;
    int* dpred = (int*)(io + nt * 8);
    int* dua = (int*)(io + nt * 12);
    int* dhow = (int*)(io + nt * 16);
    std::vector<int> evals_host(nt);
    int* devals = nullptr;
    if ((st = grow(&d_tile_ua, &n_tile_ua_bytes, nt * 4))) return st;
    devals = (int*)d_tile_ua;
    NM_HIP(hipMemcpyAsync(d_ttabs, eps.data(), eps.size() * 8, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(dprobe, probe.data(), nt * 8, hipMemcpyHostToDevice, stream));
    if (ts.mode == 2) NM_HIP(hipMemcpyAsync(dpred, ts.pred.data(), nt * 4, hipMemcpyHostToDevice, stream));
    SeriesArgs a{};
    a.epsre = (const double*)d_ttabs;
    a.epsim = (const double*)d_ttabs + ncol;
    fill_reference(a, opt);
    if ((st = ensure_events(2))) return st;
    NM_HIP(hipEventRecord(evpool[0], stream));
    hipLaunchKernelGGL(k_tile_skip, dim3((unsigned)nt), dim3(64), 0, stream, a, (const int2*)dprobe,
                       (const int*)dpred, ts.mode, (int)nt, dua, dhow, devals);
    NM_HIP(hipGetLastError());
    NM_HIP(hipEventRecord(evpool[1], stream));
    ts.ua.resize(nt);
    ts.how.resize(nt);
    ts.evals.resize(nt);
    NM_HIP(hipMemcpyAsync(ts.ua.data(), dua, nt * 4, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipMemcpyAsync(ts.how.data(), dhow, nt * 4, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipMemcpyAsync(ts.evals.data(), devals, nt * 4, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    NM_HIP(hipEventElapsedTime(&ms, evpool[0], evpool[1]));
    ts.ms = ms;
    return NM_OK;
  }

  uint64_t render_tag() const override { return tag_; }

  int colour(const EscapeValue* host_grid, uint64_t tag, size_t offset, int nr, int nc, int sc, int N,
             const uint8_t* pal_rgb, int npal, int smooth, uint8_t* host_out, double* ms) override {
    if (sc <= 0 || nr <= 0 || nc <= 0 || N <= 0 || npal < N || !pal_rgb || !host_out) return NM_ERR_ARG;
    const size_t nin = (size_t)nr * nc, nout = (size_t)(nr / sc) * (nc / sc);
    int st;
    const EscapeValue* grid;
    if (host_grid) {
      if ((st = grow(&d_grid, &n_grid_bytes, nin * sizeof(EscapeValue)))) return st;
      NM_HIP(hipMemcpyAsync(d_grid, host_grid, nin * sizeof(EscapeValue), hipMemcpyHostToDevice, stream));
      grid = (const EscapeValue*)d_grid;
    } else {
      if (tag != tag_ || !last_result || offset + nin > last_nout) return NM_ERR_STATE;
      grid = last_result + offset;
    }
    std::vector<uint8_t> rgba((size_t)npal * 4, 0);
    for (int i = 0; i < npal; i++)
      for (int k = 0; k < 3; k++) rgba[(size_t)4 * i + k] = pal_rgb[(size_t)3 * i + k];
    if ((st = grow(&d_pal, &n_pal_bytes, rgba.size()))) return st;
    if ((st = grow(&d_img, &n_img_bytes, std::max<size_t>(nout * 3, 1)))) return st;
    NM_HIP(hipMemcpyAsync(d_pal, rgba.data(), rgba.size(), hipMemcpyHostToDevice, stream));
    if ((st = ensure_events(2))) return st;
    NM_HIP(hipEventRecord(evpool[0], stream));
    if (colour_launch(stream, num_cu, grid, nr, nc, sc, N, d_pal, smooth, (uint8_t*)d_img))
      return fail("k_colour launch", hipGetLastError());
    NM_HIP(hipEventRecord(evpool[1], stream));
    if (nout) NM_HIP(hipMemcpyAsync(host_out, d_img, nout * 3, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    float t = 0.0f;
    NM_HIP(hipEventElapsedTime(&t, evpool[0], evpool[1]));
    if (ms) *ms = t;
    return NM_OK;
  }

  int video(const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc, int rate, uint8_t* out,
            double* ms) override {
    // the previous keyframe scaled by lg >= 1 must cover the canvas (blit)
    if (!prev || !next || !out || kr < nr || kc < nc || nr <= 0 || nc <= 0 || rate <= 0) return NM_ERR_ARG;
    const std::vector<VideoFrame> sched = video_schedule(kr, kc, nr, nc, rate);
    const size_t kbytes = (size_t)kr * kc * 3, sbytes = sched.size() * sizeof(VideoFrame);
    const size_t obytes = (size_t)nr * nc * 3 * (size_t)rate;
    int st;
    // layout: prev | next | schedule | widened keyframes, each 256-byte aligned
    const size_t kstep = (kbytes + 255) & ~(size_t)255;
    if ((st = grow(&d_keys, &n_keys_bytes, 2 * kstep + sbytes + 512 + (size_t)video_scratch_bytes(kr, kc)))) return st;
    if ((st = grow(&d_vframes, &n_vframes_bytes, obytes))) return st;
    uint8_t* kp = (uint8_t*)d_keys;
    uint8_t* kn = kp + kstep;
    VideoFrame* ks = (VideoFrame*)(((uintptr_t)(kn + kbytes) + 255) & ~(uintptr_t)255);
    NM_HIP(hipMemcpyAsync(kp, prev, kbytes, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(kn, next, kbytes, hipMemcpyHostToDevice, stream));
    NM_HIP(hipMemcpyAsync(ks, sched.data(), sbytes, hipMemcpyHostToDevice, stream));
    if ((st = ensure_events(2))) return st;
    NM_HIP(hipEventRecord(evpool[0], stream));
    uint8_t* wide = (uint8_t*)(((uintptr_t)(ks + sched.size()) + 255) & ~(uintptr_t)255);
    if (video_launch((void*)stream, kp, kn, kr, kc, nr, nc, ks, rate, wide, (uint8_t*)d_vframes))
      return fail("k_video_frames launch", hipGetLastError());
    NM_HIP(hipEventRecord(evpool[1], stream));
    NM_HIP(hipMemcpyAsync(out, d_vframes, obytes, hipMemcpyDeviceToHost, stream));
    NM_HIP(hipStreamSynchronize(stream));
    float t = 0.0f;
    NM_HIP(hipEventElapsedTime(&t, evpool[0], evpool[1]));
    if (ms) *ms = t;
    return NM_OK;
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


// ================================================ allocation-free plain render
// render_hw_noalloc: the plain path (getIterationsHW, mandelbrot.cpp:231-254) of
// one frame with every device buffer carved from `scratch`, a file-scope device
// array: coordinate tables, the frame record, per-sample results, hand-on words,
// states, |z|^2, the active list, tile counts and the counters.  It calls no
// hipMalloc/hipFree, creates no stream or event (the null stream) and touches no
// engine buffer, for callers that cannot allocate device memory.  Each device
// holds its own copy of the array; its address comes from hipGetSymbolAddress
// once per device (noalloc_setup) and only that address reaches kernels and
// copies.  Every call on a device shares the array: calls that overlap on one
// device race (the C boundary serialises its own calls).
constexpr uint64_t kNoAllocMaxSamples = 1ull << 19;
constexpr int kNoAllocMaxSide = 8192;
constexpr int kNoAllocMaxDevices = 64;
constexpr size_t na_align(size_t b) { return (b + 255) & ~(size_t)255; }
// bytes of the layout for n samples on an nr x nc frame (noalloc_layout)
constexpr size_t na_bytes(uint64_t n, uint64_t nr, uint64_t nc) {
  return na_align((2 * nc + 2 * nr) * sizeof(double)) + na_align(sizeof(DevFrame)) +
         na_align(n * sizeof(EscapeValue)) + na_align(n * sizeof(double)) + na_align(n * sizeof(double4)) +
         na_align((n + kTile - 1) / kTile * kHandWordsPerTile * 8) + na_align((n + kTile - 1) / kTile * 4) +
         na_align(n * 4) +
         na_align((size_t)kPartWords * sizeof(unsigned long long)) + na_align(4);
}
constexpr size_t kNoAllocScratchBytes =
    na_bytes(kNoAllocMaxSamples, kNoAllocMaxSide, kNoAllocMaxSide);  // about 29 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocDevice {
  void* ptr = nullptr;  // device address of this device's scratch
  int num_cu = 0;
};
static std::mutex g_noalloc_mu;
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// once per device, with that device current: the scratch address and the
// log2 table of the smoothing
static int noalloc_setup(int dev, NoAllocDevice** out, std::string& err) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    err = "device index beyond the allocation-free path's table";
    return NM_ERR_RANGE;
  }
  NoAllocDevice& d = g_noalloc[dev];
  if (!d.ptr) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return NM_ERR_HIP;
    }
    int cu = 0;
    e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = fill_log2_table();
    if (e != hipSuccess) {
      err = std::string("allocation-free set-up: ") + hipGetErrorString(e);
      return NM_ERR_HIP;
    }
    d.num_cu = cu > 0 ? cu : 256;
    d.ptr = p;
  }
  *out = &d;
  return NM_OK;
}

int render_hw_noalloc(const HwFrameView& f, int N, double bailout, EscapeValue* host_out, std::string& err) {
  if (f.nr <= 0 || f.nc <= 0 || N <= 0 || !(bailout > 0.0) || !host_out || !f.cre || !f.crelo || !f.cim ||
      !f.cimlo) {
    err = "bad argument";
    return NM_ERR_ARG;
  }
  const uint64_t n = (uint64_t)f.nr * f.nc;
  if (n > kNoAllocMaxSamples || f.nr > kNoAllocMaxSide || f.nc > kNoAllocMaxSide) {
    err = "frame larger than the allocation-free path's scratch (at most 2^19 samples, 8192 rows or columns)";
    return NM_ERR_RANGE;
  }
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    err = "no HIP device";
    return NM_ERR_NODEVICE;
  }
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return NM_ERR_HIP;
  }
  NoAllocDevice* d = nullptr;
  int st = noalloc_setup(dev, &d, err);
  if (st != NM_OK) return st;
  // the layout: tables, frame, out, r2, state, hand-on words, tile counts, active,
  // counters, active count
  char* b = (char*)d->ptr;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = b + off;
    off += na_align(bytes);
    return q;
  };
  double* tabs = (double*)take((2 * (size_t)f.nc + 2 * (size_t)f.nr) * sizeof(double));
  DevFrame* frame = (DevFrame*)take(sizeof(DevFrame));
  EscapeValue* out = (EscapeValue*)take(n * sizeof(EscapeValue));
  double* r2 = (double*)take(n * sizeof(double));
  double4* state = (double4*)take(n * sizeof(double4));
  uint64_t* hand = (uint64_t*)take((size_t)hand_words((uint32_t)n) * 8);
  uint32_t* tiles = (uint32_t*)take((n + kTile - 1) / kTile * 4);
  uint32_t* active = (uint32_t*)take(n * 4);
  unsigned long long* part = (unsigned long long*)take((size_t)kPartWords * sizeof(unsigned long long));
  uint32_t* nactive = (uint32_t*)take(4);
  if (off > kNoAllocScratchBytes) {  // cannot happen for frames within the limits above
    err = "allocation-free layout overflow";
    return NM_ERR_RANGE;
  }
  DevFrame df = make_frame(0, f.nr, f.nc, 0, (uint32_t)n);
  df.card_mode = f.card_mode;
  hipStream_t s = nullptr;  // the null stream: nothing is created
#define NA_HIP(expr)                                               \
  do {                                                             \
    hipError_t e_ = (expr);                                        \
    if (e_ != hipSuccess) {                                        \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);     \
      return NM_ERR_HIP;                                           \
    }                                                              \
  } while (0)
  NA_HIP(hipMemcpyAsync(frame, &df, sizeof df, hipMemcpyHostToDevice, s));
  NA_HIP(hipMemcpyAsync(tabs, f.cre, (size_t)f.nc * 8, hipMemcpyHostToDevice, s));
  NA_HIP(hipMemcpyAsync(tabs + f.nc, f.crelo, (size_t)f.nc * 8, hipMemcpyHostToDevice, s));
  NA_HIP(hipMemcpyAsync(tabs + 2 * f.nc, f.cim, (size_t)f.nr * 8, hipMemcpyHostToDevice, s));
  NA_HIP(hipMemcpyAsync(tabs + 2 * f.nc + f.nr, f.cimlo, (size_t)f.nr * 8, hipMemcpyHostToDevice, s));
  PlainArgs a{};
  Common& c = a.c;
  c.frames = frame;
  c.nframes = 1;
  c.frame_size = (uint32_t)n;
  c.inv_frame = 1.0 / (double)n;
  c.total = (uint32_t)n;
  c.out = out;
  c.r2 = r2;
  c.state = state;
  c.state_i = nullptr;  // the plain path keeps no per-sample count
  c.hand = hand;
  c.tile_count = tiles;
  c.active = active;
  c.nactive = nactive;
  c.part = part;
  c.N = N;
  c.b2 = bailout * bailout;
  c.b2m = c.b2 * (1.0 - 1.0 / 1099511627776.0);
  c.log_b = std::log(bailout);
  c.smooth_k = bailout >= 2.0 ? 0.5 * std::log(2.0) / c.log_b : 0.0;
  a.pre_iter = kPreIterDefault;
  a.k_start = (std::min(a.pre_iter, N) + 3) / 4 * 4;
  a.cre = tabs;
  a.crelo = tabs + f.nc;
  a.cim = tabs + 2 * f.nc;
  a.cimlo = tabs + 2 * f.nc + f.nr;
  const bool fma2 = f.fma_ok && bailout <= 0x1p400, grouped = grouped_test_ok(bailout, f.c_bound);
  NA_HIP(fma2 && grouped ? launch_pass_on(s, d->num_cu, part, nactive, a, k_plain_start<true, true>,
                                       k_plain_iterate<64, true, true>, kPlainWaves, nullptr, nullptr)
         : fma2       ? launch_pass_on(s, d->num_cu, part, nactive, a, k_plain_start<true, false>,
                                       k_plain_iterate<64, true, false>, kPlainWaves, nullptr, nullptr)
         : grouped    ? launch_pass_on(s, d->num_cu, part, nactive, a, k_plain_start<false, true>,
                                       k_plain_iterate<64, false, true>, kPlainWaves, nullptr, nullptr)
                      : launch_pass_on(s, d->num_cu, part, nactive, a, k_plain_start<false, false>,
                                       k_plain_iterate<64, false, false>, kPlainWaves, nullptr, nullptr));
  NA_HIP(hipMemcpyAsync(host_out, out, n * sizeof(EscapeValue), hipMemcpyDeviceToHost, s));
  NA_HIP(hipStreamSynchronize(s));
#undef NA_HIP
  return NM_OK;
}

}  // namespace nm
