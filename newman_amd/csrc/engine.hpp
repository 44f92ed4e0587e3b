// engine.hpp - host-side interface of the HIP render engine (no HIP types here,
// so the host classes compile with any C++ compiler).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/newman_amd.h"

namespace nm {

// grid.h:8-15: {int iterations; float smoothing}, 8 bytes, row-major AoS.
struct EscapeValue {
  int32_t iterations;
  float smoothing;
};
static_assert(sizeof(EscapeValue) == 8, "EscapeValue must keep grid.h's 8-byte layout");

// Status codes (NM_OK, NM_ERR_*) are the C ABI's.
// Cardioid/bulb pre-classification of a frame.
enum CardMode : int32_t { CARD_NONE = 0, CARD_ALL = 1, CARD_TEST_DD = 2, CARD_MASK = 3 };

// Arithmetic used for the deltas of the perturbation kernel.
enum DeltaMode : int32_t { DELTA_FP64 = 0, DELTA_FP32X = 1 };
// the series test and skip in the reference's doubles, or in floatexp (double
// mantissa + int exponent: the same results where the doubles are normal
// numbers, and a working test where they under- or overflow)
enum SeriesMode : int32_t { SERIES_DOUBLE = 0, SERIES_FLOATEXP = 1 };
constexpr int32_t kFxZeroExp = -0x40000000;  // the exponent of a floatexp zero

// One frame of the plain-double path (mandelbrot.cpp:231-254): per-column and
// per-row truncated coordinates (descend of the HP pixel map) plus the remainder
// as a double-double low part for the cardioid test.
struct HwFrameHost {
  int nr = 0, nc = 0;
  int card_mode = CARD_TEST_DD;
  // every coordinate is 0 or at least 2^-200 in magnitude: the device may fold
  // (2x) y + c into fma(x y, 2, c) bit for bit (see plain_update)
  bool fma_ok = false;
  double c_bound = HUGE_VAL;        // at least every |c| of the frame (grouped_test_ok; unknown: +inf)
  std::vector<double> cre, crelo;  // nc
  std::vector<double> cim, cimlo;  // nr
};

// One frame of the series-approximation path (mandelbrot.cpp:144-229): per-column
// and per-row epsilon = descend(Y0 - X[0]), and the cardioid classification.
struct SaFrameHost {
  int nr = 0, nc = 0;
  int card_mode = CARD_NONE;
  std::vector<double> epsre;      // nc
  std::vector<double> epsim;      // nr
  std::vector<uint32_t> mask;     // CARD_MASK: nr*nc bits, 1 = in cardioid/bulb
};

// Reference orbit and series coefficients in device-ready form.
struct SaReferenceHost {
  int len = 0;                 // len(X)
  std::vector<double> Zd;      // 2*(len+2): Zd[0] = 0, Zd[k] = nearest(X[k-1]); Zd[len+1] = X[len]
  std::vector<double> Xlo;     // 2*len: nearest(X[k] - Zd[k+1])
  std::vector<double> A, B, C; // 2*len: descend(A[i]) etc. (mandelbrot.cpp:166-168)
  // The same coefficients in floatexp: per index and component the truncated
  // 53-bit mantissa in [1/2, 1) (0 for zero) and its exponent (kFxZeroExp for zero)
  std::vector<double> Am, Bm, Cm;    // 2*len
  std::vector<int32_t> Ae, Be, Ce;   // 2*len
  int fxs = 0;  // 1: the tables below come from the floatexp form, and the kernels test in it
  // Stop-index tables (prefix over j in [1, i]):
  std::vector<double> PM;      // min log2(tol |B_j|^2 / |C_j|^2)
  std::vector<double> BX, BN;  // max / min log2 |B_j|^2
  std::vector<double> CX, CN;  // max / min log2 |C_j|^2
  std::vector<int32_t> SP;     // count of special j (zero or non-finite B_j or C_j)
  // For views whose series terms under/overflow in the exact test (widths below
  // about 1e-90): the thresholds T_j themselves, their sparse table of range
  // minima (level k at [k * len], minima over [i, i + 2^k)), and suffix minima of
  // log2 |B_j|^2 and log2 |C_j|^2 (over j >= i).
  std::vector<double> T, ST, BNs, CNs;
  int st_levels = 0;
  // Lookup of the stop search: PMT[k] = first i >= 1 with PM[i] <= pmt_base + k h
  // (len if none), k = 0 .. pmt_n, h = 1 / pmt_inv_h, over the finite range of
  // PM; pmt_n = 0 when PM has fewer than two distinct finite values.
  std::vector<int32_t> PMT;
  double pmt_base = 0.0, pmt_inv_h = 0.0;
  int pmt_n = 0;
};

// The series iterate kernel's active-list chunk setting (RenderOpts::chunk_run;
// 0 or 64: 64 consecutive entries per chunk; 1, 2, 4, 8, 16, 32: interleaved
// runs of that many entries, chunk_pos_run) as the run's log2, or -1 when
// invalid.
inline int chunk_run_shift(int run) {
  switch (run) {
    case 0: case 64: return 6;
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    case 32: return 5;
    default: return -1;
  }
}

struct RenderOpts {
  int N = 256;
  double bailout = 1024.0;
  double tol = 1e-10;
  int delta_mode = DELTA_FP64;
  int block_k = 0;        // iterations per block between lane checks (0 = tuned default)
  int waves_per_simd = 0; // persistent waves per SIMD (0 = tuned default)
  int repeats = 1;        // timing: launch the whole frame set this many times
  int tile = 0;           // sample order: 0 row-major, T = T x T tiles of 8 x 8 blocks (8..256)
  bool graph = false;     // replay repeats 2.. from one captured HIP graph
  int pre_iter = 0;       // plain path: updates in the start kernel (0 = default)
  int chunk_run = 0;      // series path: active-list chunks of 64 consecutive samples (64)
                          // or interleaved runs of 1, 2, 4, 8, 16, 32 (chunk_run_shift);
                          // 0: runs of 4 for deltas on a floatexp series, else 64
  const int* tile_ua = nullptr;  // series path: per-tile first unstable index (null: per sample)
  int skip_T = 0;                // its tile side
  uint64_t set_serial = 0; // identifies the frame set: the engine keeps its tables
                           // on the device and skips the upload when it repeats (0 = always upload)
};

struct RenderStats {
  uint64_t samples = 0;
  uint64_t n_cardioid = 0;   // samples resolved by the cardioid/bulb test
  uint64_t n_sa_resolved = 0; // escaped inside the series skip, or the skip reached N
  uint64_t n_iterated = 0;   // samples handed to the iteration kernel
  uint64_t executed = 0;     // iteration steps executed (start and iteration kernels)
  uint64_t executed_start = 0; // of executed: steps run inside the start kernel
  uint64_t skipped = 0;      // sum of series-skip start indices of iterated samples
  uint64_t n_rebased = 0;    // blocks of block_k steps in which a sample rebased (glitch avoidance)
  uint64_t credited = 0;     // sum of reported iterations over non-cardioid samples
  double ms_start = 0;       // classification / series start kernel
  double ms_compact = 0;     // tile scan + scatter
  double ms_iterate = 0;     // iteration kernel
  double ms_finalize = 0;    // smoothing / output kernel
  double ms_total = 0;       // all of one pass; every time is a mean over repeats
  int graph = 0;             // 1: ms_total is the mean of HIP graph replays
  int tile = 0;              // the sample order used
  int fma2 = 0;              // plain path: the folded update ran (plain_update)
  int grouped = 0;           // plain path: updates were tested in groups of four (grouped_test_ok)
  int chunk_run = 0;         // the iterate kernel's chunk order used: 64 consecutive samples, or runs of R
  int block_k = 0, waves_per_simd = 0;
};

// A plain-path frame as tables the caller keeps (render_hw_noalloc).
struct HwFrameView {
  int nr = 0, nc = 0;
  int card_mode = CARD_TEST_DD;
  bool fma_ok = false;
  double c_bound = HUGE_VAL;
  const double *cre = nullptr, *crelo = nullptr;  // nc each
  const double *cim = nullptr, *cimlo = nullptr;  // nr each
};

// The plain path of one frame (at most 2^19 samples, rows and columns at most
// 8192) with every device buffer in a file-scope device array: no device
// allocation, no stream or event created (the null stream), no engine buffer
// touched.  Results in host_out (nr x nc, row-major).  Calls overlapping on one
// device share that array and race.
int render_hw_noalloc(const HwFrameView& f, int N, double bailout, EscapeValue* host_out, std::string& err);

// Device engine: one per process per device.  All calls are synchronous.
class Engine {
 public:
  static int instance(Engine** out);
  virtual ~Engine() = default;
  // Make `ref` the device reference; `serial` identifies it so an unchanged
  // reference is not uploaded again.
  virtual int set_reference(const SaReferenceHost& ref, uint64_t serial) = 0;
  virtual int render_hw(const std::vector<const HwFrameHost*>& frames, const RenderOpts& opt,
                        EscapeValue* host_out, RenderStats* stats) = 0;
  virtual int render_sa(const std::vector<const SaFrameHost*>& frames, const RenderOpts& opt,
                        EscapeValue* host_out, RenderStats* stats) = 0;
  // per-tile series skip at each tile's probe (see k_tile_skip)
  struct TileSearch {
    int T = 64;             // tile side
    int mode = 0;           // 0 the reference's scan, 1 threshold tables, 2 validate pred
    std::vector<int> pred;  // mode 2: one per tile (frame-major, row-major tiles)
    std::vector<int> ua, how, evals;
    double ms = 0.0;
  };
  virtual int tile_skip(const std::vector<const SaFrameHost*>& frames, const RenderOpts& opt, TileSearch& ts) = 0;
  // multisample average (mandelbrot.cpp:320-332) of a fine grid on the device
  virtual int downsample(const EscapeValue* fine, int nr, int nc, int sc, EscapeValue* out) = 0;
  // Counts renders: the result of the render that returned tag t stays on the
  // device (row-major grids, frame after frame) while render_tag() == t.
  virtual uint64_t render_tag() const = 0;
  // The viewer's colouring (viewer.cpp:89-122) into (nr / sc) x (nc / sc) x 3
  // interleaved RGB bytes: of host_grid (uploaded) when it is given, else of the
  // device result of render `tag` from escape value `offset` on.  pal_rgb: npal
  // entries of 3 bytes (npal >= N).  *ms: the colouring kernel's time.
  virtual int colour(const EscapeValue* host_grid, uint64_t tag, size_t offset, int nr, int nc, int sc, int N,
                     const uint8_t* pal_rgb, int npal, int smooth, uint8_t* host_out, double* ms) = 0;
  // The zoom video's `rate` in-between frames (video.cpp:14-34) from keyframes
  // prev and next (kr x kc x 3 bytes, host) into out (rate x nr x nc x 3, host).
  virtual int video(const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc, int rate,
                    uint8_t* out, double* ms) = 0;
  virtual const char* last_error() const = 0;
};

}  // namespace nm
