// mandelbrot.hpp - drop-in for newman's render core (mandelbrot.h:7-53).
//
// Same public surface: error_tolerance, N, center, sz; rows/cols, useHardware,
// precompute, computeRow, pointAt, translate, zoom, zoomAt, at(r,c), at(r,c,sc),
// scaleUp, scaleDown, loadLegacy, load, save.  The reference orbit and series are
// computed serially on the host in HP (findProbe is threaded over probes, with
// the same "longest orbit, first on ties" choice) and uploaded once; computeRow
// renders the whole frame on the GPU on its first call for a view and then
// serves rows from the host copy of the grid.
#pragma once
#include <string>
#include <vector>

#include "engine.hpp"
#include "hpfloat.hpp"
#include "skipnet.hpp"

namespace nm {

// grid.h:6-25
class RenderGrid {
 public:
  using EscapeValue = nm::EscapeValue;
  int nr = 0, nc = 0;
  std::vector<EscapeValue> values;
  RenderGrid() = default;
  RenderGrid(int nr_, int nc_) : nr(nr_), nc(nc_), values((size_t)nr_ * nc_, EscapeValue{0, 0.0f}) {}
  EscapeValue& at(int r, int c) { return values[(size_t)r * nc + c]; }
  const EscapeValue& at(int r, int c) const { return values[(size_t)r * nc + c]; }
};

class Mandelbrot {
 public:
  double error_tolerance;  // mandelbrot.h:23
  int N;                   // mandelbrot.h:24
  HPComplex center, sz;    // mandelbrot.h:25

  // Extensions of this engine (not in the reference): bailout radius (the
  // reference's constant 1024, mandelbrot.cpp:58) and the delta arithmetic.
  double bailout = 1024.0;
  int delta_mode = DELTA_FP64;
  // the series test and skip in the reference's doubles or in floatexp
  // (SeriesMode); setSeriesMode rebuilds the tables of a precomputed reference
  int series_mode = SERIES_DOUBLE;
  int setSeriesMode(int mode);
  // Launch settings for later renders (0 = the engine's defaults): iterations
  // per block, persistent waves per SIMD, sample order (0 row-major, T = T x T
  // tiles) and HIP-graph replay of repeated passes.  autotuneFrames sets them.
  struct Launch {
    int block_k = 0, waves_per_simd = 0, tile = 0, graph = 0, pre_iter = 0, chunk_run = 0;
    double ms = 0.0;
  };
  Launch launch;

  // Series skip granularity on the series path: per sample (the reference's
  // per-pixel scan, default) or one skip per T x T tile taken at the tile's probe
  // (its sample farthest from the reference point), found by the reference's scan
  // (SKIP_TILE_SCAN), the threshold tables (SKIP_TILE_TABLE) or predicted by a
  // SkipNet and validated, with the scan as fallback (SKIP_TILE_LEARNED).
  // SKIP_NONE starts every sample at index 0 (no series: plain perturbation from
  // delta_0 = e), for views where the reference's double-precision series test
  // degenerates and for precision studies of the delta arithmetic.
  enum SkipMode { SKIP_PIXEL = 0, SKIP_TILE_SCAN = 1, SKIP_TILE_TABLE = 2, SKIP_TILE_LEARNED = 3, SKIP_NONE = 4 };
  int skip_mode = SKIP_PIXEL;
  int skip_tile = 64;
  SkipNet* skipnet = nullptr;  // not owned
  // findProbe on the host for every probe (PROBE_HOST), or screened on the GPU
  // and verified in HP for the candidates (PROBE_GPU, the default with a device)
  enum ProbeMode { PROBE_HOST = 0, PROBE_GPU = 1 };
  int probe_mode = PROBE_GPU;
  int probe_candidates_ = -1;  // probes iterated in HP by the last GPU screening
  uint64_t dev_tag_ = 0;       // engine render tag of this object's last render
  int dev_frames_ = 0;         // frames in that render
  bool dev_grid_ = false;      // that render was of this object's grid (not a frame set)
  struct TileStats {
    int ntiles = 0, accepted = 0, rejected = 0, bounded = 0;
    uint64_t evals = 0;         // exact stability tests evaluated at the probes
    double ms_search = 0.0;     // the tile kernel
    double ms_predict = 0.0;    // features + network + snapping (host wall time)
  };
  TileStats tile_stats;

  Mandelbrot();
  Mandelbrot(int nr, int nc);

  int loadLegacy(const char* fn);  // mandelbrot.cpp:19-35 (returns a status)
  int rows() const { return grid.nr; }
  int cols() const { return grid.nc; }

  bool useHardware();
  int precompute();
  int computeRow(int r);
  int computeAll();  // every row in one launch

  HPComplex pointAt(int r, int c, int sc = 1) const;
  void translate(int dr, int dc, int sc = 1);
  void zoom(float scale);
  void zoomAt(float scale, int r, int c, int sc = 1);

  const EscapeValue& at(int r, int c);
  EscapeValue at(int r, int c, int sc);

  int scaleUp(int sc);
  int scaleDown(int sc);

  void load(const char* fn) { (void)fn; }  // mandelbrot.cpp:362 (stub)
  void save(const char* fn) { (void)fn; }  // mandelbrot.cpp:364 (stub)

  // ---- extensions used by the C ABI, tests and benchmarks
  void setPrecision();                          // mandelbrot.cpp:37-56
  long precisionBits() const { return bits_; }
  int setView(int N, const char* cre, const char* cim, const char* width);
  int precomputeAt(const char* re, const char* im);  // explicit reference point
  int referenceLength() const { return (int)ref_.len; }
  int probeWinner() const { return winner_; }
  // the viewer's colouring (viewer.cpp:89-122) of this object's grid (frame < 0)
  // or of frame `frame` of its last renderFrames; from the device copy of its
  // last render while that is still the engine's, else from the host grid
  int colour(int frame, const uint8_t* pal_rgb, int npal, int smooth, int sc, uint8_t* out, double* ms);
  // the viewer's zoom video: initAutoZoom (viewer.cpp:441-453) and one keyframe
  // of update()'s zoom branch (viewer.cpp:270-286) with VideoZoom::nextFrame's
  // in-between frames (video.cpp:14-34)
  int zoomVideoStart(int sc);
  // beautyRender (viewer.cpp:186-253): this view on a 3240 x 5760 grid, coloured
  // 3 x 3-averaged into 1080 x 1920 x 3 bytes
  int beautyRender(const uint8_t* pal_rgb, int npal, int smooth, uint8_t* out, double* ms) const;
  int zoomVideoStep(const uint8_t* pal_rgb, int npal, int smooth, int rate, const uint8_t* prev_key,
                    uint8_t* key_out, uint8_t* frames_out, int* nframes, double* ms);
  int probeCandidates() const { return probe_candidates_; }
  bool hasReference() const { return have_ref_; }
  const HPComplex& referencePoint() const { return X[0]; }
  const SaReferenceHost& reference() const { return ref_; }
  RenderGrid& grid_ref() { return grid; }
  const RenderStats& lastStats() const { return stats_; }
  int render(RenderOpts* opts = nullptr, bool copy_back = true);
  // frames with this object's centre, reference and grid size, one per width
  // (width = cols * sz); outputs concatenated
  int renderFrames(const std::vector<std::string>& widths, RenderOpts* opts, EscapeValue* out);
  // Sweeps block size, waves per SIMD and sample order over (a subsample of) the
  // frame set, keeps the fastest in `launch` and returns the table.
  int autotuneFrames(const std::vector<std::string>& widths, std::vector<Launch>* table);
  struct FrameSet {
    bool hw = false;
    std::vector<HwFrameHost> h;
    std::vector<SaFrameHost> s;
  };
  int buildFrames(const std::vector<std::string>& widths, FrameSet& set);
  // frames first, first + step, ... of a built set, one launch set
  int renderSet(const FrameSet& set, const RenderOpts& o, EscapeValue* out, size_t first, size_t step);
  int renderSetOn(Engine* eng, const FrameSet& set, const RenderOpts& o, EscapeValue* out, size_t first,
                  size_t step);
  static uint64_t next_serial();
  // per-tile features of the learned skip predictor for this view (kIn floats per
  // tile, tiles row-major): log10 width; the probe's offset from the reference
  // point in widths (re, im); log2 |a_k(i)| of the 8-term Taylor series of the
  // delta in epsilon (k = 1..8) at 16 orbit indices; log10 N
  int tileFeatures(int T, std::vector<float>& feats, const SaFrameHost* frame = nullptr);
  std::vector<int> pred_;  // the last snapped predictions and their key
  std::string pred_key_;
  // per-tile first unstable index at the probes; mode 0 scan, 1 tables, 2 learned
  int tileSkips(int T, int mode, SkipNet* net, std::vector<int>& ua, std::vector<int>* how);
  // the learned network's target for a first unstable index, and back
  void ensureRecords();
  double targetOf(int ua) const;
  int snapIndex(double frac) const;
  int tileSkipsFor(Engine* eng, const std::vector<const SaFrameHost*>& frames, const RenderOpts& o,
                   std::vector<int>& ua, Engine::TileSearch* out = nullptr);
  std::vector<double> taylor_feats_;  // 128 per reference, cached by serial
  uint64_t taylor_serial_ = 0;
  std::vector<int> records_;          // indices where the threshold prefix minimum drops
  uint64_t records_serial_ = 0;

  // the last frame set built, its key (view snapshot + widths) and serial
  FrameSet cache_;
  std::string cache_key_;
  uint64_t cache_serial_ = 0;
  void applyLaunch(RenderOpts& o) const;
  int buildHwFrame(HwFrameHost& f) { return buildHwFrame(f, center, sz); }
  int buildSaFrame(SaFrameHost& f) { return buildSaFrame(f, center, sz); }
  int buildHwFrame(HwFrameHost& f, const HPComplex& c, const HPComplex& s);
  void fillHwTables(const HPComplex& c, const HPComplex& s, double* cre, double* crelo, double* cim, double* cimlo,
                    int* card_mode, bool* fma_ok, double* c_bound) const;
  // computeAll into out (rows x cols) on the plain path without allocating
  // device memory (render_hw_noalloc); noalloc_err_ holds the last error
  int computeAllNoAlloc(EscapeValue* out, size_t n);
  std::string noalloc_err_;
  int buildSaFrame(SaFrameHost& f, const HPComplex& c, const HPComplex& s);
  void invalidate() { rendered_ = false; }

 protected:
  RenderGrid grid;
  std::vector<HPComplex> X;  // reference orbit, plus one extra element X[len]
  SaReferenceHost ref_;
  long bits_ = 64;
  int winner_ = -1;
  bool have_ref_ = false;
  bool rendered_ = false;
  uint64_t ref_serial_ = 0;  // identifies this object's uploaded reference
  std::string snap_;  // view snapshot of the last render
  RenderStats stats_;

  void findProbe();
  int findProbeGPU(Engine* eng);
  void fillOrbit(SaReferenceHost& r) const;
  int computeOrbit(const HPComplex& X0, bool store);
  void computeSeries();
  void buildThresholds();
  void buildThresholdsFx();
  void buildStopLookup();
  std::string snapshot() const;
  int uploadReference();
};

// per-pixel helpers shared by the frame builders and tests
bool inCardioidHP(const hpf& re, const hpf& im);  // mandelbrot.cpp:63-71

}  // namespace nm
