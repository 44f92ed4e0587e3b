// capi.cpp - extern "C" boundary (include/newman_amd.h) over nm::Mandelbrot.
// No exception crosses this boundary; every entry point returns a status.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/newman_amd.h"
#include "mandelbrot.hpp"
#include "palette.hpp"

using nm::Mandelbrot;

struct nm_mandelbrot {
  Mandelbrot m;
  nm_mandelbrot(int nr, int nc) : m(nr, nc) {}
  explicit nm_mandelbrot(const Mandelbrot& o) : m(o) {}
};

static thread_local std::string g_err;

static int put_str(const std::string& s, char* out, size_t cap) {
  if (!out || cap == 0) return NM_ERR_ARG;
  if (s.size() + 1 > cap) return NM_ERR_RANGE;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return NM_OK;
}

// Every call that reaches the device engine runs under this process-wide lock:
// the engine (one per process) keeps one stream and grow-only buffers, so two
// threads rendering through different handles would otherwise overwrite each
// other's tables and results, or free a buffer the other's kernels still use.
static std::recursive_mutex g_call_mu;

template <class F>
static int guard(F&& f) {
  std::lock_guard<std::recursive_mutex> lk(g_call_mu);
  try {
    return f();
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return NM_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NM_ERR_ARG;
  } catch (...) {
    g_err = "unknown error";
    return NM_ERR_ARG;
  }
}

static int engine_status(int st) {
  if (st != NM_OK) {
    nm::Engine* e = nullptr;
    if (nm::Engine::instance(&e) == NM_OK && e) g_err = e->last_error();
  }
  return st;
}

static void copy_stats(const nm::RenderStats& s, nm_render_stats* o) {
  if (!o) return;
  o->samples = s.samples;
  o->n_cardioid = s.n_cardioid;
  o->n_sa_resolved = s.n_sa_resolved;
  o->n_iterated = s.n_iterated;
  o->executed = s.executed;
  o->skipped = s.skipped;
  o->n_rebased = s.n_rebased;
  o->credited = s.credited;
  o->ms_start = s.ms_start;
  o->ms_iterate = s.ms_iterate;
  o->ms_total = s.ms_total;
  o->block_k = s.block_k;
  o->waves_per_simd = s.waves_per_simd;
  o->executed_start = s.executed_start;
  o->ms_compact = s.ms_compact;
  o->ms_finalize = s.ms_finalize;
  o->graph = s.graph;
  o->tile = s.tile;
  o->fma2 = s.fma2;
  o->grouped = s.grouped;
  o->chunk_run = s.chunk_run;
}

extern "C" {

int nm_create(nm_mandelbrot** out, int nr, int nc) {
  if (!out || nr <= 0 || nc <= 0) return NM_ERR_ARG;
  return guard([&] {
    *out = new nm_mandelbrot(nr, nc);
    return NM_OK;
  });
}

int nm_clone(nm_mandelbrot** out, const nm_mandelbrot* src) {
  if (!out || !src) return NM_ERR_ARG;
  return guard([&] {
    *out = new nm_mandelbrot(src->m);
    return NM_OK;
  });
}

void nm_destroy(nm_mandelbrot* h) { delete h; }

int nm_rows(const nm_mandelbrot* h) { return h ? h->m.rows() : NM_ERR_ARG; }
int nm_cols(const nm_mandelbrot* h) { return h ? h->m.cols() : NM_ERR_ARG; }

int nm_set_N(nm_mandelbrot* h, int N) {
  if (!h || N <= 0) return NM_ERR_ARG;
  h->m.N = N;
  return NM_OK;
}
int nm_get_N(const nm_mandelbrot* h) { return h ? h->m.N : NM_ERR_ARG; }

int nm_set_error_tolerance(nm_mandelbrot* h, double tol) {
  if (!h || !(tol > 0)) return NM_ERR_ARG;
  h->m.error_tolerance = tol;
  return NM_OK;
}
double nm_get_error_tolerance(const nm_mandelbrot* h) { return h ? h->m.error_tolerance : -1.0; }

int nm_set_center(nm_mandelbrot* h, const char* re, const char* im) {
  if (!h || !re || !im) return NM_ERR_ARG;
  return guard([&] {
    nm::hpf r(h->m.center.re.L), i(h->m.center.im.L);
    if (!nm::hp_set_str(r, re) || !nm::hp_set_str(i, im)) return (int)NM_ERR_ARG;
    h->m.center.re = r;
    h->m.center.im = i;
    return (int)NM_OK;
  });
}

int nm_get_center(const nm_mandelbrot* h, char* re, char* im, size_t cap) {
  if (!h) return NM_ERR_ARG;
  int a = put_str(h->m.center.re.to_string(), re, cap);
  return a ? a : put_str(h->m.center.im.to_string(), im, cap);
}

int nm_set_size(nm_mandelbrot* h, const char* re, const char* im) {
  if (!h || !re || !im) return NM_ERR_ARG;
  return guard([&] {
    nm::hpf r(h->m.sz.re.L), i(h->m.sz.im.L);
    if (!nm::hp_set_str(r, re) || !nm::hp_set_str(i, im)) return (int)NM_ERR_ARG;
    h->m.sz.re = r;
    h->m.sz.im = i;
    return (int)NM_OK;
  });
}

int nm_get_size(const nm_mandelbrot* h, char* re, char* im, size_t cap) {
  if (!h) return NM_ERR_ARG;
  int a = put_str(h->m.sz.re.to_string(), re, cap);
  return a ? a : put_str(h->m.sz.im.to_string(), im, cap);
}

int nm_get_size_d(const nm_mandelbrot* h, double* re, double* im) {
  if (!h || !re || !im) return NM_ERR_ARG;
  *re = h->m.sz.re.get_d();
  *im = h->m.sz.im.get_d();
  return NM_OK;
}

int nm_set_precision(nm_mandelbrot* h) {
  if (!h) return NM_ERR_ARG;
  h->m.setPrecision();
  return NM_OK;
}
long nm_precision_bits(const nm_mandelbrot* h) { return h ? h->m.precisionBits() : NM_ERR_ARG; }

int nm_use_hardware(nm_mandelbrot* h) { return h ? (h->m.useHardware() ? 1 : 0) : NM_ERR_ARG; }

int nm_precompute(nm_mandelbrot* h) {
  if (!h) return NM_ERR_ARG;
  return guard([&] { return engine_status(h->m.precompute()); });
}

int nm_compute_row(nm_mandelbrot* h, int r) {
  if (!h) return NM_ERR_ARG;
  return guard([&] { return engine_status(h->m.computeRow(r)); });
}

int nm_compute_all(nm_mandelbrot* h) {
  if (!h) return NM_ERR_ARG;
  return guard([&] { return engine_status(h->m.computeAll()); });
}

int nm_compute_all_noalloc(nm_mandelbrot* h, nm_escape_value* out, size_t n) {
  if (!h || !out) return NM_ERR_ARG;
  return guard([&] {
    int st = h->m.computeAllNoAlloc((nm::EscapeValue*)out, n);
    if (st != NM_OK) g_err = h->m.noalloc_err_.empty() ? "nm_compute_all_noalloc: argument, view or size" : h->m.noalloc_err_;
    return st;
  });
}

int nm_at(nm_mandelbrot* h, int r, int c, nm_escape_value* out) {
  if (!h || !out || r < 0 || c < 0 || r >= h->m.rows() || c >= h->m.cols()) return NM_ERR_ARG;
  const nm::EscapeValue& e = h->m.at(r, c);
  out->iterations = e.iterations;
  out->smoothing = e.smoothing;
  return NM_OK;
}

int nm_at_avg(nm_mandelbrot* h, int r, int c, int sc, nm_escape_value* out) {
  if (!h || !out || sc <= 0 || r < 0 || c < 0 || (r + 1) * sc > h->m.rows() || (c + 1) * sc > h->m.cols())
    return NM_ERR_ARG;
  nm::EscapeValue e = h->m.at(r, c, sc);
  out->iterations = e.iterations;
  out->smoothing = e.smoothing;
  return NM_OK;
}

int nm_get_grid(nm_mandelbrot* h, nm_escape_value* out, size_t n) {
  if (!h || !out) return NM_ERR_ARG;
  auto& v = h->m.grid_ref().values;
  if (n < v.size()) return NM_ERR_RANGE;
  std::memcpy(out, v.data(), v.size() * sizeof(nm_escape_value));
  return NM_OK;
}

int nm_point_at(const nm_mandelbrot* h, int r, int c, int sc, char* re, char* im, size_t cap) {
  if (!h) return NM_ERR_ARG;
  nm::HPComplex p = h->m.pointAt(r, c, sc);
  int a = put_str(p.re.to_string(), re, cap);
  return a ? a : put_str(p.im.to_string(), im, cap);
}

int nm_translate(nm_mandelbrot* h, int dr, int dc, int sc) {
  if (!h) return NM_ERR_ARG;
  h->m.translate(dr, dc, sc);
  return NM_OK;
}

int nm_zoom(nm_mandelbrot* h, float scale) {
  if (!h || !(scale > 0)) return NM_ERR_ARG;
  h->m.zoom(scale);
  return NM_OK;
}

int nm_zoom_at(nm_mandelbrot* h, float scale, int r, int c, int sc) {
  if (!h || !(scale > 0)) return NM_ERR_ARG;
  h->m.zoomAt(scale, r, c, sc);
  return NM_OK;
}

int nm_scale_up(nm_mandelbrot* h, int sc) {
  if (!h || sc <= 0) return NM_ERR_ARG;
  return guard([&] { return engine_status(h->m.scaleUp(sc)); });
}

int nm_scale_down(nm_mandelbrot* h, int sc) {
  if (!h || sc <= 0) return NM_ERR_ARG;
  return guard([&] { return engine_status(h->m.scaleDown(sc)); });
}

int nm_load_legacy(nm_mandelbrot* h, const char* path) {
  if (!h || !path) return NM_ERR_ARG;
  return guard([&] { return h->m.loadLegacy(path); });
}

// viewer.cpp:6-29: N, then centre.re, centre.im, sz.re, sz.im, one per line
int nm_save_view(const nm_mandelbrot* h, const char* path) {
  if (!h || !path) return NM_ERR_ARG;
  FILE* fp = std::fopen(path, "w");
  if (!fp) return NM_ERR_IO;
  std::fprintf(fp, "%d\n%s\n%s\n%s\n%s\n", h->m.N, h->m.center.re.to_string().c_str(),
               h->m.center.im.to_string().c_str(), h->m.sz.re.to_string().c_str(), h->m.sz.im.to_string().c_str());
  std::fclose(fp);
  return NM_OK;
}

int nm_set_bailout(nm_mandelbrot* h, double bailout) {
  if (!h || !(bailout > 1.0)) return NM_ERR_ARG;
  h->m.bailout = bailout;
  return NM_OK;
}

int nm_set_delta_mode(nm_mandelbrot* h, int mode) {
  if (!h || (mode != NM_DELTA_FP64 && mode != NM_DELTA_FP32X)) return NM_ERR_ARG;
  h->m.delta_mode = mode;
  return NM_OK;
}

int nm_set_series_mode(nm_mandelbrot* h, int mode) {
  if (!h || (mode != NM_SERIES_DOUBLE && mode != NM_SERIES_FLOATEXP)) return NM_ERR_ARG;
  return guard([&] { return h->m.setSeriesMode(mode); });
}

int nm_set_view(nm_mandelbrot* h, int N, const char* cre, const char* cim, const char* width) {
  if (!h) return NM_ERR_ARG;
  return guard([&] { return h->m.setView(N, cre, cim, width); });
}

int nm_precompute_at(nm_mandelbrot* h, const char* re, const char* im) {
  if (!h || !re || !im) return NM_ERR_ARG;
  return guard([&] { return engine_status(h->m.precomputeAt(re, im)); });
}

int nm_reference_length(const nm_mandelbrot* h) { return h ? h->m.referenceLength() : NM_ERR_ARG; }

int nm_reference_point(const nm_mandelbrot* h, char* re, char* im, size_t cap) {
  if (!h || !re || !im || cap < 2) return NM_ERR_ARG;
  return guard([&] {
    if (!h->m.hasReference()) return (int)NM_ERR_STATE;
    const nm::HPComplex& x0 = h->m.referencePoint();
    std::string a = x0.re.to_string(), b = x0.im.to_string();
    if (a.size() + 1 > cap || b.size() + 1 > cap) return (int)NM_ERR_RANGE;
    std::memcpy(re, a.c_str(), a.size() + 1);
    std::memcpy(im, b.c_str(), b.size() + 1);
    return (int)NM_OK;
  });
}
int nm_probe_winner(const nm_mandelbrot* h) { return h ? h->m.probeWinner() : NM_ERR_ARG; }
int nm_set_probe_mode(nm_mandelbrot* h, int mode) {
  if (!h || (mode != NM_PROBE_HOST && mode != NM_PROBE_GPU)) return NM_ERR_ARG;
  h->m.probe_mode = mode;
  return NM_OK;
}
int nm_probe_candidates(const nm_mandelbrot* h) { return h ? h->m.probeCandidates() : NM_ERR_ARG; }

int nm_get_reference(const nm_mandelbrot* h, double* X, double* A, double* B, double* C) {
  if (!h) return NM_ERR_ARG;
  const nm::SaReferenceHost& r = h->m.reference();
  if (r.len <= 0) return NM_ERR_STATE;
  size_t n = 2 * (size_t)r.len;
  if (X) std::memcpy(X, r.Zd.data() + 2, n * sizeof(double));
  if (A) std::memcpy(A, r.A.data(), n * sizeof(double));
  if (B) std::memcpy(B, r.B.data(), n * sizeof(double));
  if (C) std::memcpy(C, r.C.data(), n * sizeof(double));
  return NM_OK;
}

int nm_render(nm_mandelbrot* h, int block_k, int waves_per_simd, int repeats, int copy_back,
              nm_render_stats* stats) {
  if (!h || repeats < 1 || block_k < 0 || waves_per_simd < 0) return NM_ERR_ARG;
  return guard([&] {
    nm::RenderOpts o;
    o.block_k = block_k;
    o.waves_per_simd = waves_per_simd;
    o.repeats = repeats;
    int st = h->m.render(&o, copy_back != 0);
    if (st == NM_OK) copy_stats(h->m.lastStats(), stats);
    return engine_status(st);
  });
}

int nm_render_frames(nm_mandelbrot* h, const char* const* widths, int nframes, int block_k, int waves_per_simd,
                     int repeats, nm_escape_value* out, nm_render_stats* stats) {
  if (!h || !widths || nframes <= 0 || repeats < 1) return NM_ERR_ARG;
  return guard([&] {
    nm::RenderOpts o;
    o.block_k = block_k;
    o.waves_per_simd = waves_per_simd;
    o.repeats = repeats;
    std::vector<std::string> w(widths, widths + nframes);
    int st = h->m.renderFrames(w, &o, (nm::EscapeValue*)out);
    if (st == NM_OK) copy_stats(h->m.lastStats(), stats);
    return engine_status(st);
  });
}

int nm_set_launch(nm_mandelbrot* h, const nm_launch* l) {
  if (!h || !l || l->block_k < 0 || l->waves_per_simd < 0 || l->waves_per_simd > 16) return NM_ERR_ARG;
  if (l->pre_iter < 0 || l->pre_iter > 1024) return NM_ERR_ARG;
  const int t = l->tile;
  if (t != 0 && t != 8 && t != 16 && t != 32 && t != 64 && t != 128 && t != 256) return NM_ERR_ARG;
  if (nm::chunk_run_shift(l->chunk_run) < 0) return NM_ERR_ARG;
  h->m.launch.block_k = l->block_k;
  h->m.launch.waves_per_simd = l->waves_per_simd;
  h->m.launch.tile = t;
  h->m.launch.graph = l->graph != 0;
  h->m.launch.pre_iter = l->pre_iter;
  h->m.launch.chunk_run = l->chunk_run;
  h->m.launch.ms = 0.0;
  return NM_OK;
}

int nm_get_launch(const nm_mandelbrot* h, nm_launch* l) {
  if (!h || !l) return NM_ERR_ARG;
  l->block_k = h->m.launch.block_k;
  l->waves_per_simd = h->m.launch.waves_per_simd;
  l->tile = h->m.launch.tile;
  l->graph = h->m.launch.graph;
  l->pre_iter = h->m.launch.pre_iter;
  l->chunk_run = h->m.launch.chunk_run;
  l->ms = h->m.launch.ms;
  return NM_OK;
}

int nm_autotune_frames(nm_mandelbrot* h, const char* const* widths, int nframes, nm_launch* best,
                       nm_launch* table, int cap, int* n_table) {
  if (!h || !widths || nframes <= 0 || cap < 0) return NM_ERR_ARG;
  return guard([&] {
    std::vector<std::string> w(widths, widths + nframes);
    std::vector<nm::Mandelbrot::Launch> rows;
    int st = h->m.autotuneFrames(w, &rows);
    if (st != NM_OK) return engine_status(st);
    if (best) nm_get_launch(h, best);
    int n = 0;
    for (const auto& r : rows) {
      if (table && n < cap) table[n] = nm_launch{r.block_k, r.waves_per_simd, r.tile, r.graph, r.pre_iter, r.chunk_run, r.ms};
      n++;
    }
    if (n_table) *n_table = std::min(n, cap);
    return (int)NM_OK;
  });
}

struct nm_skipnet {
  nm::SkipNet net;
};

int nm_skipnet_create(nm_skipnet** out) {
  if (!out) return NM_ERR_ARG;
  return guard([&] {
    *out = new nm_skipnet();
    return (int)NM_OK;
  });
}

void nm_skipnet_destroy(nm_skipnet* net) { delete net; }

int nm_skipnet_get(const nm_skipnet* net, float* params, float* mean, float* scale) {
  if (!net) return NM_ERR_ARG;
  if (params) std::copy(net->net.params.begin(), net->net.params.end(), params);
  if (mean) std::copy(net->net.mean.begin(), net->net.mean.end(), mean);
  if (scale) std::copy(net->net.scale.begin(), net->net.scale.end(), scale);
  return NM_OK;
}

int nm_skipnet_set(nm_skipnet* net, const float* params, const float* mean, const float* scale) {
  if (!net || !params) return NM_ERR_ARG;
  std::copy(params, params + nm::SkipNet::kParams, net->net.params.begin());
  if (mean) std::copy(mean, mean + nm::SkipNet::kIn, net->net.mean.begin());
  if (scale) std::copy(scale, scale + nm::SkipNet::kIn, net->net.scale.begin());
  net->net.trained = true;
  net->net.version++;
  return NM_OK;
}

static int skipnet_status(nm_skipnet* net, int st) {
  if (st == NM_ERR_HIP) g_err = net->net.last_error();
  return st;
}

int nm_skipnet_train(nm_skipnet* net, const float* x, const float* t, int n, int epochs, int batch, float lr,
                     uint64_t seed, float* curve) {
  if (!net || !x || !t || n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f)) return NM_ERR_ARG;
  return guard([&] {
    nm::SkipNet::TrainOpts o;
    o.epochs = epochs;
    o.batch = batch;
    o.lr = lr;
    o.seed = seed;
    std::vector<float> c;
    int st = net->net.train(x, t, (size_t)n, o, curve ? &c : nullptr);
    if (st == NM_OK && curve) std::copy(c.begin(), c.end(), curve);
    return skipnet_status(net, st);
  });
}

int nm_skipnet_predict(nm_skipnet* net, const float* x, int n, float* y) {
  if (!net || !x || !y || n <= 0) return NM_ERR_ARG;
  return guard([&] { return skipnet_status(net, net->net.predict(x, (size_t)n, y)); });
}

int nm_skipnet_gradient(nm_skipnet* net, const float* xn, const float* t, int n, float* grad, float* loss) {
  if (!net || !xn || !t || !grad || n <= 0) return NM_ERR_ARG;
  return guard([&] { return skipnet_status(net, net->net.gradient(xn, t, (size_t)n, grad, loss)); });
}

int nm_tile_features(nm_mandelbrot* h, int tile, float* x, int cap_tiles, int* ntiles) {
  if (!h || tile < 8 || cap_tiles < 0) return NM_ERR_ARG;
  return guard([&] {
    std::vector<float> f;
    int st = h->m.tileFeatures(tile, f);
    if (st != NM_OK) return engine_status(st);
    const int nt = (int)(f.size() / nm::SkipNet::kIn);
    if (ntiles) *ntiles = nt;
    if (x) std::copy(f.begin(), f.begin() + (size_t)std::min(nt, cap_tiles) * nm::SkipNet::kIn, x);
    return (int)NM_OK;
  });
}

static void copy_tile_stats(const nm::Mandelbrot::TileStats& s, nm_tile_stats* o) {
  if (!o) return;
  o->ntiles = s.ntiles;
  o->accepted = s.accepted;
  o->rejected = s.rejected;
  o->bounded = s.bounded;
  o->evals = s.evals;
  o->ms_search = s.ms_search;
  o->ms_predict = s.ms_predict;
}

int nm_tile_skips(nm_mandelbrot* h, int tile, int mode, nm_skipnet* net, int* ua, int* how, int cap_tiles,
                  nm_tile_stats* stats) {
  if (!h || tile < 8 || mode < 0 || mode > 2 || (mode == 2 && !net) || cap_tiles < 0) return NM_ERR_ARG;
  return guard([&] {
    std::vector<int> u, w;
    int st = h->m.tileSkips(tile, mode, net ? &net->net : nullptr, u, &w);
    if (st != NM_OK) return engine_status(st);
    const size_t n = std::min(u.size(), (size_t)cap_tiles);
    if (ua) std::copy(u.begin(), u.begin() + n, ua);
    if (how) std::copy(w.begin(), w.begin() + n, how);
    copy_tile_stats(h->m.tile_stats, stats);
    return (int)NM_OK;
  });
}

int nm_tile_targets(nm_mandelbrot* h, const int* ua, int n, float* t) {
  if (!h || !ua || !t || n < 0) return NM_ERR_ARG;
  return guard([&] {
    if (!h->m.hasReference()) return (int)NM_ERR_STATE;
    h->m.ensureRecords();
    for (int i = 0; i < n; i++) t[i] = (float)h->m.targetOf(ua[i]);
    return (int)NM_OK;
  });
}

int nm_set_skip_mode(nm_mandelbrot* h, int mode, int tile, nm_skipnet* net) {
  if (!h || mode < NM_SKIP_PIXEL || mode > NM_SKIP_NONE || tile < 8) return NM_ERR_ARG;
  if (mode == NM_SKIP_TILE_LEARNED && !net) return NM_ERR_ARG;
  h->m.skip_mode = mode;
  h->m.skip_tile = tile;
  h->m.skipnet = net ? &net->net : nullptr;
  return NM_OK;
}

int nm_get_tile_stats(const nm_mandelbrot* h, nm_tile_stats* stats) {
  if (!h || !stats) return NM_ERR_ARG;
  copy_tile_stats(h->m.tile_stats, stats);
  return NM_OK;
}

const char* nm_last_error(void) { return g_err.c_str(); }

int nm_hp_op(int op, const char* a, const char* b, long iv, double dv, long bits, int* sign, long* exp,
             uint32_t* words) {
  if (!sign || !exp || !words) return NM_ERR_ARG;
  return guard([&] {
    int L = nm::hp_limbs_for_bits(bits);
    nm::hpf x(L), y(L), r(L);
    if (a && !nm::hp_set_str(x, a)) return (int)NM_ERR_ARG;
    if (b && !nm::hp_set_str(y, b)) return (int)NM_ERR_ARG;
    switch (op) {
      case 0: nm::hp_add(r, x, y); break;
      case 1: nm::hp_sub(r, x, y); break;
      case 2: nm::hp_mul(r, x, y); break;
      case 3: nm::hp_mul_si(r, x, iv); break;
      case 4: nm::hp_div_ui(r, x, (unsigned long)iv); break;
      case 5: nm::hp_set(r, x); break;
      case 6: nm::hp_mul_d(r, x, dv); break;
      default: return (int)NM_ERR_ARG;
    }
    *sign = r.sign;
    *exp = (long)r.exp;
    for (int i = 0; i < L; i++) {
      uint64_t limb = r.m[L - 1 - i];
      words[2 * i] = (uint32_t)(limb >> 32);
      words[2 * i + 1] = (uint32_t)limb;
    }
    return 2 * L;
  });
}

// ---- colouring
struct nm_palette {
  nm::MultiWave mw;
  bool loaded = false;
};

int nm_palette_create(nm_palette** out) {
  if (!out) return NM_ERR_ARG;
  return guard([&] {
    *out = new nm_palette();
    return (int)NM_OK;
  });
}

void nm_palette_destroy(nm_palette* p) { delete p; }

int nm_palette_parse(nm_palette* p, const char* text) {
  if (!p || !text) return NM_ERR_ARG;
  return guard([&] {
    nm::MultiWave w;
    if (!w.load(text) || !w.valid()) return (int)NM_ERR_ARG;
    p->mw = w;
    p->loaded = true;
    return (int)NM_OK;
  });
}

int nm_palette_load(nm_palette* p, const char* filename) {
  if (!p || !filename) return NM_ERR_ARG;
  return guard([&] {
    FILE* fp = std::fopen(filename, "r");
    if (!fp) return (int)NM_ERR_IO;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, n);
    std::fclose(fp);
    nm::MultiWave w;
    if (!w.load(text) || !w.valid()) return (int)NM_ERR_IO;
    p->mw = w;
    p->loaded = true;
    return (int)NM_OK;
  });
}

int nm_palette_save(const nm_palette* p, const char* filename) {
  if (!p || !filename || !p->loaded) return NM_ERR_ARG;
  return guard([&] {
    FILE* fp = std::fopen(filename, "w");
    if (!fp) return (int)NM_ERR_IO;
    const std::string s = p->mw.save();
    std::fwrite(s.data(), 1, s.size(), fp);
    std::fclose(fp);
    return (int)NM_OK;
  });
}

int nm_palette_cache(const nm_palette* p, int N, uint8_t* rgb) {
  if (!p || !rgb || N <= 0 || !p->loaded) return NM_ERR_ARG;
  return guard([&] {
    const std::vector<nm::Rgb> pal = p->mw.cache(N);
    for (int i = 0; i < N; i++) {
      rgb[3 * (size_t)i] = pal[i].r;
      rgb[3 * (size_t)i + 1] = pal[i].g;
      rgb[3 * (size_t)i + 2] = pal[i].b;
    }
    return (int)NM_OK;
  });
}

int nm_colour(nm_mandelbrot* h, int frame, const uint8_t* pal, int npal, int smooth, int sc, uint8_t* out,
              double* ms) {
  if (!h || !pal || !out || sc <= 0) return NM_ERR_ARG;
  return guard([&] { return h->m.colour(frame, pal, npal, smooth, sc, out, ms); });
}

int nm_colour_grid(const nm_escape_value* grid, int nr, int nc, int N, const uint8_t* pal, int npal, int smooth,
                   int sc, uint8_t* out, double* ms) {
  if (!grid || !pal || !out || nr <= 0 || nc <= 0 || sc <= 0 || N <= 0 || npal < N) return NM_ERR_ARG;
  // getColor reads pal[it] and pal[it - 1] (viewer.cpp:89-95): a negative count
  // would index before the palette on the device
  const size_t n = (size_t)nr * nc;
  for (size_t i = 0; i < n; i++)
    if (grid[i].iterations < 0) {
      g_err = "nm_colour_grid: negative iteration count";
      return NM_ERR_ARG;
    }
  return guard([&] {
    nm::Engine* eng = nullptr;
    int st = nm::Engine::instance(&eng);
    if (st != NM_OK) return st;
    return eng->colour((const nm::EscapeValue*)grid, 0, 0, nr, nc, sc, N, pal, npal, smooth, out, ms);
  });
}

// ---- zoom video
int nm_video_interpolate(const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc, int rate,
                         uint8_t* out, double* ms) {
  if (!prev || !next || !out || kr < nr || kc < nc || nr <= 0 || nc <= 0 || rate <= 0) return NM_ERR_ARG;
  return guard([&] {
    nm::Engine* eng = nullptr;
    int st = nm::Engine::instance(&eng);
    if (st != NM_OK) return st;
    return eng->video(prev, next, kr, kc, nr, nc, rate, out, ms);
  });
}

int nm_zoom_video_start(nm_mandelbrot* h, int sc) {
  if (!h) return NM_ERR_ARG;
  return guard([&] { return h->m.zoomVideoStart(sc); });
}

int nm_zoom_video_step(nm_mandelbrot* h, const uint8_t* pal, int npal, int smooth, int rate, const uint8_t* prev_key,
                       uint8_t* key_out, uint8_t* frames, int* nframes, double* ms) {
  if (!h) return NM_ERR_ARG;
  return guard([&] { return h->m.zoomVideoStep(pal, npal, smooth, rate, prev_key, key_out, frames, nframes, ms); });
}

int nm_beauty_render(const nm_mandelbrot* h, const uint8_t* pal, int npal, int smooth, uint8_t* out, double* ms) {
  if (!h || !pal || !out) return NM_ERR_ARG;
  return guard([&] { return h->m.beautyRender(pal, npal, smooth, out, ms); });
}

}  // extern "C"
