/* newman_amd.h - C ABI of the MI355X deep-zoom render core.
 *
 * Drop-in boundary for newman's render core, class Mandelbrot (mandelbrot.h:7-53):
 * every entry point below replaces one member or public field of that class and
 * cites it.  Entry points marked "extension" have no counterpart in the
 * reference; they expose what the GPU engine adds (batched frames, the bailout
 * parameter, the delta arithmetic, render statistics).
 *
 * Conventions: plain C types only; every function returns an int status (0 = ok,
 * < 0 = error, see NM_ERR_*) and never throws; all calls are synchronous; output
 * buffers are owned by the caller.  Calls are serialised by a process-wide lock,
 * so different handles may be used from different threads; one handle must not be
 * used from two threads at once.  HP values cross the boundary as decimal strings ("-0.75",
 * "0.123e-30"), as the reference's view files carry them (viewer.cpp:12-22).
 */
#ifndef NEWMAN_AMD_H
#define NEWMAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NM_OK 0
#define NM_ERR_ARG (-1)      /* bad argument (sizes, N <= 0, malformed string) */
#define NM_ERR_HIP (-2)      /* HIP runtime error (nm_last_error has the text) */
#define NM_ERR_STATE (-3)    /* call order: e.g. deep view rendered before precompute */
#define NM_ERR_NODEVICE (-4) /* no HIP device: there is no CPU fallback */
#define NM_ERR_RANGE (-5)    /* frame set larger than the engine handles */
#define NM_ERR_IO (-6)       /* view file missing or malformed */
#define NM_ERR_NOMEM (-7)

#define NM_DELTA_FP64 0  /* perturbation deltas in FP64 */
#define NM_DELTA_FP32X 1 /* extension: FP32 mantissa + int32 exponent deltas */
#define NM_SERIES_DOUBLE 0   /* the reference's series test and skip in double */
#define NM_SERIES_FLOATEXP 1 /* extension: the same in double mantissa + int exponent */

/* grid.h:8-15 RenderGrid::EscapeValue, 8 bytes, same layout */
typedef struct nm_escape_value {
  int32_t iterations;
  float smoothing;
} nm_escape_value;

/* extension: counters of the last render */
typedef struct nm_render_stats {
  uint64_t samples;      /* samples in the frame set */
  uint64_t n_cardioid;   /* resolved by the cardioid/bulb test (mandelbrot.cpp:63-71) */
  uint64_t n_sa_resolved; /* resolved inside the series skip: escaped (mandelbrot.cpp:186-206) or skip reached N */
  uint64_t n_iterated;   /* handed to the iteration kernel */
  uint64_t executed;     /* iteration steps executed (start and iteration kernels) */
  uint64_t skipped;      /* sum of series-skip start indices of iterated samples */
  uint64_t n_rebased;    /* blocks of block_k steps in which a sample rebased its delta */
  uint64_t credited;     /* sum of reported iterations over non-cardioid samples */
  double ms_start;       /* start kernel, ms (HIP events), mean over repeats */
  double ms_iterate;     /* iteration kernel, ms, mean over repeats */
  double ms_total;       /* all kernels of one pass, ms, mean over repeats */
  int32_t block_k;       /* iterations per block used */
  int32_t waves_per_simd;
  uint64_t executed_start; /* of executed: steps run inside the start kernel */
  double ms_compact;     /* tile scan + scatter of the active list, ms */
  double ms_finalize;    /* smoothing / output kernel, ms */
  int32_t graph;         /* 1: ms_total is the mean of HIP graph replays of the pass */
  int32_t tile;          /* sample order used: 0 row-major, T = T x T tiles */
  int32_t fma2;          /* plain path: 1 when y' = (2x) y + c ran as fma(x y, 2, c), bit-identical for frames
                            whose coordinates are 0 or at least 2^-200 in magnitude */
  int32_t grouped;       /* plain path: 1 when the escape test ran once per group of four updates, the first
                            escaping update found from the group's earlier |z|^2 (bailout radius 2 to 2^30 and
                            every |c| <= 0.999 (R^2 - R): an escape at one update implies one at the next) */
  int32_t chunk_run;     /* the iterate kernel's chunk order used: 64 = 64 consecutive samples of the active
                            list per chunk, R < 64 = runs of R (nm_launch.chunk_run; the plain path: 1) */
} nm_render_stats;

typedef struct nm_mandelbrot nm_mandelbrot; /* class Mandelbrot (mandelbrot.h:7) */

/* Mandelbrot(int nr, int nc), mandelbrot.cpp:8-17: tol 1e-10, N 256, centre
 * -0.5+0i, sz 4/nc x 3/nr, then setPrecision. */
int nm_create(nm_mandelbrot** out, int nr, int nc);
/* copy construction / assignment (viewer.cpp:193, 251 copy a Mandelbrot by value) */
int nm_clone(nm_mandelbrot** out, const nm_mandelbrot* src);
void nm_destroy(nm_mandelbrot* h);

int nm_rows(const nm_mandelbrot* h);                 /* rows(), mandelbrot.h:33 */
int nm_cols(const nm_mandelbrot* h);                 /* cols(), mandelbrot.h:34 */
int nm_set_N(nm_mandelbrot* h, int N);               /* public field N, mandelbrot.h:24 */
int nm_get_N(const nm_mandelbrot* h);
int nm_set_error_tolerance(nm_mandelbrot* h, double tol); /* error_tolerance, mandelbrot.h:23 */
double nm_get_error_tolerance(const nm_mandelbrot* h);
/* public field center (mandelbrot.h:25): assignment keeps the field's precision,
 * as mpf_class assignment does */
int nm_set_center(nm_mandelbrot* h, const char* re, const char* im);
int nm_get_center(const nm_mandelbrot* h, char* re, char* im, size_t cap);
/* public field sz (mandelbrot.h:25) */
int nm_set_size(nm_mandelbrot* h, const char* re, const char* im);
int nm_get_size(const nm_mandelbrot* h, char* re, char* im, size_t cap);
int nm_get_size_d(const nm_mandelbrot* h, double* re, double* im);
/* setPrecision (mandelbrot.cpp:37-56, protected in the reference; callers that
 * write sz directly call it to re-derive the precision) */
int nm_set_precision(nm_mandelbrot* h);
long nm_precision_bits(const nm_mandelbrot* h);

int nm_use_hardware(nm_mandelbrot* h);        /* useHardware(), mandelbrot.cpp:256-259: 1/0 */
int nm_precompute(nm_mandelbrot* h);          /* precompute(), mandelbrot.cpp:261-267 */
int nm_compute_row(nm_mandelbrot* h, int r);  /* computeRow(int), mandelbrot.cpp:269-283 */
/* computeRow for every row, one launch (viewer.cpp:166-171 loops computeRow) */
int nm_compute_all(nm_mandelbrot* h);

/* No counterpart in the reference: nm_compute_all for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Renders h's view on the plain path
 * (useHardware() true; deep views give NM_ERR_STATE) into out (n >= rows x cols
 * escape values, row-major) with every device buffer in the library's static
 * device scratch array: no hipMalloc/hipFree, no stream or event created.  At
 * most 2^19 samples and 8192 rows or columns (NM_ERR_RANGE beyond).  Results
 * equal nm_compute_all's.  All calls on one GPU share the scratch array, which
 * keeps the last call's contents; the library's process-wide lock serialises
 * them (calls that overlapped on one GPU would race). */
int nm_compute_all_noalloc(nm_mandelbrot* h, nm_escape_value* out, size_t n);

int nm_at(nm_mandelbrot* h, int r, int c, nm_escape_value* out);         /* at(r,c), mandelbrot.cpp:318 */
int nm_at_avg(nm_mandelbrot* h, int r, int c, int sc, nm_escape_value* out); /* at(r,c,sc), :320-332 */
/* the whole grid, row-major (RenderGrid::values, grid.h:18) */
int nm_get_grid(nm_mandelbrot* h, nm_escape_value* out, size_t n);

/* pointAt(r,c,sc), mandelbrot.cpp:285-290, as decimal strings */
int nm_point_at(const nm_mandelbrot* h, int r, int c, int sc, char* re, char* im, size_t cap);
int nm_translate(nm_mandelbrot* h, int dr, int dc, int sc);    /* mandelbrot.cpp:292-295 */
int nm_zoom(nm_mandelbrot* h, float scale);                    /* mandelbrot.cpp:297-302 */
int nm_zoom_at(nm_mandelbrot* h, float scale, int r, int c, int sc); /* mandelbrot.cpp:304-316 */
int nm_scale_up(nm_mandelbrot* h, int sc);                     /* mandelbrot.cpp:334-346 */
int nm_scale_down(nm_mandelbrot* h, int sc);                   /* mandelbrot.cpp:348-360 */
int nm_load_legacy(nm_mandelbrot* h, const char* path);        /* mandelbrot.cpp:19-35 */
/* the view-file writer of viewer.cpp:6-29 (N, centre, size as decimal strings) */
int nm_save_view(const nm_mandelbrot* h, const char* path);

/* ---- extensions ----------------------------------------------------------- */
/* bailout radius; the reference's constant is 1024 (mandelbrot.cpp:58-59) */
int nm_set_bailout(nm_mandelbrot* h, double bailout);
/* NM_DELTA_FP64 (default) or NM_DELTA_FP32X */
int nm_set_delta_mode(nm_mandelbrot* h, int mode);
/* The series test and skip of the series path (mandelbrot.cpp:138-142, 160-178):
 * NM_SERIES_DOUBLE (default) in the reference's doubles; NM_SERIES_FLOATEXP
 * (extension) in the same operation order with each real held as a double
 * mantissa and an int exponent: the same results wherever the reference's
 * doubles are normal numbers, and a working test below widths of about 1e-90,
 * where the reference's |B e^2|^2 and |C e^3|^2 leave the double range and every
 * index tests stable.  Takes effect on a precomputed reference at once. */
int nm_set_series_mode(nm_mandelbrot* h, int mode);
/* a view by width: centre strings and width = cols * sz.re (square samples),
 * parsed at the precision the width needs, then setPrecision */
int nm_set_view(nm_mandelbrot* h, int N, const char* cre, const char* cim, const char* width);
/* precompute around an explicit reference point instead of findProbe */
int nm_precompute_at(nm_mandelbrot* h, const char* re, const char* im);
int nm_reference_length(const nm_mandelbrot* h);  /* len(X) */
/* the reference point X[0] as decimal strings (precompute first) */
int nm_reference_point(const nm_mandelbrot* h, char* re, char* im, size_t cap);
int nm_probe_winner(const nm_mandelbrot* h);      /* index of findProbe's winner, -1 if none */
/* how precompute runs findProbe (mandelbrot.cpp:73-95): NM_PROBE_HOST iterates
 * every probe in HP on the host; NM_PROBE_GPU (the default) screens the probes
 * by FP64 perturbation on the device and iterates in HP only those within 3
 * iterations of the best estimate, falling back to the host when no device is
 * present.  nm_probe_candidates: probes the last search iterated in HP. */
#define NM_PROBE_HOST 0
#define NM_PROBE_GPU 1
int nm_set_probe_mode(nm_mandelbrot* h, int mode);
int nm_probe_candidates(const nm_mandelbrot* h);
/* reference series as the device sees it: X (nearest), A, B, C (truncated),
 * 2*len doubles each (re, im interleaved); any pointer may be NULL */
int nm_get_reference(const nm_mandelbrot* h, double* X, double* A, double* B, double* C);
/* render every row with explicit engine settings and `repeats` timed launches;
 * copy_back = 0 leaves the grid on the device (benchmarks).  block_k: 0 for the
 * default (plain 64; series FP64 64, FP32X 8), else plain 8/16/32/64, series
 * FP64 8/12/16/32/64, FP32X 8/12/16/32; other values give NM_ERR_ARG. */
int nm_render(nm_mandelbrot* h, int block_k, int waves_per_simd, int repeats, int copy_back,
              nm_render_stats* stats);
/* render a batch of frames that share h's reference orbit (zoom video): frame
 * k has h's centre and size width_k = cols * sz; outputs are concatenated
 * row-major grids (out = NULL keeps them on the device).  widths are decimal
 * strings. */
int nm_render_frames(nm_mandelbrot* h, const char* const* widths, int nframes, int block_k, int waves_per_simd,
                     int repeats, nm_escape_value* out, nm_render_stats* stats);
/* extension: launch settings of later renders on h.  0 keeps the engine's
 * default; pre_iter must be 0 or 1..1024; tile = sample order (0 row-major, or T x T tiles of 8 x 8 blocks, T in
 * 8..256, a power of two); graph = 1 replays repeated passes from one captured
 * HIP graph; ms = the kernel time the autotuner measured (output only). */
typedef struct nm_launch {
  int32_t block_k;
  int32_t waves_per_simd;
  int32_t tile;
  int32_t graph;
  int32_t pre_iter; /* plain path: updates tried in the start kernel (0 = default 16) */
  int32_t chunk_run; /* series path: each wave takes 64 consecutive samples of the active list (64)
                        or 64 / R runs of R consecutive samples spread over the list (R = 1, 2, 4, 8,
                        16, 32): spreads deep views' long samples over the waves; 0 = the default,
                        runs of 4 when the deltas follow a floatexp series, else 64 */
  double ms;
} nm_launch;
int nm_set_launch(nm_mandelbrot* h, const nm_launch* l);
int nm_get_launch(const nm_mandelbrot* h, nm_launch* l);
/* extension: the launch autotuner for a frame set (see nm_render_frames):
 * sweeps sample order x waves per SIMD x block size on an evenly spaced
 * subsample of at most 16 frames, makes the fastest h's launch settings and
 * writes it to *best; up to cap sweep rows go to table (NULL ok), their count
 * to *n_table. */
int nm_autotune_frames(nm_mandelbrot* h, const char* const* widths, int nframes, nm_launch* best,
                       nm_launch* table, int cap, int* n_table);
/* ---- learned series skip (extension) ---------------------------------------
 * A per-tile alternative to the reference's per-pixel series scan
 * (mandelbrot.cpp:165-181): every sample of a T x T tile starts at the skip of the
 * tile's probe, its sample farthest from the reference point (the one whose scan
 * stops first).  The probe's first unstable index comes from the reference's scan,
 * from the threshold tables, or from a small network (nm_skipnet: 132 inputs, two
 * hidden layers of 64, ReLU, FP32) whose every prediction is validated by the
 * reference's stability test, with the scan as fallback. */
#define NM_SKIP_PIXEL 0
#define NM_SKIP_TILE_SCAN 1
#define NM_SKIP_TILE_TABLE 2
#define NM_SKIP_TILE_LEARNED 3
#define NM_SKIP_NONE 4 /* no series: every sample starts at d_0 = e (plain perturbation) */
#define NM_SKIPNET_INPUTS 132
#define NM_SKIPNET_PARAMS 12737
typedef struct nm_skipnet nm_skipnet;
typedef struct nm_tile_stats {
  int32_t ntiles;
  int32_t accepted;   /* predictions that were the first unstable index */
  int32_t rejected;   /* predictions the check rejected: the tile was scanned from index 1 */
  int32_t bounded;    /* predictions the check accepted as a lower bound: scanned from them */
  uint64_t evals;     /* exact stability tests evaluated at the probes */
  double ms_search;   /* tile kernel, ms */
  double ms_predict;  /* features, network and snapping, ms (host wall time) */
} nm_tile_stats;
int nm_skipnet_create(nm_skipnet** out);
void nm_skipnet_destroy(nm_skipnet* net);
/* params: NM_SKIPNET_PARAMS floats (W1[64][132], b1, W2[64][64], b2, W3[64], b3);
 * mean, scale: the input standardisation (NM_SKIPNET_INPUTS each) */
int nm_skipnet_get(const nm_skipnet* net, float* params, float* mean, float* scale);
int nm_skipnet_set(nm_skipnet* net, const float* params, const float* mean, const float* scale);
/* trains on the GPU (HIP forward, backward and Adam kernels): standardises the
 * inputs on this data, initialises from seed, minimises the mean squared error;
 * curve (epochs floats, NULL ok) gets each epoch's first-batch loss */
int nm_skipnet_train(nm_skipnet* net, const float* x, const float* t, int n, int epochs, int batch, float lr,
                     uint64_t seed, float* curve);
int nm_skipnet_predict(nm_skipnet* net, const float* x, int n, float* y);
/* loss and gradient of a batch of standardised inputs, on the GPU */
int nm_skipnet_gradient(nm_skipnet* net, const float* xn, const float* t, int n, float* grad, float* loss);
/* the view's tile features (NM_SKIPNET_INPUTS per tile, row-major tiles) */
int nm_tile_features(nm_mandelbrot* h, int tile, float* x, int cap_tiles, int* ntiles);
/* first unstable index of every tile probe (0 = none): mode 0 scan, 1 tables,
 * 2 learned (net required); how (NULL ok): 0 scan, 1 tables, 2 prediction exact,
 * 3 rejected (scanned from 1), 4 a certified lower bound (scanned from it) */
int nm_tile_skips(nm_mandelbrot* h, int tile, int mode, nm_skipnet* net, int* ua, int* how, int cap_tiles,
                  nm_tile_stats* stats);
/* the network's training target of each first unstable index for this view: the
 * rank of the index among the indices where the threshold prefix minimum drops,
 * t = (rank + 1) / (R + 1), and 1 for 0 ("never unstable") */
int nm_tile_targets(nm_mandelbrot* h, const int* ua, int n, float* t);
/* skip granularity of later renders on the series path (NM_SKIP_*); the net is
 * not owned and must outlive its use */
int nm_set_skip_mode(nm_mandelbrot* h, int mode, int tile, nm_skipnet* net);
int nm_get_tile_stats(const nm_mandelbrot* h, nm_tile_stats* stats);

/* ---- colouring: the viewer's palette and colour lookup (next to the grid) ----
 * MultiWaveGenerator (multiwave.h:7-37): hue cycles, a saturation cycle and
 * luminance waves; the .pal text format of multiwave.cpp:19-73. */
typedef struct nm_palette nm_palette;
int nm_palette_create(nm_palette** out); /* empty: load or parse before use */
void nm_palette_destroy(nm_palette* p);
int nm_palette_load(nm_palette* p, const char* filename); /* load_filename, multiwave.cpp:19-48 */
int nm_palette_save(const nm_palette* p, const char* filename); /* save_filename, multiwave.cpp:50-73 */
int nm_palette_parse(nm_palette* p, const char* text); /* the same format from a string */
/* MultiWaveGenerator::cache(N) (multiwave.cpp:75-116): N colours, 3 bytes (R, G,
 * B) each, into rgb (3 N bytes).  hsl2rgb, interp and clip come from the viewer's
 * image library, which is not part of the reference: HSL is the standard hexcone
 * (hue in degrees), HSL channels round to the nearest byte, interpolations and
 * averages clamp and truncate. */
int nm_palette_cache(const nm_palette* p, int N, uint8_t* rgb);
/* FractalViewer::getColor / colorLine (viewer.cpp:89-122) on the GPU: the grid
 * (frame < 0) or frame `frame` of the last nm_render_frames of h, coloured with
 * pal (npal >= N entries of 3 bytes; black at or beyond N; smooth != 0
 * interpolates pal[it - 1] -> pal[it] by the smoothing, with pal[0] standing in
 * for pal[-1]) and averaged over sc x sc samples, into (rows / sc) x (cols / sc)
 * x 3 bytes of interleaved RGB, row-major.  The grid is taken from the device
 * when it still holds h's last render (no host round trip), else from h's host
 * grid.  ms (NULL ok): the colouring kernel's time. */
int nm_colour(nm_mandelbrot* h, int frame, const uint8_t* pal, int npal, int smooth, int sc, uint8_t* out,
              double* ms);
/* the same for a grid given by the caller (nr x nc escape values, row-major) */
int nm_colour_grid(const nm_escape_value* grid, int nr, int nc, int N, const uint8_t* pal, int npal, int smooth,
                   int sc, uint8_t* out, double* ms);

/* ---- the zoom video: keyframes 1.5x apart and the frames between them ----
 * VideoZoom::nextFrame (video.cpp:14-34) on the GPU: the `rate` frames (nr x nc
 * x 3 bytes each, into out) between keyframes prev and next (kr x kc x 3 bytes,
 * kr >= nr, kc >= nc): frame i blits prev scaled by lg and blends next scaled by
 * sm with weight i / rate (lg from 1, sm from 2/3, both times 1.5^(1/rate) per
 * frame).  byteimage's scaled/blit/blend are restated: bilinear with
 * pixel-centre alignment rounded to the nearest byte, truncating blend. */
int nm_video_interpolate(const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc, int rate,
                         uint8_t* out, double* ms);
/* initAutoZoom (viewer.cpp:441-453) for a grid holding sc x sc samples per
 * output pixel: full view (sz = 4 / cols, 3 / rows), scaleDown(sc), scaleUp(3) */
int nm_zoom_video_start(nm_mandelbrot* h, int sc);
/* one keyframe of the viewer's zoom loop (viewer.cpp:270-286): precompute and
 * render the 3 x 3-sampled grid, colour it 2 x 2-averaged into key_out
 * ((rows / 2) x (cols / 2) x 3), write the rate in-between frames from prev_key
 * (NULL for the first keyframe: no frames) into frames ((rows / 3) x (cols / 3)
 * x 3 each), *nframes = their count, then zoom(1.5).  ms (NULL ok): colouring +
 * interpolation kernel time. */
int nm_zoom_video_step(nm_mandelbrot* h, const uint8_t* pal, int npal, int smooth, int rate, const uint8_t* prev_key,
                       uint8_t* key_out, uint8_t* frames, int* nframes, double* ms);
/* beautyRender (viewer.cpp:186-253): h's view (N, centre, sz scaled by rows /
 * 3240; error tolerance and bailout at their defaults, as there) rendered on a
 * 3240 x 5760 grid and coloured 3 x 3-averaged into out (1080 x 1920 x 3 bytes);
 * h is unchanged.  ms (NULL ok): colouring time. */
int nm_beauty_render(const nm_mandelbrot* h, const uint8_t* pal, int npal, int smooth, uint8_t* out, double* ms);

/* the text of the last engine error */
const char* nm_last_error(void);

/* HP arithmetic probe for tests: op 0 add, 1 sub, 2 mul, 3 mul by int, 4 divide by
 * unsigned int, 5 parse, 6 mul by double; writes sign, exponent (value = sign *
 * 0.m * 2^exp) and the mantissa as 32-bit words, most significant first.
 * Returns the number of words. */
int nm_hp_op(int op, const char* a, const char* b, long iv, double dv, long bits, int* sign, long* exp,
             uint32_t* words);

#ifdef __cplusplus
}
#endif
#endif
