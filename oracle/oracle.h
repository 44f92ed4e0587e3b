/* oracle.h - CPU oracle for newman's per-pixel deep-zoom iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may build, load or call anything under oracle/.  The product
 * (newman_amd/) never links or calls it.
 *
 * Two restatements live here:
 *  (1) the reference method, step for step (mandelbrot.cpp:8-17, 37-360,
 *      complex.h:19-35, grid.h:8-24): setPrecision, useHardware, the pixel map,
 *      inCardioid in HP, findProbe, computeOrbit, the quirky computeSeries,
 *      getIterations (SA scan + binary search + HP continuation),
 *      getIterationsHW (plain double), the multisample average at(r,c,sc),
 *      scaleUp/scaleDown;
 *  (2) the GPU method ("pert64"): the same per-pixel SA start, then FP64
 *      perturbation with rebasing against the reference orbit.  It uses exactly
 *      the arithmetic of the HIP kernels so GPU output can be checked bit for bit.
 * Compiled with -ffp-contract=off and without -march, as the reference's
 * makefile:3 is (no FMA contraction on x86-64).
 */
#ifndef NEWMAN_ORACLE_H
#define NEWMAN_ORACLE_H

#include <stdint.h>

#include "ohp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t iterations;
  float smoothing;
} orc_escape; /* grid.h:8-15 */

typedef struct orc_view orc_view;

orc_view* orc_create(int nr, int nc); /* Mandelbrot(nr, nc), mandelbrot.cpp:8-17 */
void orc_destroy(orc_view* v);
void orc_set_N(orc_view* v, int N);
void orc_set_tol(orc_view* v, double tol);
void orc_set_bailout(orc_view* v, double bailout);
/* 1: the series test and skip in floatexp (an extension; set before precompute
 * or after it, the coefficients are kept in both forms) */
void orc_set_series_mode(orc_view* v, int fxs);
void orc_get_series_fx(const orc_view* v, double* Am, int* Ae, double* Bm, int* Be, double* Cm, int* Ce);
int orc_set_center(orc_view* v, const char* re, const char* im); /* at the current precision */
int orc_set_size(orc_view* v, const char* re, const char* im);
int orc_set_width(orc_view* v, const char* width); /* sz = width / cols at the current precision */
/* view by width: parses at the precision the target needs, sz.re = sz.im = width/nc,
 * then setPrecision */
int orc_set_view(orc_view* v, int N, const char* cre, const char* cim, const char* width);
void orc_set_precision(orc_view* v); /* mandelbrot.cpp:37-56 */
long orc_precision_bits(const orc_view* v);
int orc_use_hardware(const orc_view* v); /* mandelbrot.cpp:256-259 */
int orc_load_legacy(orc_view* v, const char* path); /* mandelbrot.cpp:19-35 */
void orc_zoom(orc_view* v, float scale);            /* mandelbrot.cpp:297-302 */
void orc_zoom_at(orc_view* v, float scale, int r, int c, int sc);
void orc_translate(orc_view* v, int dr, int dc, int sc);
void orc_scale_size(orc_view* v, int mul, int div); /* sz = sz*mul/div then setPrecision */
void orc_get_center(const orc_view* v, char* re, char* im, int cap);
void orc_get_size_d(const orc_view* v, double* re, double* im);

/* precompute: findProbe + computeSeries (mandelbrot.cpp:261-267). Returns len(X). */
int orc_precompute(orc_view* v, int nthreads);
/* precompute around an explicit reference point instead of findProbe */
int orc_precompute_at(orc_view* v, const char* re, const char* im);
int orc_probe_winner(const orc_view* v); /* index into the probe list, -1 if none */
int orc_ref_len(const orc_view* v);
/* orbit/series accessors: X (trunc and nearest doubles), A,B,C (trunc doubles) */
void orc_get_series(const orc_view* v, double* Xtrunc, double* Xnear, double* A, double* B, double* C);

/* reference semantics, full rows / selected samples */
void orc_render_rows(orc_view* v, const int* rows, int nrows, orc_escape* out, int nthreads);
void orc_render_samples(orc_view* v, const int* rc, int n, orc_escape* out, int nthreads);
/* GPU-method restatement (pert64) on selected samples */
void orc_pert_samples(orc_view* v, const int* rc, int n, orc_escape* out, int nthreads);
/* the same with FP32X deltas, as the kernel with block size K runs them */
void orc_pert_samples_fx(orc_view* v, const int* rc, const int* ua, int n, int K, orc_escape* out, int nthreads);
/* the same with a given first unstable index per sample (a tile's skip; 0 = none) */
void orc_pert_samples_ua(orc_view* v, const int* rc, const int* ua, int n, orc_escape* out, int nthreads);
/* per T x T tile (row-major): first unstable index of the reference's scan at the
 * tile's probe (its sample farthest from the reference point); returns the tile
 * count; probe_rc (NULL ok) gets (row, column) per tile */
int orc_tile_probe_ua(orc_view* v, int T, int* ua, int* probe_rc);

/* the learned skip predictor's network in double (mlp.c) */
int orc_mlp_nparams(void);
void orc_mlp_forward(const float* p, const float* x, int n, double* y);
double orc_mlp_grad(const float* p, const float* x, const float* t, int n, double* g);
/* per-sample SA start of the reference (mandelbrot.cpp:155-207): returns the first
 * unstable index (0 if none), `found`, d[found], and the SA-escape result if any. */
int orc_sa_start(orc_view* v, int r, int c, int* found, double* d, orc_escape* esc);
void orc_pixel_eps(orc_view* v, int r, int c, double* eps);   /* trunc(Y0 - X0) */
void orc_pixel_coord_d(orc_view* v, int r, int c, double* z); /* trunc(Y0) */
int orc_pixel_in_cardioid(orc_view* v, int r, int c);

/* single-point helpers at an explicit precision (bits) */
orc_escape orc_point_hw(const char* re, const char* im, int N, long bits);
int orc_point_in_cardioid(const char* re, const char* im, long bits);

/* multisample average at(r,c,sc) (mandelbrot.cpp:320-332) over a fine grid */
void orc_downsample(const orc_escape* fine, int nr, int nc, int sc, orc_escape* out);
void orc_upsample(const orc_escape* coarse, int nr, int nc, int sc, orc_escape* out);

/* HP arithmetic probe for pinning: op 0 add, 1 sub, 2 mul, 3 mul_si, 4 div_ui,
 * 5 parse(a), 6 mul_d.  Writes sign, exponent and n 32-bit limbs. */
int orc_hp_op(int op, const char* a, const char* b, long iv, double dv, long bits, int* sign, long* exp,
              uint32_t* limbs);

/* seeded zoom walk: centre strings for a config (see oracle.c) */
int orc_zoom_walk(uint64_t seed, int N, const char* target_width, int res, int nthreads, char* out_re, char* out_im,
                  int cap, int* steps);

/* pert64 render of a full small frame with the reference orbit at the centre
 * (used by the walk; exposed for tests) */
void orc_center_render(const char* cre, const char* cim, const char* width, int res, int N, int nthreads,
                       int32_t* counts);

#ifdef __cplusplus
}
#endif
#endif
