/* colour.c - CPU oracle for the viewer's colouring.  TEST INFRASTRUCTURE ONLY
 * (the same rule as oracle.h: only tests/, smoke() and bench.py's cpu_baseline
 * leg may use it; the product never does).
 *
 * Restates MultiWaveGenerator (multiwave.cpp:5-17, 19-48, 75-116) in the
 * reference's float/double mix, and FractalViewer::getColor / colorLine
 * (viewer.cpp:89-122).  hsl2rgb, interp, clip and the Pt3f average belong to
 * the viewer's image library (byteimage), which is not in the reference
 * sources: they are restated with the conventions the product documents
 * (standard HSL hexcone, hue in degrees; HSL channels rounded to the nearest
 * byte; interpolations and averages clamped and truncated).  Their agreement
 * with byteimage's own rounding is parity unpinned; the checks in
 * tests/test_colour.py pin them to closed forms (primary and grey colours,
 * endpoints of interpolation, the logistic and cycle formulas by hand).
 * Compiled with -ffp-contract=off, like the rest of the oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_CYCLES 256
#define ORC_MAX_VALUES 4096

typedef struct {
  int n;
  float v[ORC_MAX_VALUES];
  int period;
} orc_cycle;

typedef struct {
  int ncyc;
  orc_cycle hue[ORC_MAX_CYCLES];
  int hue_period;
  orc_cycle sat;
  int nwaves;
  float amp[ORC_MAX_CYCLES];
  int wper[ORC_MAX_CYCLES];
} orc_multiwave;

static orc_multiwave g_mw;

/* load_filename (multiwave.cpp:19-48): whitespace-separated numbers read in
 * order; returns 0 on success */
int orc_pal_parse(const char* text) {
  const char* p = text;
  char* end;
  long n;
#define NEXT_INT(dst)                     \
  do {                                    \
    long t_ = strtol(p, &end, 10);        \
    if (end == p) return 1;               \
    p = end;                              \
    (dst) = (int)t_;                      \
  } while (0)
#define NEXT_FLT(dst)                     \
  do {                                    \
    float t_ = strtof(p, &end);           \
    if (end == p) return 1;               \
    p = end;                              \
    (dst) = t_;                           \
  } while (0)
  memset(&g_mw, 0, sizeof g_mw);
  NEXT_INT(n);
  if (n < 1 || n > ORC_MAX_CYCLES) return 1;
  g_mw.ncyc = (int)n;
  for (int i = 0; i < g_mw.ncyc; i++) {
    NEXT_INT(n);
    if (n < 1 || n > ORC_MAX_VALUES) return 1;
    g_mw.hue[i].n = (int)n;
    for (int j = 0; j < g_mw.hue[i].n; j++) NEXT_FLT(g_mw.hue[i].v[j]);
    NEXT_INT(g_mw.hue[i].period);
  }
  NEXT_INT(g_mw.hue_period);
  NEXT_INT(n);
  if (n < 1 || n > ORC_MAX_VALUES) return 1;
  g_mw.sat.n = (int)n;
  for (int j = 0; j < g_mw.sat.n; j++) NEXT_FLT(g_mw.sat.v[j]);
  NEXT_INT(g_mw.sat.period);
  NEXT_INT(n);
  if (n < 0 || n > ORC_MAX_CYCLES) return 1;
  g_mw.nwaves = (int)n;
  for (int i = 0; i < g_mw.nwaves; i++) {
    NEXT_FLT(g_mw.amp[i]);
    NEXT_INT(g_mw.wper[i]);
  }
#undef NEXT_INT
#undef NEXT_FLT
  return 0;
}

/* the library helpers, with the documented conventions */
uint8_t orc_clip(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)(int)v;
}

void orc_hsl2rgb(float h, float s, float l, uint8_t* rgb) {
  float hh = fmodf(h, 360.0f);
  if (hh < 0.0f) hh += 360.0f;
  float c = (1.0f - fabsf(2.0f * l - 1.0f)) * s;
  float hp = hh / 60.0f;
  float x = c * (1.0f - fabsf(fmodf(hp, 2.0f) - 1.0f));
  float r = 0.0f, g = 0.0f, b = 0.0f;
  int sector = (int)hp;
  if (sector == 0) { r = c; g = x; }
  else if (sector == 1) { r = x; g = c; }
  else if (sector == 2) { g = c; b = x; }
  else if (sector == 3) { g = x; b = c; }
  else if (sector == 4) { r = x; b = c; }
  else { r = c; b = x; }
  float m = l - 0.5f * c;
  rgb[0] = orc_clip((r + m) * 255.0f + 0.5f);
  rgb[1] = orc_clip((g + m) * 255.0f + 0.5f);
  rgb[2] = orc_clip((b + m) * 255.0f + 0.5f);
}

void orc_interp(const uint8_t* a, const uint8_t* b, float t, uint8_t* out) {
  float u = 1.0f - t;
  for (int k = 0; k < 3; k++) out[k] = orc_clip(u * (float)a[k] + t * (float)b[k]);
}

/* FloatCycle::value (multiwave.cpp:5-12): values.size() * (step % period) is an
 * unsigned product converted to float; (1.0 - t) is a double */
static float cycle_value(const orc_cycle* c, int step) {
  float t = (float)((unsigned long)c->n * (unsigned long)(step % c->period)) / (float)c->period;
  int i0 = (int)t;
  int i1 = (int)((unsigned long)(i0 + 1) % (unsigned long)c->n);
  t -= (float)i0;
  float prod = t * c->v[i1];
  return (float)((1.0 - (double)t) * (double)c->v[i0] + (double)prod);
}

/* FloatWave::value (multiwave.cpp:14-17): sin of the float argument in double */
static float wave_value(float amp, int period, int step) {
  const float tau = (float)(2.0 * 3.14159265358979);
  float arg = (float)step * tau / (float)period;
  return (float)((double)amp * sin((double)arg));
}

/* MultiWaveGenerator::cache (multiwave.cpp:75-116) into rgb (3 N bytes) */
void orc_pal_cache(int N, uint8_t* rgb) {
  for (int i = 0; i < N; i++) {
    float sat = cycle_value(&g_mw.sat, i);
    float lum = 0.0f;
    for (int w = 0; w < g_mw.nwaves; w++) lum += wave_value(g_mw.amp[w], g_mw.wper[w], i);
    lum = (float)(1.0 / (1.0 + exp(-(double)lum)));
    float ty = (float)((unsigned long)g_mw.ncyc * (unsigned long)(i % g_mw.hue_period)) / (float)g_mw.hue_period;
    int y0 = (int)ty;
    int y1 = (int)((unsigned long)(y0 + 1) % (unsigned long)g_mw.ncyc);
    ty -= (float)y0;
    uint8_t c0[3], c1[3], first[3], second[3];
    const orc_cycle* cy = &g_mw.hue[y0];
    float tx = (float)((unsigned long)cy->n * (unsigned long)(i % cy->period)) / (float)cy->period;
    int x0 = (int)tx;
    int x1 = (int)((unsigned long)(x0 + 1) % (unsigned long)cy->n);
    tx -= (float)x0;
    orc_hsl2rgb(cy->v[x0], sat, lum, c0);
    orc_hsl2rgb(cy->v[x1], sat, lum, c1);
    orc_interp(c0, c1, tx, first);
    cy = &g_mw.hue[y1];
    tx = (float)((unsigned long)cy->n * (unsigned long)(i % cy->period)) / (float)cy->period;
    x0 = (int)tx;
    x1 = (int)((unsigned long)(x0 + 1) % (unsigned long)cy->n);
    tx -= (float)x0;
    orc_hsl2rgb(cy->v[x0], sat, lum, c0);
    orc_hsl2rgb(cy->v[x1], sat, lum, c1);
    orc_interp(c0, c1, tx, second);
    orc_interp(first, second, ty, rgb + 3 * (size_t)i);
  }
}

/* getColor (viewer.cpp:89-95); pal[-1] does not exist in the reference's
 * palette: entry 0 stands in */
static void get_color(int it, float sm, int N, const uint8_t* pal, int smooth, uint8_t* out) {
  if (it >= N) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  if (!smooth) {
    memcpy(out, pal + 3 * (size_t)it, 3);
    return;
  }
  orc_interp(pal + 3 * (size_t)(it > 0 ? it - 1 : 0), pal + 3 * (size_t)it, sm, out);
}

/* colorLine (viewer.cpp:97-122) for every row: grid of nr x nc
 * {iterations, smoothing} pairs, out (nr / sc) x (nc / sc) x 3 bytes */
void orc_colour(const int32_t* its, const float* sms, int nr, int nc, int N, const uint8_t* pal, int smooth,
                int sc, uint8_t* out) {
  int nro = nr / sc, nco = nc / sc;
  for (int r = 0; r < nro; r++)
    for (int c = 0; c < nco; c++) {
      uint8_t* o = out + 3 * ((size_t)r * nco + c);
      if (sc == 1) {
        get_color(its[(size_t)r * nc + c], sms[(size_t)r * nc + c], N, pal, smooth, o);
        continue;
      }
      float sum[3] = {0.0f, 0.0f, 0.0f};
      for (int r1 = r * sc; r1 < (r + 1) * sc; r1++)
        for (int c1 = c * sc; c1 < (c + 1) * sc; c1++) {
          uint8_t col[3];
          get_color(its[(size_t)r1 * nc + c1], sms[(size_t)r1 * nc + c1], N, pal, smooth, col);
          for (int k = 0; k < 3; k++) sum[k] += (float)col[k];
        }
      for (int k = 0; k < 3; k++) o[k] = orc_clip(sum[k] / (float)(sc * sc));
    }
}

/* ---- VideoZoom::nextFrame (video.cpp:14-34) with its images materialised as
 * the reference does: small = next.scaled(), large = prev.scaled(), then
 * canvas.blit(large, ...) and canvas.blend(small, t, ...).  scaled/blit/blend
 * belong to byteimage (absent): bilinear with pixel-centre alignment and clamped
 * edges rounded to the nearest byte; blit copies with clipping; blend clamps
 * and truncates (1 - t) canvas + t image. */
static void scale_image(const uint8_t* src, int sr, int sc, uint8_t* dst, int dr, int dc) {
  for (int i = 0; i < dr; i++)
    for (int j = 0; j < dc; j++) {
      float y = ((float)i + 0.5f) * ((float)sr / (float)dr) - 0.5f;
      float x = ((float)j + 0.5f) * ((float)sc / (float)dc) - 0.5f;
      if (y < 0.0f) y = 0.0f;
      if (y > (float)(sr - 1)) y = (float)(sr - 1);
      if (x < 0.0f) x = 0.0f;
      if (x > (float)(sc - 1)) x = (float)(sc - 1);
      int y0 = (int)y, x0 = (int)x;
      int y1 = y0 + 1 < sr ? y0 + 1 : sr - 1;
      int x1 = x0 + 1 < sc ? x0 + 1 : sc - 1;
      float fy = y - (float)y0, fx = x - (float)x0;
      for (int k = 0; k < 3; k++) {
        float a = src[3 * ((size_t)y0 * sc + x0) + k], b = src[3 * ((size_t)y0 * sc + x1) + k];
        float c = src[3 * ((size_t)y1 * sc + x0) + k], d = src[3 * ((size_t)y1 * sc + x1) + k];
        float top = (1.0f - fx) * a + fx * b;
        float bot = (1.0f - fx) * c + fx * d;
        dst[3 * ((size_t)i * dc + j) + k] = orc_clip((1.0f - fy) * top + fy * bot + 0.5f);
      }
    }
}

/* frames: rate x nr x nc x 3 */
int orc_video_frames(const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc, int rate,
                     uint8_t* frames) {
  uint8_t* canvas = (uint8_t*)calloc((size_t)nr * nc * 3, 1);
  uint8_t* small = (uint8_t*)malloc((size_t)kr * kc * 3 + 3);
  uint8_t* large = (uint8_t*)malloc((size_t)kr * kc * 12 + 3);
  if (!canvas || !small || !large) return 1;
  float v = (float)pow(1.5, 1.0 / rate), sm = (float)(2.0 / 3.0), lg = 1.0f;
  for (int i = 0; i < rate; i++) {
    float t = (float)i / (float)rate;
    int srr = (int)((float)kr * sm), scc = (int)((float)kc * sm);
    int lrr = (int)((float)kr * lg), lcc = (int)((float)kc * lg);
    if ((size_t)lrr * lcc > (size_t)kr * kc * 4) return 2; /* lg stays below 1.5 */
    scale_image(next, kr, kc, small, srr, scc);
    scale_image(prev, kr, kc, large, lrr, lcc);
    int lr0 = (nr - lrr) / 2, lc0 = (nc - lcc) / 2;
    for (int r = 0; r < lrr; r++)
      for (int c = 0; c < lcc; c++) {
        int rr = r + lr0, cc = c + lc0;
        if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
        memcpy(canvas + 3 * ((size_t)rr * nc + cc), large + 3 * ((size_t)r * lcc + c), 3);
      }
    int sr0 = (nr - srr) / 2, sc0 = (nc - scc) / 2;
    for (int r = 0; r < srr; r++)
      for (int c = 0; c < scc; c++) {
        int rr = r + sr0, cc = c + sc0;
        if (rr < 0 || rr >= nr || cc < 0 || cc >= nc) continue;
        uint8_t* o = canvas + 3 * ((size_t)rr * nc + cc);
        const uint8_t* q = small + 3 * ((size_t)r * scc + c);
        for (int k = 0; k < 3; k++) o[k] = orc_clip((1.0f - t) * (float)o[k] + t * (float)q[k]);
      }
    sm *= v;
    lg *= v;
    memcpy(frames + (size_t)i * nr * nc * 3, canvas, (size_t)nr * nc * 3);
  }
  free(canvas);
  free(small);
  free(large);
  return 0;
}
