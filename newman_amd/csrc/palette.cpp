// palette.cpp - the multi-wave palette (multiwave.cpp) in the reference's float
// arithmetic, and the colour helpers its library (byteimage, not part of the
// reference sources) provides there: clip, hsl2rgb, interp.  Their exact
// rounding in byteimage cannot be checked here; the conventions below are this
// project's (clamp-and-truncate for clip, round-to-nearest for HSL channels,
// the standard HSL hexcone for the colour model), shared with the GPU kernel and
// the oracle.
#include "palette.hpp"

#include <cmath>
#include <cstdio>
#include <sstream>

namespace nm {

uint8_t clip_byte(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint8_t)(int)v;
}

Rgb hsl_to_rgb(float h, float s, float l) {
  float hh = std::fmod(h, 360.0f);
  if (hh < 0.0f) hh += 360.0f;
  const float c = (1.0f - std::fabs(2.0f * l - 1.0f)) * s;
  const float hp = hh / 60.0f;
  const float x = c * (1.0f - std::fabs(std::fmod(hp, 2.0f) - 1.0f));
  float r1 = 0.0f, g1 = 0.0f, b1 = 0.0f;
  switch ((int)hp) {
    case 0: r1 = c; g1 = x; break;
    case 1: r1 = x; g1 = c; break;
    case 2: g1 = c; b1 = x; break;
    case 3: g1 = x; b1 = c; break;
    case 4: r1 = x; b1 = c; break;
    default: r1 = c; b1 = x; break;
  }
  const float m = l - 0.5f * c;
  Rgb o;
  o.r = clip_byte((r1 + m) * 255.0f + 0.5f);
  o.g = clip_byte((g1 + m) * 255.0f + 0.5f);
  o.b = clip_byte((b1 + m) * 255.0f + 0.5f);
  return o;
}

Rgb interp(Rgb a, Rgb b, float t) {
  const float u = 1.0f - t;
  Rgb o;
  o.r = clip_byte(u * a.r + t * b.r);
  o.g = clip_byte(u * a.g + t * b.g);
  o.b = clip_byte(u * a.b + t * b.b);
  return o;
}

// t = size * (step % period) / (float) period in the reference's types: the
// product is an unsigned integer converted to float
static float cycle_pos(size_t n, int step, int period) {
  return (float)(n * (size_t)(step % period)) / (float)period;
}

// multiwave.cpp:5-12; (1.0 - t) is a double, t * values[i1] a float product
float MultiWave::FloatCycle::value(int step) const {
  float t = cycle_pos(values.size(), step, period);
  const int i0 = (int)t;
  const int i1 = (int)((size_t)(i0 + 1) % values.size());
  t -= (float)i0;
  const double a = (1.0 - (double)t) * (double)values[i0];
  const float b = t * values[i1];
  return (float)(a + (double)b);
}

// multiwave.cpp:14-17; sin of the float argument in double
float MultiWave::FloatWave::value(int step) const {
  const float tau = (float)(2.0 * 3.14159265358979);
  const float arg = (float)step * tau / (float)period;
  return (float)((double)amplitude * std::sin((double)arg));
}

bool MultiWave::valid() const {
  if (hue_cycles.empty() || hue_period < 1) return false;
  for (const auto& c : hue_cycles)
    if (c.values.empty() || c.period < 1) return false;
  if (sat_cycle.values.empty() || sat_cycle.period < 1) return false;
  for (const auto& w : lum_waves)
    if (w.period < 1) return false;
  return true;
}

// multiwave.cpp:75-116
std::vector<Rgb> MultiWave::cache(int N) const {
  std::vector<Rgb> pal((size_t)(N > 0 ? N : 0));
  for (int i = 0; i < N; i++) {
    const float sat = sat_cycle.value(i);
    float lum = 0.0f;
    for (const auto& w : lum_waves) lum += w.value(i);
    lum = (float)(1.0 / (1.0 + std::exp(-(double)lum)));  // logistic curve for the sum
    // ty interpolates between two hue cycles
    float ty = cycle_pos(hue_cycles.size(), i, hue_period);
    const int y0 = (int)ty;
    const int y1 = (int)((size_t)(y0 + 1) % hue_cycles.size());
    ty -= (float)y0;
    Rgb first;
    for (int pass = 0; pass < 2; pass++) {
      const FloatCycle& cy = hue_cycles[pass ? y1 : y0];
      float tx = cycle_pos(cy.values.size(), i, cy.period);
      const int x0 = (int)tx;
      const int x1 = (int)((size_t)(x0 + 1) % cy.values.size());
      tx -= (float)x0;
      const Rgb c0 = hsl_to_rgb(cy.values[x0], sat, lum);
      const Rgb c1 = hsl_to_rgb(cy.values[x1], sat, lum);
      if (pass == 0) first = interp(c0, c1, tx);
      else pal[i] = interp(first, interp(c0, c1, tx), ty);
    }
  }
  return pal;
}

// multiwave.cpp:19-48: whitespace-separated numbers
bool MultiWave::load(const std::string& text) {
  std::istringstream in(text);
  MultiWave w;
  int n = 0;
  if (!(in >> n) || n < 0 || n > (1 << 20)) return false;
  w.hue_cycles.resize(n);
  for (auto& c : w.hue_cycles) {
    int m = 0;
    if (!(in >> m) || m < 0 || m > (1 << 20)) return false;
    c.values.resize(m);
    for (auto& v : c.values)
      if (!(in >> v)) return false;
    if (!(in >> c.period)) return false;
  }
  if (!(in >> w.hue_period)) return false;
  if (!(in >> n) || n < 0 || n > (1 << 20)) return false;
  w.sat_cycle.values.resize(n);
  for (auto& v : w.sat_cycle.values)
    if (!(in >> v)) return false;
  if (!(in >> w.sat_cycle.period)) return false;
  if (!(in >> n) || n < 0 || n > (1 << 20)) return false;
  w.lum_waves.resize(n);
  for (auto& lw : w.lum_waves)
    if (!(in >> lw.amplitude >> lw.period)) return false;
  *this = w;
  return true;
}

// multiwave.cpp:50-73, the same layout and printf formats
std::string MultiWave::save() const {
  std::string s;
  char buf[64];
  auto put = [&](const char* fmt, auto v) {
    std::snprintf(buf, sizeof buf, fmt, v);
    s += buf;
  };
  put("%d\n", (int)hue_cycles.size());
  for (const auto& c : hue_cycles) {
    put("%d ", (int)c.values.size());
    for (float v : c.values) put("%f ", (double)v);
    put("%d\n", c.period);
  }
  put("%d\n\n", hue_period);
  put("%d\n", (int)sat_cycle.values.size());
  for (float v : sat_cycle.values) put("%f ", (double)v);
  put("\n%d\n\n", sat_cycle.period);
  put("%d\n", (int)lum_waves.size());
  for (const auto& lw : lum_waves) {
    put("%f ", (double)lw.amplitude);
    put("%d\n", lw.period);
  }
  return s;
}

}  // namespace nm
