// palette.hpp - newman's multi-wave palette (multiwave.h:7-37) and the colour
// helpers it needs: an N-entry RGB palette from hue cycles, a saturation cycle
// and logistic-summed luminance waves, with the text .pal format.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nm {

// RGB bytes of a colour
struct Rgb {
  uint8_t r = 0, g = 0, b = 0;
};

// clamp to [0, 255] and truncate to a byte; NaN maps to 0
uint8_t clip_byte(float v);
// HSL -> RGB: h in degrees (any real, taken mod 360), s and l in [0, 1]
Rgb hsl_to_rgb(float h, float s, float l);
// (1 - t) a + t b per channel, then clip_byte
Rgb interp(Rgb a, Rgb b, float t);

class MultiWave {
 public:
  struct FloatCycle {  // multiwave.h:10-16
    std::vector<float> values;
    int period = 1;
    float value(int step) const;  // multiwave.cpp:5-12
  };
  struct FloatWave {  // multiwave.h:18-24
    float amplitude = 1.0f;
    int period = 1;
    float value(int step) const;  // multiwave.cpp:14-17
  };

  std::vector<FloatCycle> hue_cycles;
  int hue_period = 1;
  FloatCycle sat_cycle;
  std::vector<FloatWave> lum_waves;

  // multiwave.cpp:19-48 / 50-73; load returns false on a missing or short file
  bool load(const std::string& text);
  std::string save() const;
  // the parameters are usable by cache(): every cycle non-empty, periods >= 1
  bool valid() const;
  // multiwave.cpp:75-116: palette entries 0 .. N-1
  std::vector<Rgb> cache(int N) const;
};

}  // namespace nm
