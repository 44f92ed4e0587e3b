// video.cpp - the in-between frame schedule of the zoom video (video.cpp:14-34).
#include "video.hpp"

#include <cmath>

namespace nm {

std::vector<VideoFrame> video_schedule(int kr, int kc, int nr, int nc, int rate) {
  std::vector<VideoFrame> s;
  if (rate <= 0) return s;
  // float t, v = pow(1.5, 1.0 / rate), sm = 2.0 / 3.0, lg = 1.0
  const float v = (float)std::pow(1.5, 1.0 / rate);
  float sm = (float)(2.0 / 3.0), lg = 1.0f;
  for (int i = 0; i < rate; i++) {
    VideoFrame f;
    f.t = (float)i / (float)rate;
    f.sr = (int)((float)kr * sm);
    f.sc = (int)((float)kc * sm);
    f.lr = (int)((float)kr * lg);
    f.lc = (int)((float)kc * lg);
    f.lr0 = (nr - f.lr) / 2;
    f.lc0 = (nc - f.lc) / 2;
    f.sr0 = (nr - f.sr) / 2;
    f.sc0 = (nc - f.sc) / 2;
    f.lry = f.lr > 0 ? (float)kr / (float)f.lr : 0.0f;
    f.lrx = f.lc > 0 ? (float)kc / (float)f.lc : 0.0f;
    f.sry = f.sr > 0 ? (float)kr / (float)f.sr : 0.0f;
    f.srx = f.sc > 0 ? (float)kc / (float)f.sc : 0.0f;
    s.push_back(f);
    sm *= v;
    lg *= v;
  }
  return s;
}

}  // namespace nm
