// video.hpp - the zoom video's in-between frames (video.cpp:14-34): per frame, the
// previous keyframe scaled by lg and blitted centred, the next keyframe scaled by
// sm and blended over it with weight t.
#pragma once
#include <cstdint>
#include <vector>

namespace nm {

// One in-between frame: the scaled sizes and centring offsets of the previous
// ("large") and next ("small") keyframes, and the blend weight.
struct VideoFrame {
  int lr, lc, lr0, lc0;  // large: rows, cols, offset in the canvas (may be negative)
  int sr, sc, sr0, sc0;  // small: the same
  float t;
  float lry, lrx, sry, srx;  // source pixels per scaled pixel: kr / lr, kc / lc, kr / sr, kc / sc
};

// The schedule of VideoZoom::nextFrame (video.cpp:16-33) in its float
// arithmetic: v = 1.5^(1/rate), sm from 2/3 and lg from 1, both times v per frame,
// t = i / rate; sizes truncate (int)(n * s), offsets (canvas - size) / 2.
std::vector<VideoFrame> video_schedule(int kr, int kc, int nr, int nc, int rate);

// Row pitch, in pixels, of a keyframe widened to 4 bytes per pixel: one padding
// pixel past the last column, rounded up to 4 pixels (16-byte aligned rows).
uint32_t widen_pitch(int kc);

// Device scratch video_launch needs for the two widened keyframes.
uint64_t video_scratch_bytes(int kr, int kc);

// Renders the frames of `sched` on the device from keyframes prev and next
// (kr x kc x 3 bytes each, interleaved RGB, on the device) into out (frames x nr x
// nc x 3, on the device); prev and next 4-byte aligned; rgba:
// video_scratch_bytes(kr, kc) bytes of 16-byte aligned device scratch for the
// widened keyframes.  Defined in video.hip; 0 when the launches were accepted.
int video_launch(void* stream, const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc,
                 const VideoFrame* sched_dev, int frames, uint8_t* rgba, uint8_t* out);

}  // namespace nm
