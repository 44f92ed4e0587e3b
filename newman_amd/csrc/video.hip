// video.hip - the zoom video's in-between frames on the device (video.cpp:14-34):
// for each frame of the schedule, the previous keyframe scaled by lg and blitted
// centred (it covers the canvas), then the next keyframe scaled by sm and blended
// over it with weight t.  One thread per output pixel of every frame (grid: column
// blocks x rows x frames, no index division); the scaled images are never
// materialised: each output pixel samples the two keyframes, widened once to 4
// bytes per pixel so a bilinear tap is one 32-bit load, and they stay in L2 (the
// output, frames x canvas x 3 bytes, is the HBM traffic).
//
// byteimage's scaled() and blend() are not in the reference: scaling is bilinear
// with pixel-centre alignment and clamped edges, rounded to the nearest byte;
// blending clamps and truncates (1 - t) canvas + t small, like the colour
// interpolation.  Both are restated identically in the oracle.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "video.hpp"

namespace nm {

namespace {

__device__ __forceinline__ uint32_t clip_u8(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return (uint32_t)(int)v;
}

// pixel (i, j) of src (sr x sc pixels of 4 bytes) scaled by the source-per-scaled
// pixel ratios ry, rx, all three channels
__device__ __forceinline__ void scaled_px(const uchar4* src, int sr, int sc, float ry, float rx, int i, int j,
                                          float out[3]) {
  float y = ((float)i + 0.5f) * ry - 0.5f;
  float x = ((float)j + 0.5f) * rx - 0.5f;
  y = fminf(fmaxf(y, 0.0f), (float)(sr - 1));
  x = fminf(fmaxf(x, 0.0f), (float)(sc - 1));
  const int y0 = (int)y, x0 = (int)x;
  const int y1 = min(y0 + 1, sr - 1), x1 = min(x0 + 1, sc - 1);
  const float fy = y - (float)y0, fx = x - (float)x0;
  const uchar4 a = src[(size_t)y0 * sc + x0], b = src[(size_t)y0 * sc + x1];
  const uchar4 c = src[(size_t)y1 * sc + x0], d = src[(size_t)y1 * sc + x1];
  const float av[3] = {(float)a.x, (float)a.y, (float)a.z}, bv[3] = {(float)b.x, (float)b.y, (float)b.z};
  const float cv[3] = {(float)c.x, (float)c.y, (float)c.z}, dv[3] = {(float)d.x, (float)d.y, (float)d.z};
  for (int k = 0; k < 3; k++) {
    const float top = (1.0f - fx) * av[k] + fx * bv[k];
    const float bot = (1.0f - fx) * cv[k] + fx * dv[k];
    out[k] = (float)clip_u8((1.0f - fy) * top + fy * bot + 0.5f);
  }
}

// the two keyframes, 3 bytes per pixel, widened to 4 (one 32-bit load per tap)
__global__ __launch_bounds__(256) void k_widen(const uint8_t* a, const uint8_t* b, uint64_t n, uchar4* wa,
                                               uchar4* wb) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    wa[i] = make_uchar4(a[3 * i], a[3 * i + 1], a[3 * i + 2], 0);
    wb[i] = make_uchar4(b[3 * i], b[3 * i + 1], b[3 * i + 2], 0);
  }
}

// grid (column blocks of 256, rows): one thread per canvas pixel, for every frame
// in turn, so a workgroup's keyframe taps of consecutive frames (nearby, the
// scales change by 1.5^(1/rate)) hit L2
__global__ __launch_bounds__(256) void k_video_frames(const uchar4* prev, const uchar4* next, int kr, int kc, int nr,
                                                      int nc, const VideoFrame* sched, int frames, uint8_t* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= nc) return;
  for (int f = 0; f < frames; f++) {
    const VideoFrame s = sched[f];
    float px[3] = {0.0f, 0.0f, 0.0f};
    // blit(large, lr0, lc0): the canvas pixel is the large image's (r - lr0, c - lc0)
    const int li = r - s.lr0, lj = c - s.lc0;
    if (li >= 0 && li < s.lr && lj >= 0 && lj < s.lc) scaled_px(prev, kr, kc, s.lry, s.lrx, li, lj, px);
    // blend(small, t, sr0, sc0)
    const int si = r - s.sr0, sj = c - s.sc0;
    if (si >= 0 && si < s.sr && sj >= 0 && sj < s.sc) {
      float q[3];
      scaled_px(next, kr, kc, s.sry, s.srx, si, sj, q);
      const float u = 1.0f - s.t;
      for (int k = 0; k < 3; k++) px[k] = (float)clip_u8(u * px[k] + s.t * q[k]);
    }
    uint8_t* o = out + 3 * (((size_t)f * nr + r) * nc + c);
    o[0] = (uint8_t)px[0];
    o[1] = (uint8_t)px[1];
    o[2] = (uint8_t)px[2];
  }
}

}  // namespace

int video_launch(void* stream, const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc,
                 const VideoFrame* sched_dev, int frames, uint8_t* rgba, uint8_t* out) {
  if ((uint64_t)nr * nc * (uint64_t)frames == 0) return 0;
  const uint64_t nk = (uint64_t)kr * kc;
  uchar4* wa = (uchar4*)rgba;
  uchar4* wb = wa + nk;
  const uint64_t wantk = (nk + 255) / 256;
  hipLaunchKernelGGL(k_widen, dim3((unsigned)(wantk < 4096 ? wantk : 4096)), dim3(256), 0, (hipStream_t)stream, prev,
                     next, nk, wa, wb);
  if (hipGetLastError() != hipSuccess) return 1;
  hipLaunchKernelGGL(k_video_frames, dim3((nc + 255) / 256, nr), dim3(256), 0, (hipStream_t)stream, wa, wb, kr, kc,
                     nr, nc, sched_dev, frames, out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace nm
