// video.hip - the zoom video's in-between frames on the device (video.cpp:14-34):
// for each frame of the schedule, the previous keyframe scaled by lg and blitted
// centred (it covers the canvas), then the next keyframe scaled by sm and blended
// over it with weight t.  One thread per 4 output pixels of every frame (grid:
// column blocks x rows, frames in a loop, no index division); the scaled images are never
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

// 0 for v <= 0 or NaN, 255 for v >= 255, else v truncated (fmaxf returns the
// operand that is not NaN)
__device__ __forceinline__ uint32_t clip_u8(float v) { return (uint32_t)fminf(fmaxf(v, 0.0f), 255.0f); }

// 8 bytes (two widened pixels) from a 4-byte aligned address: one dwordx2 load
__device__ __forceinline__ uint2 load_pair(const uint32_t* p) {
  uint2 v;
  __builtin_memcpy(&v, p, 8);
  return v;
}

__device__ __forceinline__ float chan(uint32_t px, int k) { return (float)((px >> (8 * k)) & 0xFFu); }

// pixel (i, j) of src (sr x sc pixels of 4 bytes, rows `pitch` pixels apart, one
// padding pixel past each row's end) scaled by the source-per-scaled pixel ratios
// ry, rx, all three channels.  The taps (x0, x0 + 1) of a row are one 8-byte
// load: x0 + 1 <= sc lies inside the padded row, and when x0 = sc - 1 (x clamped
// to the last column) fx = 0, so the padding pixel's weight is 0 and the sum is
// the same as with the reference's clamped tap x1 = sc - 1.
__device__ __forceinline__ void scaled_px(const uint32_t* src, int sr, int sc, uint32_t pitch, float ry, float rx,
                                          int i, int j, float out[3]) {
  float y = ((float)i + 0.5f) * ry - 0.5f;
  float x = ((float)j + 0.5f) * rx - 0.5f;
  y = fminf(fmaxf(y, 0.0f), (float)(sr - 1));
  x = fminf(fmaxf(x, 0.0f), (float)(sc - 1));
  const int y0 = (int)y, x0 = (int)x;
  const int y1 = min(y0 + 1, sr - 1);
  const float fy = y - (float)y0, fx = x - (float)x0;
  const uint2 ab = load_pair(src + (size_t)y0 * pitch + x0);
  const uint2 cd = load_pair(src + (size_t)y1 * pitch + x0);
  float av[3], bv[3], cv[3], dv[3];
  for (int k = 0; k < 3; k++) {
    av[k] = chan(ab.x, k);
    bv[k] = chan(ab.y, k);
    cv[k] = chan(cd.x, k);
    dv[k] = chan(cd.y, k);
  }
  for (int k = 0; k < 3; k++) {
    const float top = (1.0f - fx) * av[k] + fx * bv[k];
    const float bot = (1.0f - fx) * cv[k] + fx * dv[k];
    out[k] = (float)clip_u8((1.0f - fy) * top + fy * bot + 0.5f);
  }
}

// the two keyframes, 3 bytes per pixel, widened to 4 (taps are 4-byte loads)
// into rows of widen_pitch(kc) pixels: pixels kc .. pitch - 1 of a row repeat its
// last pixel.  Grid (pitch / 4 / 256 column blocks, kr rows): a thread writes 4
// pixels of a row (one 16-byte store; rows are 16-byte aligned) from 12 source
// bytes read as three aligned dwords when kc is a multiple of 4.
__device__ __forceinline__ uint32_t rgb_at(const uint8_t* row, int x) {
  return (uint32_t)row[3 * x] | (uint32_t)row[3 * x + 1] << 8 | (uint32_t)row[3 * x + 2] << 16;
}

__device__ __forceinline__ void widen_row4(const uint8_t* src, uint32_t* dst, int kc, uint32_t pitch, int y, int x0) {
  const uint8_t* row = src + (size_t)y * kc * 3;
  uint4 o;
  if ((kc & 3) == 0 && x0 + 4 <= kc) {
    const uint32_t* s = (const uint32_t*)(row + 3 * x0);  // little-endian: byte 3i is pixel i's low byte
    const uint32_t w0 = s[0], w1 = s[1], w2 = s[2];
    o.x = w0 & 0xFFFFFFu;
    o.y = (w0 >> 24) | ((w1 & 0xFFFFu) << 8);
    o.z = (w1 >> 16) | ((w2 & 0xFFu) << 16);
    o.w = w2 >> 8;
  } else {
    o.x = rgb_at(row, min(x0, kc - 1));
    o.y = rgb_at(row, min(x0 + 1, kc - 1));
    o.z = rgb_at(row, min(x0 + 2, kc - 1));
    o.w = rgb_at(row, min(x0 + 3, kc - 1));
  }
  *(uint4*)(dst + (size_t)y * pitch + x0) = o;
}

__global__ __launch_bounds__(256) void k_widen(const uint8_t* a, const uint8_t* b, int kr, int kc, uint32_t pitch,
                                               uint32_t* wa, uint32_t* wb) {
  const int x0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x), y = blockIdx.y;
  if ((uint32_t)x0 >= pitch || y >= kr) return;
  widen_row4(a, wa, kc, pitch, y, x0);
  widen_row4(b, wb, kc, pitch, y, x0);
}

// grid (column blocks of 1024, rows): one thread per 4 consecutive canvas
// pixels, for every frame in turn, so a workgroup's keyframe taps of consecutive
// frames (nearby, the scales change by 1.5^(1/rate)) hit L2; with the canvas
// width a multiple of 4 the 12 bytes of a thread's pixels are one aligned
// 3-dword store
__global__ __launch_bounds__(256) void k_video_frames(const uint32_t* prev, const uint32_t* next, int kr, int kc,
                                                      uint32_t pitch, int nr, int nc, const VideoFrame* sched,
                                                      int frames, uint8_t* out) {
  const int c0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x), r = blockIdx.y;
  if (c0 >= nc) return;
  const int n = min(4, nc - c0);
  const bool packed = (nc & 3) == 0;
  for (int f = 0; f < frames; f++) {
    const VideoFrame s = sched[f];
    uint32_t b[12];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int c = c0 + q;
      float px[3] = {0.0f, 0.0f, 0.0f};
      if (q < n) {
        // blit(large, lr0, lc0): the canvas pixel is the large image's (r - lr0, c - lc0)
        const int li = r - s.lr0, lj = c - s.lc0;
        if (li >= 0 && li < s.lr && lj >= 0 && lj < s.lc) scaled_px(prev, kr, kc, pitch, s.lry, s.lrx, li, lj, px);
        // blend(small, t, sr0, sc0)
        const int si = r - s.sr0, sj = c - s.sc0;
        if (si >= 0 && si < s.sr && sj >= 0 && sj < s.sc) {
          float qv[3];
          scaled_px(next, kr, kc, pitch, s.sry, s.srx, si, sj, qv);
          const float u = 1.0f - s.t;
          for (int k = 0; k < 3; k++) px[k] = (float)clip_u8(u * px[k] + s.t * qv[k]);
        }
      }
      for (int k = 0; k < 3; k++) b[3 * q + k] = (uint32_t)px[k];
    }
    uint8_t* o = out + 3 * (((size_t)f * nr + r) * nc + c0);
    if (packed) {
      uint3 w;
      w.x = b[0] | b[1] << 8 | b[2] << 16 | b[3] << 24;
      w.y = b[4] | b[5] << 8 | b[6] << 16 | b[7] << 24;
      w.z = b[8] | b[9] << 8 | b[10] << 16 | b[11] << 24;
      *(uint3*)o = w;
    } else {
      for (int i = 0; i < 3 * n; i++) o[i] = (uint8_t)b[i];
    }
  }
}

}  // namespace

uint32_t widen_pitch(int kc) { return ((uint32_t)kc + 1u + 3u) & ~3u; }

uint64_t video_scratch_bytes(int kr, int kc) { return 2ull * 4ull * (uint64_t)kr * widen_pitch(kc); }

int video_launch(void* stream, const uint8_t* prev, const uint8_t* next, int kr, int kc, int nr, int nc,
                 const VideoFrame* sched_dev, int frames, uint8_t* rgba, uint8_t* out) {
  if ((uint64_t)nr * nc * (uint64_t)frames == 0) return 0;
  if (kr <= 0 || kc <= 0) return 1;
  const uint32_t pitch = widen_pitch(kc);
  uint32_t* wa = (uint32_t*)rgba;
  uint32_t* wb = wa + (uint64_t)kr * pitch;  // 16-byte aligned like wa: pitch is a multiple of 4
  hipLaunchKernelGGL(k_widen, dim3((pitch / 4 + 255) / 256, kr), dim3(256), 0, (hipStream_t)stream, prev, next, kr,
                     kc, pitch, wa, wb);
  if (hipGetLastError() != hipSuccess) return 1;
  hipLaunchKernelGGL(k_video_frames, dim3((nc + 1023) / 1024, nr), dim3(256), 0, (hipStream_t)stream, wa, wb, kr, kc,
                     pitch, nr, nc, sched_dev, frames, out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace nm
