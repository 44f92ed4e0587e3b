// colour.hip - the viewer's colouring of an escape grid on the device
// (viewer.cpp:89-122): palette lookup with the smoothing interpolation, and the
// sc x sc colour average of a multisampled grid, into interleaved RGB bytes.
//
// One thread per output pixel; consecutive threads take consecutive output
// columns, so the sc samples a thread reads per grid row sit next to its
// neighbours' (a wave reads 64 sc contiguous escape values per grid row).  The
// kernel is HBM-bound: 8 sc^2 bytes in and 3 bytes out per output pixel; the
// palette (4 bytes per entry) stays in L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "colour.hpp"

namespace nm {

namespace {

// the same conventions as clip_byte / interp in palette.cpp: clamp and truncate
// (0 for v <= 0 or NaN, 255 for v >= 255; fmaxf returns the operand that is not
// NaN), without branches
__device__ __forceinline__ uint32_t clip_u8(float v) { return (uint32_t)fminf(fmaxf(v, 0.0f), 255.0f); }

// getColor (viewer.cpp:89-95): black at or beyond N; the palette entry; or with
// smoothing interp(pal[it - 1], pal[it], smoothing).  pal[-1] (iterations 0 with
// smoothing) lies outside the reference's palette: entry 0 stands in.
__device__ __forceinline__ void get_color(const EscapeValue e, const int N, const uchar4* pal, const int smooth,
                                          uint32_t& r, uint32_t& g, uint32_t& b) {
  if (e.iterations >= N) {
    r = g = b = 0u;
    return;
  }
  const uchar4 c1 = pal[e.iterations];
  if (!smooth) {
    r = c1.x;
    g = c1.y;
    b = c1.z;
    return;
  }
  const uchar4 c0 = pal[e.iterations > 0 ? e.iterations - 1 : 0];
  const float t = e.smoothing, u = 1.0f - t;
  r = clip_u8(u * (float)c0.x + t * (float)c1.x);
  g = clip_u8(u * (float)c0.y + t * (float)c1.y);
  b = clip_u8(u * (float)c0.z + t * (float)c1.z);
}

// SC > 0: the sample count per side is a compile-time constant, so all SC^2
// escape values are loaded before the palette lookups that depend on them;
// SC = 0: any sc at run time
template <int SC>
__global__ __launch_bounds__(256) void k_colour(const EscapeValue* grid, int nc, int sc_rt, int nro, int nco, int N,
                                                const uchar4* pal, int smooth, uint8_t* out) {
  const int sc = SC > 0 ? SC : sc_rt;
  const uint64_t n = (uint64_t)nro * nco;
  const float area = (float)(sc * sc);
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (uint64_t)gridDim.x * blockDim.x) {
    const int r = (int)(o / (uint64_t)nco), c = (int)(o - (uint64_t)r * nco);
    uint32_t cr, cg, cb;
    if (sc == 1) {
      get_color(grid[(uint64_t)r * nc + c], N, pal, smooth, cr, cg, cb);
    } else if (SC > 0) {
      // colorLine (viewer.cpp:110-121): float sums in row-then-column order,
      // divided by sc * sc, then clipped
      EscapeValue e[SC > 0 ? SC * SC : 1];
#pragma unroll
      for (int i = 0; i < SC; i++)
#pragma unroll
        for (int j = 0; j < SC; j++) e[i * SC + j] = grid[(uint64_t)(r * SC + i) * nc + (uint64_t)c * SC + j];
      float sr = 0.0f, sg = 0.0f, sb = 0.0f;
#pragma unroll
      for (int k = 0; k < SC * SC; k++) {
        uint32_t xr, xg, xb;
        get_color(e[k], N, pal, smooth, xr, xg, xb);
        sr += (float)xr;
        sg += (float)xg;
        sb += (float)xb;
      }
      cr = clip_u8(sr / area);
      cg = clip_u8(sg / area);
      cb = clip_u8(sb / area);
    } else {
      float sr = 0.0f, sg = 0.0f, sb = 0.0f;
      for (int r1 = r * sc; r1 < (r + 1) * sc; r1++)
        for (int c1 = c * sc; c1 < (c + 1) * sc; c1++) {
          uint32_t xr, xg, xb;
          get_color(grid[(uint64_t)r1 * nc + c1], N, pal, smooth, xr, xg, xb);
          sr += (float)xr;
          sg += (float)xg;
          sb += (float)xb;
        }
      cr = clip_u8(sr / area);
      cg = clip_u8(sg / area);
      cb = clip_u8(sb / area);
    }
    uint8_t* p = out + 3 * o;
    p[0] = (uint8_t)cr;
    p[1] = (uint8_t)cg;
    p[2] = (uint8_t)cb;
  }
}

}  // namespace

int colour_launch(hipStream_t stream, int num_cu, const EscapeValue* grid, int nr, int nc, int sc, int N,
                  const void* pal_rgba, int smooth, uint8_t* out) {
  const int nro = nr / sc, nco = nc / sc;
  const uint64_t n = (uint64_t)nro * nco;
  if (n == 0) return 0;
  const uint64_t want = (n + 255) / 256;
  const int grid_dim = (int)(want < (uint64_t)num_cu * 16 ? want : (uint64_t)num_cu * 16);
  const uchar4* pal = (const uchar4*)pal_rgba;
  switch (sc) {
    case 1: hipLaunchKernelGGL(k_colour<1>, dim3(grid_dim), dim3(256), 0, stream, grid, nc, sc, nro, nco, N, pal,
                               smooth, out); break;
    case 2: hipLaunchKernelGGL(k_colour<2>, dim3(grid_dim), dim3(256), 0, stream, grid, nc, sc, nro, nco, N, pal,
                               smooth, out); break;
    case 3: hipLaunchKernelGGL(k_colour<3>, dim3(grid_dim), dim3(256), 0, stream, grid, nc, sc, nro, nco, N, pal,
                               smooth, out); break;
    case 4: hipLaunchKernelGGL(k_colour<4>, dim3(grid_dim), dim3(256), 0, stream, grid, nc, sc, nro, nco, N, pal,
                               smooth, out); break;
    default: hipLaunchKernelGGL(k_colour<0>, dim3(grid_dim), dim3(256), 0, stream, grid, nc, sc, nro, nco, N, pal,
                                smooth, out);
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace nm
