// sample_index.hpp - how a render's sample ids map to frames, pixels and table
// entries: the frame records, the sample orders (row-major, or T x T tiles of
// 8 x 8 micro-blocks) and the start kernels' tiling of the sample ids.  Every
// function here is compiled for the device (the kernels use it) and for the host
// (the engine builds the frame records with it, and a host-side checker walks
// every thread of a launch through it, with bounds checks).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace nm {

constexpr int kThreads = 256;           // workgroup size of every kernel
constexpr int kPerThread = 4;           // samples per thread per start tile
constexpr int kTile = kThreads * kPerThread;

struct DevFrame {
  uint32_t sample_base;  // first global sample id of this frame
  uint32_t nr, nc;
  uint32_t tab_c, tab_r;  // offsets into the column / row tables
  int32_t card_mode;
  uint64_t out_base;   // first output element (row-major grid of this frame)
  uint64_t mask_base;  // first bit in the cardioid mask (row-major)
  // sample order: tile == 0 is row-major; otherwise T x T tiles (row-major tile
  // grid, edge tiles padded) of 8 x 8 micro-blocks, one wave per micro-block
  uint32_t tile, ntc;     // T, tiles per frame row
  uint32_t tshift, mshift;  // log2(T*T), log2(T/8)
  double inv_nc, inv_ntc;
};

// x / d for x < 2^32 with inv = 1.0 / d: (x + 1/2) / d is at least 1/(2d) away
// from an integer and the product errs by at most (x / d) 2^-52 < 1/(2d), so the
// truncated product is the exact quotient
__host__ __device__ inline uint32_t udiv(uint32_t x, double inv) { return (uint32_t)(((double)x + 0.5) * inv); }

// every frame of a frame set has the same size: frame = sample id / frame size
__host__ __device__ inline uint32_t frame_of(uint32_t gid, double inv_frame) { return udiv(gid, inv_frame); }

// the (row, column) of a frame-local sample id; padding samples of edge tiles
// are not valid
struct Loc {
  uint32_t r, c;
  bool valid;
};
__host__ __device__ inline Loc locate(const DevFrame& f, uint32_t local) {
  Loc l;
  if (f.tile == 0) {
    l.r = udiv(local, f.inv_nc);
    l.c = local - l.r * f.nc;
    l.valid = true;
    return l;
  }
  const uint32_t t = local >> f.tshift, w = local & ((1u << f.tshift) - 1u);
  const uint32_t tr = udiv(t, f.inv_ntc), tc = t - tr * f.ntc;
  const uint32_t mb = w >> 6, in = w & 63u;
  l.r = tr * f.tile + ((mb >> f.mshift) << 3) + (in >> 3);
  l.c = tc * f.tile + ((mb & ((1u << f.mshift) - 1u)) << 3) + (in & 7u);
  l.valid = l.r < f.nr && l.c < f.nc;
  return l;
}

// The start kernels: workgroup b takes the 1024-sample tiles b, b + grid, ...;
// thread t handles samples tile * 1024 + j * 256 + t, j = 0 .. 3 (ids at or past
// the total are idle lanes)
__host__ __device__ inline uint32_t start_gid(uint32_t tile, int j, uint32_t tid) {
  return tile * kTile + (uint32_t)j * kThreads + tid;
}
__host__ __device__ inline uint32_t start_tiles(uint32_t total) { return (total + kTile - 1) / kTile; }
// the frames of a tile's first and last sample (equal when one frame holds the
// whole tile); tile < start_tiles(total)
__host__ __device__ inline void tile_frame_span(uint32_t tile, uint32_t total, double inv_frame, uint32_t* f0,
                                                uint32_t* f1) {
  const uint32_t a = tile * kTile, b = (a + kTile < total ? a + kTile : total) - 1;
  *f0 = frame_of(a, inv_frame);
  *f1 = frame_of(b, inv_frame);
}

// The start kernels' hand-on bits: one 64-bit word per wave and sample slot (64
// consecutive sample ids from a multiple of 64; bit l = lane l's sample goes to
// the iterate kernel), written by every wave of every tile, so no clearing is
// needed; a tile owns words 16 t .. 16 t + 15.
constexpr int kHandWordsPerTile = kTile / 64;
__host__ __device__ inline uint32_t hand_words(uint32_t total) { return start_tiles(total) * kHandWordsPerTile; }
__host__ __device__ inline uint32_t hand_word(uint32_t gid) { return gid >> 6; }
// k_scatter: the active-list position of bit l of word k of a tile, given the
// tile's offset (the exclusive scan of the tile counts) and the set bits of the
// tile's earlier words (before); sample ids keep their order
__host__ __device__ inline uint32_t hand_position(uint32_t tile_off, uint32_t before, uint64_t word, uint32_t l) {
  const uint64_t lower = l ? word & (~0ull >> (64 - l)) : 0ull;
  return tile_off + before + (uint32_t)__builtin_popcountll(lower);
}

// k_scatter on a sparse tile (at most kScatterSparse marked samples): lane j
// writes active-list entries off + j, off + j + 64, ...; entry p is bit
// select_bit(w_k, p - before_k) of the last word k with before_k <= p (the
// words' exclusive prefix of set bits), which is never an empty word
constexpr uint32_t kScatterSparse = 256;
// the position of set bit r (0-based) of w; r < popcount(w)
__host__ __device__ inline uint32_t select_bit(uint64_t w, uint32_t r) {
  uint32_t pos = 0;
  for (uint32_t width = 32; width >= 1; width >>= 1) {
    const uint32_t c = (uint32_t)__builtin_popcountll(w & ((1ull << width) - 1ull));
    if (r >= c) {
      r -= c;
      w >>= width;
      pos += width;
    }
  }
  return pos;
}

// ---- the exclusive scan of the per-tile counts: rounds of kScanRound counts
// (one 1024-thread workgroup pass), `per` consecutive rounds to each of `groups`
// workgroups (at most kScanGroups); group g covers counts [g per kScanRound,
// min((g + 1) per kScanRound, ntiles)), and every group has at least one count
// when ntiles > 0 (ntiles = 0: one group, no counts)
constexpr uint32_t kScanRound = 16384;
constexpr uint32_t kScanGroups = 64;
struct ScanSplit {
  uint32_t per, groups;
};
__host__ __device__ inline ScanSplit scan_split(uint32_t ntiles) {
  const uint32_t rounds = ntiles ? (ntiles - 1) / kScanRound + 1 : 1u;
  const uint32_t per = (rounds - 1) / kScanGroups + 1;
  return {per, (rounds - 1) / per + 1};
}
__host__ __device__ inline void scan_group_range(uint32_t ntiles, uint32_t per, uint32_t g, uint32_t& lo,
                                                 uint32_t& hi) {
  const uint64_t span = (uint64_t)per * kScanRound, n = ntiles;
  const uint64_t l = (uint64_t)g * span;
  lo = (uint32_t)(l < n ? l : n);
  hi = (uint32_t)(l + span < n ? l + span : n);
}

// ---- the active list (sample ids that iterate) as the iterate kernels read it
constexpr int kChunk = 64;   // active samples per chunk (one per lane)
constexpr int kQueues = 8;   // claim ranges of the chunks
__host__ __device__ inline uint32_t chunk_count(uint32_t n) { return (n + kChunk - 1) / kChunk; }
// interleaved chunks: lane l of chunk c takes entry R ((l / R) nch + c) + l % R,
// R = 2^shift, when that is below n (64 / R runs of R consecutive entries, nch
// chunks apart; R = 64 is 64 consecutive entries); the entry grows with the
// lane, so the lanes that take one are a prefix of the wave
__host__ __device__ inline uint32_t chunk_pos_run(uint32_t lane, uint32_t c, uint32_t nch, uint32_t shift) {
  return (((lane >> shift) * nch + c) << shift) + (lane & ((1u << shift) - 1u));
}
// the plain iterate kernel's chunks: 64 entries nch apart (runs of 1)
__host__ __device__ inline uint32_t chunk_pos(uint32_t lane, uint32_t c, uint32_t nch) {
  return chunk_pos_run(lane, c, nch, 0);
}
// the chunks of claim range h: [claim_lo(nch, h), claim_lo(nch, h + 1))
__host__ __device__ inline uint32_t claim_lo(uint32_t nch, uint32_t h) {
  return (uint32_t)((uint64_t)nch * h / kQueues);
}

// samples per frame in the sample order (tiles pad the frame); 0 = bad tile
inline uint64_t frame_samples(int nr, int nc, int tile) {
  if (tile == 0) return (uint64_t)nr * nc;
  if (tile != 8 && tile != 16 && tile != 32 && tile != 64 && tile != 128 && tile != 256) return 0;
  uint64_t ntr = (uint64_t)(nr + tile - 1) / tile, ntc = (uint64_t)(nc + tile - 1) / tile;
  return ntr * ntc * (uint64_t)tile * tile;
}

// DevFrame for frame k of a set: sample order and bases (table offsets and the
// cardioid class are the caller's)
inline DevFrame make_frame(uint64_t k, int nr, int nc, int tile, uint32_t frame_size) {
  DevFrame d{};
  d.sample_base = (uint32_t)(k * frame_size);
  d.nr = (uint32_t)nr;
  d.nc = (uint32_t)nc;
  d.out_base = k * (uint64_t)nr * nc;
  d.inv_nc = 1.0 / (double)nc;
  d.tile = (uint32_t)tile;
  if (tile) {
    d.ntc = (uint32_t)((nc + tile - 1) / tile);
    d.inv_ntc = 1.0 / (double)d.ntc;
    int lt = 0;
    while ((1 << lt) < tile) lt++;
    d.tshift = (uint32_t)(2 * lt);
    d.mshift = (uint32_t)(lt - 3);
  }
  return d;
}

// workgroups of a start launch: one per tile up to 16 per CU and up to the
// counter slots the start kernels own (slot_cap)
inline int start_grid(uint64_t total, int num_cu, uint64_t slot_cap) {
  const uint64_t ntiles = (total + kTile - 1) / kTile;
  const uint64_t cap = std::min<uint64_t>((uint64_t)num_cu * 16, slot_cap);
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(ntiles, cap));
}

}  // namespace nm
