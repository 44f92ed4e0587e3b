// index_check.cpp - a host-side walk of every thread of the render kernels'
// launches through the same index code the kernels run (sample_index.hpp), with
// every index checked against the length of the buffer it addresses (the
// lengths the engine allocates: prepare_common, render_hw, render_sa).
//
// For a frame set (nframes frames of nr x nc in a sample order) it checks:
//  * the start kernels (plain and series): each (workgroup, tile, j, thread) of
//    the launch grid: the tile's frame span, the sample's frame (and that the
//    one-frame shortcut names the sample's own frame), its position, the column /
//    row table entries, the cardioid-mask word, the tile-skip entry, and the
//    per-sample, per-wave (hand-on word) and per-tile stores;
//  * k_scatter: for hand-on words with pseudo-random bits on the valid sample
//    ids, its word loads, and that the positions it writes are exactly
//    0 .. count - 1, each once, in sample order;
//  * k_emit: every sample's output element of the row-major grids, and that the
//    valid samples cover every pixel of every frame exactly once;
//  * the active-list chunks of both iterate kernels (every run length) for a
//    range of list lengths: every entry read below the length, each entry taken
//    once, and the claim ranges partitioning the chunks.
// Test support, not part of the render path: built as its own library and called
// from the CPU tests.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sample_index.hpp"

namespace {

struct Report {
  unsigned long long violations = 0, checked = 0;
  char* msg;
  int cap;
  void fail(const char* fmt, ...) {
    if (violations++ == 0 && msg && cap > 0) {
      va_list ap;
      va_start(ap, fmt);
      std::vsnprintf(msg, (size_t)cap, fmt, ap);
      va_end(ap);
    }
  }
  bool in(uint64_t idx, uint64_t len, const char* what, uint64_t gid) {
    checked++;
    if (idx < len) return true;
    fail("%s index %llu >= length %llu (sample %llu)", what, (unsigned long long)idx, (unsigned long long)len,
         (unsigned long long)gid);
    return false;
  }
};

// the slot caps of the engine's start and iterate launches
constexpr uint64_t kStartSlots = 4096;

}  // namespace

extern "C" {

// Returns the number of violations (0: every index in bounds, every pixel
// covered once); msg gets the first.  mask: frames carry cardioid masks;
// skip_T: tile side of per-tile series skips (0: none).  plant (the checker's
// own negative controls, 0 for none): 1 moves the last frame's column tables
// one entry on, 2 gives the frames one sample more than the sample order has,
// 3 interleaves the plain chunks with one chunk too few, 4 counts a lane's own
// bit into its scatter position, 5 launches the tile scan with one group too few,
// 6 selects the next set bit in the sparse scatter.
long long nm_index_check(int nr, int nc, int nframes, int tile, int num_cu, int mask, int skip_T, int plant,
                         char* msg, int cap, unsigned long long* checked) {
  using namespace nm;
  Report R;
  R.msg = msg;
  R.cap = cap;
  if (msg && cap > 0) msg[0] = 0;
  if (nr <= 0 || nc <= 0 || nframes <= 0 || num_cu <= 0) return -1;
  const uint64_t fs = frame_samples(nr, nc, tile);
  if (fs == 0) return -1;
  const uint64_t total64 = fs * (uint64_t)nframes;
  if (total64 >= 0xFFFFF000ull) return -2;  // the engine refuses it (prepare_common)
  const uint32_t total = (uint32_t)total64;
  // the frame records and buffer lengths as render_hw / render_sa build them
  const uint64_t mask_words = ((uint64_t)nr * nc + 31) / 32;
  std::vector<DevFrame> df;
  for (int k = 0; k < nframes; k++) {
    DevFrame d = make_frame((uint64_t)k, nr, nc, tile, (uint32_t)fs);
    d.tab_c = (uint32_t)((uint64_t)k * nc);
    d.tab_r = (uint32_t)((uint64_t)k * nr);
    d.mask_base = mask ? (uint64_t)k * mask_words * 32 : 0;
    if (plant == 1 && k == nframes - 1) d.tab_c += 1;
    if (plant == 2) d.sample_base = (uint32_t)((uint64_t)k * (fs + 1));
    df.push_back(d);
  }
  const double inv_frame = 1.0 / (double)(plant == 2 ? fs + 1 : fs);
  const uint64_t ncol = (uint64_t)nframes * nc, nrow = (uint64_t)nframes * nr;
  const uint64_t nmask = mask ? (uint64_t)nframes * mask_words : 1;
  const uint64_t nout = (uint64_t)nframes * nr * nc;
  const uint32_t ntiles = start_tiles(total);
  const uint32_t sk_ntc = skip_T ? (uint32_t)((nc + skip_T - 1) / skip_T) : 0;
  const uint32_t sk_tpf = skip_T ? sk_ntc * (uint32_t)((nr + skip_T - 1) / skip_T) : 0;
  const uint64_t nskip = (uint64_t)sk_tpf * nframes;
  std::vector<uint8_t> seen(nout, 0);

  // ---- the start kernels
  const int grid = start_grid(total, num_cu, kStartSlots);
  for (int b = 0; b < grid; b++)
    for (uint32_t t = (uint32_t)b; t < ntiles; t += (uint32_t)grid) {
      R.in(t, ntiles, "tile_count", t);
      uint32_t f0, f1;
      tile_frame_span(t, total, inv_frame, &f0, &f1);
      R.in(f0, (uint64_t)nframes, "frames (tile span)", (uint64_t)t * kTile);
      R.in(f1, (uint64_t)nframes, "frames (tile span)", (uint64_t)t * kTile);
      for (int j = 0; j < kPerThread; j++)
        for (uint32_t tid = 0; tid < (uint32_t)kThreads; tid++) {
          const uint32_t gid = start_gid(t, j, tid);
          if (tid % 64 == 0) {  // lane 0 stores the wave's hand-on word, idle waves too
            if (gid % 64) R.fail("wave of sample %u does not start a hand-on word", gid);
            R.in(hand_word(gid), hand_words(total), "hand-on words", gid);
          }
          if (gid >= total) continue;  // an idle lane: the kernels load and store nothing else for it
          const uint32_t fi = f0 == f1 ? f0 : frame_of(gid, inv_frame);
          if (fi != gid / fs) R.fail("sample %u: frame %u, want %llu", gid, fi, (unsigned long long)(gid / fs));
          if (frame_of(gid, inv_frame) != gid / fs) R.fail("sample %u: frame_of is not the quotient", gid);
          if (!R.in(fi, (uint64_t)nframes, "frames", gid)) continue;
          const DevFrame& f = df[fi];
          if (gid < f.sample_base) {
            R.fail("sample %u below its frame's base %u", gid, f.sample_base);
            continue;
          }
          const uint32_t local = gid - f.sample_base;
          R.in(local, fs, "frame-local sample", gid);
          R.in(gid, total, "out/flag/state/state_i/r2", gid);
          const Loc lc = locate(f, local);
          if (!lc.valid) continue;  // tile padding: only its flag is written
          R.in(lc.r, f.nr, "row", gid);
          R.in(lc.c, f.nc, "column", gid);
          R.in((uint64_t)f.tab_c + lc.c, ncol, "column tables (cre, crelo, epsre)", gid);
          R.in((uint64_t)f.tab_r + lc.r, nrow, "row tables (cim, cimlo, epsim)", gid);
          if (mask) R.in((f.mask_base + (uint64_t)lc.r * f.nc + lc.c) >> 5, nmask, "cardioid mask", gid);
          if (skip_T)
            R.in((uint64_t)fi * sk_tpf + (lc.r / (uint32_t)skip_T) * sk_ntc + lc.c / (uint32_t)skip_T, nskip,
                 "tile skips", gid);
        }
    }

  // ---- k_scatter over hand-on words with pseudo-random bits on valid ids: tiles
  // in turn about 3 in 8, 1 in 16 and 1 in 64 marked, and none (the kernel's
  // dense and sparse paths)
  auto marked = [](uint32_t gid, uint64_t x) {
    static const uint64_t dens[4] = {24, 4, 1, 0};
    return (x & 63) < dens[(gid / kTile) % 4];
  };
  {
    const uint32_t nwords = hand_words(total);
    std::vector<uint64_t> words(nwords, 0);
    uint64_t x = 0x9E3779B97F4A7C15ull ^ total;
    for (uint32_t gid = 0; gid < total; gid++) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      if (marked(gid, x)) words[hand_word(gid)] |= 1ull << (gid % 64);
    }
    std::vector<uint32_t> off(ntiles, 0);
    uint64_t count = 0;
    for (uint32_t t = 0; t < ntiles; t++) {
      off[t] = (uint32_t)count;
      for (int k = 0; k < kHandWordsPerTile; k++) count += (uint64_t)__builtin_popcountll(words[t * kHandWordsPerTile + k]);
    }
    std::vector<int64_t> at(count, -1);
    for (uint32_t t = 0; t < ntiles; t++) {
      uint32_t before = 0;
      for (int k = 0; k < kHandWordsPerTile; k++) {
        const uint64_t wi = (uint64_t)t * kHandWordsPerTile + k;
        if (!R.in(wi, nwords, "hand-on words (scatter)", (uint64_t)t * kTile)) continue;
        const uint64_t w = words[wi];
        for (uint32_t l = 0; l < 64; l++) {
          if (!((w >> l) & 1ull)) continue;
          const uint32_t gid = t * kTile + (uint32_t)k * 64 + l;
          const uint32_t pos = hand_position(off[t], before, w, plant == 4 ? l + 1 : l);
          if (!R.in(pos, count, "active (scatter)", gid)) continue;
          if (at[pos] >= 0) R.fail("active position %u written twice (sample %u)", pos, gid);
          at[pos] = gid;
        }
        before += (uint32_t)__builtin_popcountll(w);
      }
    }
    // the sparse-tile path (tiles with at most kScatterSparse marked samples):
    // lane j of batch p0 takes entry p = p0 + j, the last word k whose prefix
    // is <= p (the kernel's 4-step search) and bit select_bit(w_k, p - before_k)
    for (uint32_t t = 0; t < ntiles; t++) {
      uint32_t pre[kHandWordsPerTile], m = 0;
      for (int k = 0; k < kHandWordsPerTile; k++) {
        pre[k] = m;
        m += (uint32_t)__builtin_popcountll(words[(uint64_t)t * kHandWordsPerTile + k]);
      }
      if (m > kScatterSparse) continue;
      for (uint32_t p0 = 0; p0 < m; p0 += 64)
        for (uint32_t j = 0; j < 64; j++) {
          const uint32_t p = p0 + j;
          int k = 0;
          for (int step = kHandWordsPerTile / 2; step >= 1; step >>= 1)
            if (pre[k + step] <= p) k += step;
          if (p >= m) continue;
          const uint64_t wk = words[(uint64_t)t * kHandWordsPerTile + k];
          const uint32_t gid = t * kTile + (uint32_t)k * 64 + select_bit(wk, p - pre[k] + (plant == 6 ? 1u : 0u));
          if (!R.in((uint64_t)off[t] + p, count, "active (sparse scatter)", gid)) continue;
          if (at[off[t] + p] != (int64_t)gid)
            R.fail("sparse scatter: tile %u entry %u takes sample %u, want %lld", t, p, gid, (long long)at[off[t] + p]);
        }
    }
    // select_bit against a bit-by-bit walk
    {
      uint64_t y = 0x2545F4914F6CDD1Dull ^ total;
      for (int trial = 0; trial < 2000; trial++) {
        y ^= y << 13;
        y ^= y >> 7;
        y ^= y << 17;
        const uint64_t w = trial < 3 ? (trial == 0 ? 1ull : trial == 1 ? ~0ull : 1ull << 63) : y & (y >> (trial % 7));
        uint32_t r = 0;
        for (uint32_t b = 0; b < 64; b++)
          if ((w >> b) & 1ull) {
            if (select_bit(w, r) != b) R.fail("select_bit(%llx, %u) = %u, want %u", (unsigned long long)w, r,
                                              select_bit(w, r), b);
            r++;
          }
      }
    }
    for (uint64_t p = 0; p < count; p++) {
      if (at[p] < 0) {
        R.fail("active position %llu written by no sample", (unsigned long long)p);
        break;
      }
      if (p && at[p] <= at[p - 1]) {
        R.fail("active position %llu out of sample order", (unsigned long long)p);
        break;
      }
    }
  }

  // ---- the tile scan (k_tile_sums, then k_scan_tiles) over the scatter's tile
  // counts: every thread's loads and stores, the groups' carries and the total;
  // plant 5 launches one group too few
  {
    std::vector<uint32_t> cnt(ntiles), seq(ntiles);
    uint64_t acc = 0;
    {  // the marked samples of the scatter check's generator, per tile
      uint64_t x = 0x9E3779B97F4A7C15ull ^ total;
      for (uint32_t gid = 0; gid < total; gid++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        if (marked(gid, x)) cnt[gid / kTile]++;
      }
    }
    for (uint32_t t = 0; t < ntiles; t++) {
      seq[t] = (uint32_t)acc;
      acc += cnt[t];
    }
    const ScanSplit sp = scan_split(ntiles);
    const uint32_t groups = plant == 5 ? sp.groups - 1 : sp.groups;
    if (sp.groups < 1 || sp.groups > kScanGroups) R.fail("tile scan of %u counts: %u groups", ntiles, sp.groups);
    std::vector<uint32_t> gsum(groups, 0), out(ntiles, 0);
    std::vector<uint8_t> wrote(ntiles, 0);
    for (int pass = 0; pass < 2; pass++) {  // 0: k_tile_sums, 1: k_scan_tiles
      for (uint32_t g = 0; g < groups; g++) {
        uint32_t lo, hi;
        scan_group_range(ntiles, sp.per, g, lo, hi);
        uint32_t carry = 0;
        for (uint32_t q = 0; q < g; q++) carry += gsum[q];
        for (uint32_t base = lo; base < hi; base += kScanRound) {
          uint32_t before = 0;  // the counts of the round's earlier threads (the wave scans)
          for (uint32_t tid = 0; tid < 1024; tid++) {
            const uint32_t i0 = base + 16 * tid;
            uint32_t mine = 0;
            for (uint32_t k = 0; k < 16; k++) {
              if (i0 + k >= hi) continue;
              if (!R.in(i0 + k, ntiles, "tile counts (scan)", (uint64_t)g)) continue;
              if (pass == 1) {
                if (wrote[i0 + k]++) R.fail("tile offset %u written twice", i0 + k);
                out[i0 + k] = carry + before + mine;
              }
              mine += cnt[i0 + k];
            }
            before += mine;
          }
          carry += before;
          if (pass == 0) gsum[g] += before;
        }
        if (pass == 1 && g + 1 == groups && carry != (uint32_t)acc)
          R.fail("tile scan total %u, want %llu", carry, (unsigned long long)acc);
      }
    }
    for (uint32_t t = 0; t < ntiles; t++)
      if (!wrote[t] || out[t] != seq[t]) {
        R.fail("tile offset %u: %s %u, want %u", t, wrote[t] ? "wrote" : "not written", out[t], seq[t]);
        break;
      }
    // the split at edge sizes: at most kScanGroups groups of consecutive,
    // non-empty count ranges covering [0, ntiles)
    const uint32_t edge[] = {0u, 1u, kScanRound - 1, kScanRound, kScanRound + 1, kScanGroups * kScanRound,
                             kScanGroups * kScanRound + 1, 100 * kScanRound + 7, start_tiles(0xFFFFEFFFu), ntiles};
    for (uint32_t n : edge) {
      const ScanSplit e = scan_split(n);
      if (e.groups < 1 || e.groups > kScanGroups) R.fail("tile scan of %u counts: %u groups", n, e.groups);
      uint32_t next = 0;
      for (uint32_t g = 0; g < e.groups; g++) {
        uint32_t lo, hi;
        scan_group_range(n, e.per, g, lo, hi);
        if (lo != next || hi < lo || (n > 0 && hi == lo))
          R.fail("tile scan of %u counts: group %u covers [%u, %u) after %u", n, g, lo, hi, next);
        next = hi;
      }
      if (next != n) R.fail("tile scan of %u counts: the groups end at %u", n, next);
    }
  }

  // ---- k_emit: sample order -> the row-major grids
  for (uint32_t gid = 0; gid < total; gid++) {
    const uint32_t fi = frame_of(gid, inv_frame);
    if (!R.in(fi, (uint64_t)nframes, "frames (emit)", gid)) continue;
    const DevFrame& f = df[fi];
    const Loc lc = locate(f, gid - f.sample_base);
    if (!lc.valid) continue;
    const uint64_t o = f.out_base + (uint64_t)lc.r * f.nc + lc.c;
    if (!R.in(o, nout, "grid", gid)) continue;
    if (seen[o]++) R.fail("grid element %llu written twice (sample %u)", (unsigned long long)o, gid);
  }
  for (uint64_t o = 0; o < nout; o++)
    if (!seen[o]) {
      R.fail("grid element %llu written by no sample", (unsigned long long)o);
      break;
    }

  // ---- the iterate kernels' chunks of an active list of n entries
  std::vector<uint64_t> lens = {0, 1, 3, 4, 63, 64, 65, 255, 256, 257, 1000, 4095, 4097};
  lens.push_back(total / 3);
  lens.push_back(total > 0 ? total - 1 : 0);
  lens.push_back(total);
  std::vector<uint8_t> took;
  for (uint64_t n64 : lens) {
    if (n64 > total) continue;
    const uint32_t n = (uint32_t)n64, nch = chunk_count(n);
    // claim ranges: a partition of [0, nch)
    uint32_t next = 0;
    for (uint32_t h = 0; h < (uint32_t)kQueues; h++) {
      const uint32_t lo = claim_lo(nch, h), hi = claim_lo(nch, h + 1);
      if (lo != next || hi < lo) R.fail("claim range %u of %u chunks: [%u, %u) after %u", h, nch, lo, hi, next);
      next = hi;
    }
    if (next != nch) R.fail("claim ranges end at %u of %u chunks", next, nch);
    // interleaved chunks: runs of 1 (plain), runs of 1 .. 64 (series: 64 is 64
    // consecutive entries); plant 3 interleaves the plain chunks with one too few
    for (uint32_t shift = 0; shift <= 7; shift++) {  // 7: the plain kernel's own chunk_pos
      took.assign(n, 0);
      const uint32_t nchp = plant == 3 && shift == 7 && nch > 1 ? nch - 1 : nch;
      for (uint32_t c = 0; c < nch; c++) {
        uint32_t items = 0;
        auto at = [&](uint32_t l) { return shift == 7 ? chunk_pos(l, c, nchp) : chunk_pos_run(l, c, nchp, shift); };
        for (uint32_t l = 0; l < (uint32_t)kChunk; l++) items += at(l) < n;
        for (uint32_t l = 0; l < (uint32_t)kChunk; l++) {
          const uint32_t p = at(l);
          if ((l < items) != (p < n))
            R.fail("chunk %u of %u entries (shift %u): the lanes with entries are not a prefix", c, n, shift);
          if (l < items && R.in(p, n, "active (chunks)", c) && took[p]++)
            R.fail("active entry %u taken twice (shift %u)", p, shift);
        }
      }
      for (uint32_t p = 0; p < n; p++)
        if (!took[p]) {
          R.fail("active entry %u of %u taken by no chunk (shift %u)", p, n, shift);
          break;
        }
    }
  }
  if (checked) *checked = R.checked;
  return (long long)R.violations;
}

}  // extern "C"
