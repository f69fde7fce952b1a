// rz_tile.hip -- coverage pre-passes of the tile kernel: heavy-tile lists
// and masks (k_heavy_*, k_cover) and the wide pre-pass (k_wide_cover).
#pragma once

#include "rz_device.h"

namespace rzk {

// ---------------------------------------------------------- heavy tiles --
//
// A tile with at least kHeavyItems items would serialise all their coverage
// work in its four strip waves.  For such tiles the coverage masks of every
// (polygon item, strip) are computed beforehand by k_cover, one wave per
// (item, strip), in parallel; the tile kernel then only folds the masks in
// geometry order (every pixel function, ordered sum included, stays exact).

// Items of every heavy tile (0 for the others) and its heavy flag.
__global__ void k_heavy_count(const uint32_t* tile_off, uint32_t n_tiles, uint32_t min_items,
                              uint32_t* heavy_cnt, uint32_t* heavy_flag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  uint32_t c = 0;
  if (t < n_tiles) {
    c = tile_off[t + 1] - tile_off[t];
    if (c < min_items) c = 0;
  }
  heavy_cnt[t] = c;
  heavy_flag[t] = c > 0 ? 1u : 0u;
}

// The heavy tiles in tile order: heavy_list[heavy_rank[t]] = t, so the heavy
// tiles of tiles [t_lo, t_hi) are heavy_list[heavy_rank[t_lo] .. heavy_rank[t_hi]).
__global__ void k_heavy_list(const uint32_t* heavy_flag, const uint32_t* heavy_rank, uint32_t n_tiles,
                             uint32_t* heavy_list) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_tiles && heavy_flag[t]) heavy_list[heavy_rank[t]] = t;
}

// The compact list of the pairs of heavy tiles that hold polygon items other
// than wide ones (the only ones k_cover writes; k_wide_cover writes the wide
// ones), appended with one atomic per block.  For a wide item's pair it also
// records the item's place in the tile's pair list, at the pair's emission
// index (k_bin_emit's order over the item's tile rectangle), where
// k_wide_cover finds it with one load.
constexpr int kHeavyPairsBlock = 1024;
__global__ __launch_bounds__(kHeavyPairsBlock) void k_heavy_pairs(const JobDev* jobs, const uint32_t* pair_tile,
                              const uint32_t* pair_item, const ItemDesc* pair_desc, const uint64_t* n_pairs, uint64_t cap,
                              const uint32_t* tile_off, const uint64_t* heavy_base, const uint64_t* pair_off,
                              uint32_t* wide_j, uint32_t* cover_list, uint32_t* cover_count, uint32_t n_tiles) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool poly = false;
  if (i < min(*n_pairs, cap)) {
    const uint32_t t = pair_tile[i];
    // (a key past every tile: a slot of a captured batch whose pairs outgrew
    // its capacity, whose tile lists are then all empty)
    if (t < n_tiles && heavy_base[t + 1] > heavy_base[t]) {
      const ItemDesc& d = pair_desc[i];
      poly = (d.kind_layer & 3u) == 0u && !(d.flags & kItemWide);  // k_wide_cover writes those
      int32_t tx0, tx1, ty0, ty1;
      if ((d.flags & kItemWide) && tile_rect(d.r0, d.r1, d.c0, d.c1, jobs[d.job].tile_w, tx0, tx1, ty0, ty1)) {
        const JobDev& J = jobs[d.job];
        const uint32_t tl = (t - J.tile_base) % J.tiles_per_layer;
        const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
        wide_j[pair_off[pair_item[i]] + (uint64_t)((ty - ty0) * (tx1 - tx0 + 1) + (tx - tx0))] =
            (uint32_t)(i - tile_off[t]);
      }
    }
  }
  // one atomic per block (the list's order does not matter)
  __shared__ uint32_t s_n[kHeavyPairsBlock / 64], s_base;
  const int wave = threadIdx.x >> 6;
  const uint64_t m = __ballot(poly);
  if (lane == 0) s_n[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = 0;
    for (int w = 0; w < kHeavyPairsBlock / 64; ++w) n += s_n[w];
    s_base = n ? atomicAdd(cover_count, n) : 0u;
  }
  __syncthreads();
  uint32_t base = s_base;
  for (int w = 0; w < wave; ++w) base += s_n[w];
  if (poly) cover_list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
}

template <int NWORD>
__global__ __launch_bounds__(256) void k_cover(CoverArgs a) {
  constexpr int TW = 64 * NWORD;
  __shared__ PolyEdge s_edge[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST uint64_t* band_c = as_const(a.band_off);
  const uint64_t units = 4ull * *a.cover_count;
  const uint64_t ustride = (uint64_t)gridDim.x * 4u;
  uint64_t u = (uint64_t)blockIdx.x * 4u + (uint64_t)wave;
  // the next unit's pair is loaded one unit ahead
  uint32_t i_next = u < units ? a.cover_list[u >> 2] : 0u;
  for (; u < units; u += ustride) {
    const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)i_next);  // wave-uniform
    if (u + ustride < units) i_next = a.cover_list[(u + ustride) >> 2];
    const int32_t strip = (int32_t)(u & 3u);
    const uint32_t t = a.pair_tile[i];
    const ItemDesc d = load_desc(as_const(a.pair_desc) + i);
    const RZ_CONST JobDev& J = jobs_c[d.job];
    if (J.tile_w != TW) continue;
    const uint32_t tj = t - J.tile_base;
    const uint32_t layer = tj / J.tiles_per_layer;
    const uint32_t tl = tj - layer * J.tiles_per_layer;
    const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
    const int32_t C0 = tx * TW, Rw = ty * TILE_H + strip * WAVE_ROWS;
    const int32_t strip_r1 = min(Rw + WAVE_ROWS, J.nrows), tile_c1 = min(C0 + TW, J.ncols);
    if (d.r1 <= Rw || d.r0 >= strip_r1 || d.c1 <= C0 || d.c0 >= tile_c1) continue;
    const PolyEdge* ebase = a.edges + d.vbeg;
    uint32_t n_e = (uint32_t)d.nv;
    if (d.band) {
      const uint32_t bi = d.band - 1u + (uint32_t)(Rw / kBandRows - d.r0 / kBandRows);
      const uint64_t b0 = band_c[bi];
      ebase = a.band_edges + b0;
      n_e = (uint32_t)(band_c[bi + 1u] - b0);
    }
    const uint64_t cov = strip_coverage<NWORD>(ebase, n_e, Rw, strip_r1, C0, (double)J.ncols, s_edge[wave], lane);
    // strip-major slots: item j of heavy tile t, strip s at heavy_base[t] * 4 + s * cnt + j
    const uint32_t p0 = a.tile_off[t], cnt = a.tile_off[t + 1] - p0;
    const uint64_t slot = a.heavy_base[t] * 4u + (uint64_t)strip * cnt + (i - p0);
    if (q < NWORD) a.masks[slot * a.slot_words + (uint64_t)(rl * NWORD + q)] = cov;
  }
}

// The wide pre-pass: one wave per unit = (kItemWide item, 16-row strip,
// window of at most kWideWords mask words), grid-stride over the unit list.
// The strip's edges (its band, or the whole edge list) are visited once for
// the window, in the cheaper of the two layouts per chunk of 64 edges
// (row-major: four lanes per row over the staged active edges; edge-major:
// one lane per edge over the rows it crosses), and every crossing updates
// the row's LDS state with an atomic (XOR of its toggle bit or of the left
// parity, max of the column; order-free).  Then two rows at a time, lane
// (half, w) turns word w of the toggles of its row into coverage (prefix
// parity across the words from one ballot, right limit from the rightmost
// crossing) and stores it.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_wide_cover(WideCoverArgs a) {
  __shared__ PolyEdge s_edge[4][64];
  // rows padded by one word, so the same word of different rows falls in
  // different LDS banks
  __shared__ uint64_t s_tog[4][WAVE_ROWS][kWideWords + 1];
  __shared__ uint32_t s_lp[4][WAVE_ROWS];
  __shared__ int32_t s_amax[4][WAVE_ROWS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST uint64_t* band_c = as_const(a.band_off);
  const uint64_t units = min(*a.n_units, a.unit_cap);
  PolyEdge* stage = s_edge[wave];
  uint64_t(*tog)[kWideWords + 1] = s_tog[wave];
  uint32_t* lp = s_lp[wave];
  int32_t* amax = s_amax[wave];
  if (a.n_items <= 0) return;
  // the toggle rows start at zero, and each unit's finalize step clears the
  // words it read (no crossing touches a word past the unit's window)
  for (int idx = lane; idx < WAVE_ROWS * (kWideWords + 1); idx += 64) (&tog[0][0])[idx] = 0ull;
  const uint64_t ustride = (uint64_t)gridDim.x * 4u;
  uint64_t u = (uint64_t)blockIdx.x * 4u + (uint64_t)wave;
  // the next unit's item is loaded one unit ahead, and a unit's range and
  // descriptor are loaded together (clamped to a valid item) and checked
  // afterwards, so a unit starts after one round trip instead of three
  uint32_t it_next = u < units ? a.unit_item[u] : 0u;
  for (; u < units; u += ustride) {
    // (wave-uniform: the unit's header and descriptor go to scalar registers)
    const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane((int)it_next);
    if (u + ustride < units) it_next = a.unit_item[u + ustride];
    const uint32_t itc = (int64_t)it < a.n_items ? it : (uint32_t)(a.n_items - 1);
    const uint64_t ub = a.unit_base[itc], ue = a.unit_base[itc + 1];
    const ItemDesc d = load_desc(as_const(a.desc) + itc);
    // a unit of an item past the capacities, or a stale one
    if ((int64_t)it >= a.n_items || u < ub || u >= ue || !(d.flags & kItemWide)) continue;
    const RZ_CONST JobDev& J = jobs_c[d.job];
    const int32_t nw = wide_row_words(d.c0, d.c1);
    const int32_t nwin = (nw + kWideWords - 1) / kWideWords;
    const int32_t local = (int32_t)(u - ub);
    const int32_t k = local / nwin, w0 = (local - k * nwin) * kWideWords;
    const int32_t ww = min(kWideWords, nw - w0);
    const int32_t Rw = ((d.r0 >> 4) + k) * WAVE_ROWS, strip_r1 = min(Rw + WAVE_ROWS, J.nrows);
    const int32_t W0 = ((d.c0 >> 6) + w0) * 64, WEND = W0 + ww * 64;
    const double ncols_f = (double)J.ncols;
    const PolyEdge* ebase = a.edges + d.vbeg;
    uint32_t n_e = (uint32_t)d.nv;
    if (d.band) {
      const uint32_t bi = d.band - 1u + (uint32_t)k;
      const uint64_t b0 = band_c[bi];
      ebase = a.band_edges + b0;
      n_e = (uint32_t)(band_c[bi + 1u] - b0);
    }
    // the rightmost crossing of a row matters only with an odd count
    const bool odd_rows = (d.flags & kItemOddRows) != 0;
    if (lane < WAVE_ROWS) {
      lp[lane] = 0u;
      amax[lane] = -1;
    }
    wave_sync();
    // one crossing of row `row` (0..15) at x
    auto put = [&](int32_t row, double x) {
      const int32_t col = span_col(x, ncols_f);
      const int32_t rel = col - W0;
      if (rel <= 0) {
        __hip_atomic_fetch_xor(&lp[row], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        return;
      }
      if (odd_rows) __hip_atomic_fetch_max(&amax[row], col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      if (col < WEND)
        __hip_atomic_fetch_xor(&tog[row][rel >> 6], 1ull << (rel & 63), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
    };
    for (uint32_t cb = 0; cb < n_e; cb += 64) {
      const uint32_t e = cb + (uint32_t)lane;
      uint4 e_lo = make_uint4(0u, 0u, 0u, 0u), e_hi = make_uint4(0u, 0u, 0u, 0u);
      int32_t a0 = 0, span = 0;  // rows [a0, a0 + span) of the strip the edge crosses
      if (e < n_e) {
        const uint4* src = reinterpret_cast<const uint4*>(ebase + e);
        e_lo = src[0];
        e_hi = src[1];  // dxdy, ystart (z), yend (w)
        a0 = max((int32_t)e_hi.z, Rw);
        span = max(0, min((int32_t)e_hi.w, strip_r1) - a0);
      }
      const uint64_t m = __ballot(span > 0);
      const int n = __popcll(m);
      const int32_t steps = wave_max16(span);
      if (a.cov_mode == 2 || (a.cov_mode == 0 && steps <= (n + 3) / 4)) {  // edge-major
        PolyEdge ee;
        __builtin_memcpy(&ee, &e_lo, 16);
        __builtin_memcpy(reinterpret_cast<char*>(&ee) + 16, &e_hi, 16);
        for (int32_t j = 0; j < steps; ++j) {
          const int32_t rr = a0 + j;
          const double x = crossing(ee, rr);
          if (j < span) put(rr - Rw, x);
        }
        continue;
      }
      // row-major over the staged active edges
      if (span > 0) {
        uint4* dst = reinterpret_cast<uint4*>(stage + __popcll(m & ((1ull << lane) - 1ull)));
        dst[0] = e_lo;
        dst[1] = e_hi;
      }
      wave_sync();
      const int32_t r = Rw + rl;
      for (int k0 = 0; k0 < n; k0 += 4) {
        const int kk = k0 + q;
        const PolyEdge ee = stage[kk < n ? kk : n - 1];
        const double x = crossing(ee, r);
        if (kk < n && r < strip_r1 && ee.ystart <= r && r < ee.yend) put(rl, x);
      }
      wave_sync();
    }
    wave_sync();
    // two rows per step: lanes 0..31 row 2i, lanes 32..63 row 2i + 1
    static_assert(kWideWords == 32, "two rows of window words per wave step");
    uint64_t* out = a.wide_masks + (uint64_t)d.wide * 16u + (uint64_t)k * 16u * (uint64_t)nw + (uint64_t)w0;
    const int half = lane >> 5, w = lane & 31;
    // A heavy tile's fold reads its items' masks from its pre-pass slots, so
    // each word also goes to the slot of the tile that holds it when that
    // tile is heavy: tile column gw / (TW / 64), the item's place j in the
    // tile's (item-ordered) pair list found by binary search.
    uint64_t* slot = nullptr;
    if (a.heavy_cnt && !J.scan && w < ww) {
      const int32_t nwt = J.tile_w >> 6;  // words per tile row
      const int32_t gw = (W0 >> 6) + w, tx = gw / nwt;
      const uint32_t t = J.tile_base + (d.kind_layer >> 2) * J.tiles_per_layer +
                         (uint32_t)(Rw / TILE_H) * (uint32_t)J.ntx + (uint32_t)tx;
      const uint32_t cnt = a.heavy_cnt[t];
      // the pair (it, t) in k_bin_emit's order over the item's tile
      // rectangle: row Rw / TILE_H - ty0, column tx - tx0 (uniform part first)
      int32_t tx0, tx1, ty0, ty1;
      uint32_t j = ~0u;
      if (cnt && a.pair_off[it + 1] <= a.pair_cap &&
          tile_rect(d.r0, d.r1, d.c0, d.c1, J.tile_w, tx0, tx1, ty0, ty1)) {
        const int64_t jrow = (int64_t)a.pair_off[it] + (int64_t)(Rw / TILE_H - ty0) * (tx1 - tx0 + 1) - tx0;
        j = a.wide_j[jrow + tx];
      }
      if (j < cnt) {
        const uint32_t s = (uint32_t)((Rw % TILE_H) / WAVE_ROWS);
        slot = a.masks + (a.heavy_base[t] * 4u + (uint64_t)s * cnt + j) * a.slot_words + (uint32_t)(gw % nwt);
      }
    }
    const uint64_t below = ((1ull << lane) - 1ull) & (half ? 0xFFFFFFFF00000000ull : 0xFFFFFFFFull);
    for (int r2 = 0; r2 < WAVE_ROWS; r2 += 2) {
      const int row = r2 + half;
      uint64_t t = 0ull;
      if (w < ww) {
        t = tog[row][w];
        tog[row][w] = 0ull;
      }
      const uint64_t odd = __ballot(__popcll(t) & 1);
      const uint32_t carry = lp[row] ^ ((uint32_t)__popcll(odd & below) & 1u);
      uint64_t x = prefix_xor64(t) ^ (carry ? ~0ull : 0ull);
      const int32_t lim = odd_rows ? amax[row] - (W0 + 64 * w) : 64;
      x &= lim <= 0 ? 0ull : (lim >= 64 ? ~0ull : ((1ull << lim) - 1ull));
      if (w < ww) out[(uint64_t)row * (uint64_t)nw + (uint64_t)w] = x;
      if (slot) {
        // the tile's words left of the item's first word or right of its
        // last one are empty: written as such (the fold reads every word)
        const int32_t nwt = J.tile_w >> 6, gw = (W0 >> 6) + w, qq = gw % nwt;
        uint64_t* srow_p = slot + row * nwt - qq;
        srow_p[qq] = x;
        if (gw == (d.c0 >> 6))
          for (int z = 0; z < qq; ++z) srow_p[z] = 0ull;
        if (gw == ((d.c1 - 1) >> 6))
          for (int z = qq + 1; z < nwt; ++z) srow_p[z] = 0ull;
      }
    }
    wave_sync();  // the rows are read and cleared before the next unit's crossings
  }
}

}  // namespace rzk
