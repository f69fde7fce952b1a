// rz_tile.hip -- the tile kernel: scanline coverage, burn folds and the
// write-once store of every output tile.
#pragma once

#include "rz_kernels.h"

namespace rzk {

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// XOR-combine a value across the four lanes of a quad with DPP quad
// permutations ([1,0,3,2] then [2,3,0,1]): VALU data movement, no LDS.
__device__ inline uint32_t quad_xor_all(uint32_t x) {
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
  return x;
}
__device__ inline uint64_t quad_xor_all64(uint64_t x) {
  return (uint64_t)quad_xor_all((uint32_t)x) | ((uint64_t)quad_xor_all((uint32_t)(x >> 32)) << 32);
}
__device__ inline int32_t quad_max_all(int32_t x) {
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));
  return x;
}

// Persistent tile kernel.
//
// Scheduling: workgroup b runs on XCD b % 8; each XCD owns a contiguous
// eighth of the tiles (neighbouring tiles share cache lines and items in
// that XCD's L2) and its workgroups walk that range with a stride.  The next
// tile's item range is fetched with scalar loads before the stores go out.
// The four waves of a workgroup work on the four 16-row strips of a tile
// independently (no workgroup barrier).
//
// Per tile (64 rows x TW columns): the background fill of
// RasterInfo::build_raster (rust/src/geo/raster.rs:23-28) and every write of
// every item meeting the tile, folded in geometry order in registers, then
// written to HBM once.
//
// Two lane layouts per wave:
//  * coverage: lane (rl = lane/4, q = lane%4) takes every 4th active edge of
//    row rl; the four lanes of a row XOR-combine their crossing toggles and
//    publish the row's coverage mask in LDS;
//  * fold / store: lane l owns the 16-byte chunk l % LPR of rows
//    l / LPR + k * RSTEP, so every store instruction writes whole tile rows
//    (full 128-byte lines, contiguous across lanes).
// The polygon path is branch-free per lane (predicated crossings and folds).
template <typename T, int FUN, bool LINES>
__global__ __launch_bounds__(256) void k_tile(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int NWORD = TW / 64;                    // u64 words of a row mask
  constexpr int PPC = 16 / (int)sizeof(T);          // pixels per 16-byte chunk
  constexpr int LPR = TW / PPC;                     // lanes per row (store layout)
  constexpr int RSTEP = 64 / LPR;                   // rows per store instruction
  constexpr int RPL = WAVE_ROWS / RSTEP;            // rows per lane
  constexpr int NPIX = RPL * PPC;                   // pixels per lane
  constexpr int CNT_WORDS = LINES ? WAVE_ROWS * TW / 2 : 1;  // u16 counters, 2 per word
  static_assert(NPIX <= 64, "dedup bits assume <= 64 pixels per lane");
  __shared__ PolyEdge s_edge[4][64];
  __shared__ uint64_t s_mask[4][2][WAVE_ROWS][NWORD];
  __shared__ uint32_t s_cnt[4][CNT_WORDS];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;            // coverage layout
  const int srow = lane / LPR, ch = lane % LPR;      // store layout
  const int ccol = ch * PPC;                         // first local column of the chunk
  const double ncols_f = (double)a.ncols;
  const T bg = from_bits<T>(a.bg_bits);
  PolyEdge* stage = s_edge[wave];
  uint32_t* cnt = s_cnt[wave];
  if (LINES) {
    for (int w = lane; w < CNT_WORDS; w += 64) cnt[w] = 0u;  // every apply re-zeroes
    wave_sync();
  }
  unsigned long long burns = 0;
  int mbuf = 0;  // alternating row-mask buffer: one wave_sync per polygon item

  // this XCD's tile range and this workgroup's slot in it
  const uint32_t xcd = blockIdx.x % NUM_XCD, slot = blockIdx.x / NUM_XCD;
  const uint32_t slots = gridDim.x / NUM_XCD;
  const uint32_t per = a.n_tiles / NUM_XCD, rem = a.n_tiles % NUM_XCD;
  const uint32_t lo = xcd * per + min(xcd, rem), hi = lo + per + (xcd < rem ? 1u : 0u);
  const bool binned = a.pair_desc != nullptr;

  uint32_t t = lo + slot;
  uint32_t pb = 0, pe = 0;
  if (t < hi && binned) {
    pb = a.tile_off[t];
    pe = a.tile_off[t + 1];
  }
  for (; t < hi; t += slots) {
    const uint32_t layer = t / a.tiles_per_layer;
    const uint32_t tl = t - layer * a.tiles_per_layer;
    const int32_t ty = (int32_t)(tl / (uint32_t)a.ntx), tx = (int32_t)(tl % (uint32_t)a.ntx);
    const int32_t R0 = ty * TILE_H, C0 = tx * TW;
    const int32_t Rw = R0 + wave * WAVE_ROWS;  // first row of this wave's strip
    const int32_t r = Rw + rl;                 // coverage-layout row of this lane
    const int32_t strip_r1 = min(Rw + WAVE_ROWS, a.nrows);
    const int32_t tile_r1 = min(R0 + TILE_H, a.nrows);
    const int32_t tile_c1 = min(C0 + TW, a.ncols);
    const bool row_ok = r < strip_r1;

    T acc[NPIX];  // acc[k * PPC + j]: row srow + k*RSTEP, column ccol + j
#pragma unroll
    for (int j = 0; j < NPIX; ++j) acc[j] = bg;

    auto process_item = [&](const ItemDesc& d) {
      // wave-uniform skip: the item box misses this wave's strip or the tile
      if (d.r1 <= Rw || d.r0 >= strip_r1 || d.c1 <= C0 || d.c0 >= tile_c1) return;
      const int32_t kind = (int32_t)(d.kind_layer & 3u);
      const T v = from_bits<T>(d.value);
      const int64_t vb = d.vbeg, ve = d.vbeg + d.nv;
      if (kind == 0) {
        // ---- polygon: even-odd coverage of every row from crossing counts
        uint64_t tog[NWORD];
#pragma unroll
        for (int w = 0; w < NWORD; ++w) tog[w] = 0ull;
        uint32_t lp = 0;
        int32_t amax = -1;
        for (int64_t cb = vb; cb < ve; cb += 64) {
          // stage the edges of this chunk that reach the strip, compacted
          const int64_t e = cb + lane;
          PolyEdge ed;
          bool act = false;
          if (e < ve) {
            ed = a.edges[e];
            act = ed.ystart < strip_r1 && ed.yend > Rw && ed.ystart < ed.yend;
          }
          const uint64_t m = __ballot(act);
          if (act) stage[__popcll(m & ((1ull << lane) - 1ull))] = ed;
          wave_sync();
          const int n = __popcll(m);
          for (int k0 = 0; k0 < n; k0 += 4) {  // uniform trip count, predicated lanes
            const int k = k0 + q;
            const PolyEdge ee = stage[k < n ? k : n - 1];
            const bool on = k < n && row_ok && ee.ystart <= r && r < ee.yend;
            const double x = crossing(ee, r);
            toggle_crossing_if<NWORD>(on && !is_pos_nan(x), span_col(x, ncols_f), C0, tog, lp,
                                      amax);
          }
          wave_sync();
        }
        // combine the four lanes of this row, then publish the row's coverage
        // mask (the note above row_coverage_word in rz_kernels.h states why it
        // equals the reference's spans) for the store layout
        lp = quad_xor_all(lp);
        amax = quad_max_all(amax);
#pragma unroll
        for (int w = 0; w < NWORD; ++w) tog[w] = quad_xor_all64(tog[w]);
        if (q < NWORD) s_mask[wave][mbuf][rl][q] = row_coverage_word<NWORD>(tog, lp, amax, C0, q);
        wave_sync();
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const uint64_t row_word = s_mask[wave][mbuf][srow + k * RSTEP][ccol >> 6];
          const uint32_t bits = (uint32_t)(row_word >> (ccol & 63)) & ((1u << PPC) - 1u);
          if (__any(bits != 0u)) {  // wave-uniform skip of uncovered row groups
            burns += (unsigned long long)__popc(bits);
#pragma unroll
            for (int j = 0; j < PPC; ++j)
              acc[k * PPC + j] = fold_select<T, FUN>(acc[k * PPC + j], v, bg, (bits >> j) & 1u);
          }
        }
        mbuf ^= 1;  // the next polygon writes the other buffer
        return;
      }
      if (!LINES) return;  // the host only picks this variant without lines / points
      // ---- lines and points: multiplicities in wave-private u16 counters
      auto emit = [&](int64_t cx, int64_t cy) {
        if (cx < C0 || cx >= tile_c1 || cy < Rw || cy >= strip_r1) return;
        const uint32_t idx = (uint32_t)(cy - Rw) * TW + (uint32_t)(cx - C0);
        atomicAdd(&cnt[idx >> 1], 1u << ((idx & 1u) * 16u));
      };
      const bool dedup = kind == 1 && a.dedup_lines;
      uint64_t seen = 0u;  // PixelCache bits of this lane's pixels
      for (int64_t cs = vb; cs < ve; cs += 65535) {  // <= 65535 writes per pixel per chunk
        const int64_t ce = min(ve, cs + 65535);
        for (int64_t s = cs + lane; s < ce; s += 64) {
          const double2 p0 = a.px[s];
          if (kind == 2) {  // point: extract_point + burn_point
            if (p0.x >= 0.0 && p0.x < ncols_f && p0.y >= 0.0 && p0.y < (double)a.nrows)
              emit((int64_t)p0.x, (int64_t)p0.y);
            continue;
          }
          const uint8_t f = a.vflags[s];
          if (!(f & VF_HAS_NEXT)) continue;
          const double2 p1 = a.px[s + 1];
          if (!(fmin(p0.x, p1.x) < ncols_f && fmax(p0.x, p1.x) >= 0.0 &&
                fmin(p0.y, p1.y) < (double)a.nrows && fmax(p0.y, p1.y) >= 0.0))
            continue;  // extract_line bbox filter, edges.rs:130
          const int64_t ix0 = sat_isize(floor(p0.x)), iy0 = sat_isize(floor(p0.y));
          const int64_t ix1 = sat_isize(floor(p1.x)), iy1 = sat_isize(floor(p1.y));
          const int64_t lim = (int64_t)1 << 30, lim2 = (int64_t)1 << 60;
          if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim &&
              iy1 > -lim && iy1 < lim) {
            walk_window<int64_t>(ix0, iy0, ix1, iy1, C0, tile_c1, Rw, strip_r1, emit);
          } else if (ix0 > -lim2 && ix0 < lim2 && ix1 > -lim2 && ix1 < lim2 && iy0 > -lim2 &&
                     iy0 < lim2 && iy1 > -lim2 && iy1 < lim2) {
            walk_window<__int128>(ix0, iy0, ix1, iy1, C0, tile_c1, Rw, strip_r1, emit);
          } else {
            atomicOr(a.status, 1u);  // pixel coordinates beyond +-2^60: unsupported
          }
          // end cell of the last kept edge of an open line (burners.rs:86-89)
          if ((uint64_t)(s + 1) == d.last && !(f & VF_CLOSED)) emit(ix1, iy1);
        }
        wave_sync();
        // apply the multiplicities of this lane's pixels, then clear them
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          uint32_t* crow = cnt + ((srow + k * RSTEP) * TW + ccol) / 2;
#pragma unroll
          for (int j2 = 0; j2 < PPC / 2; ++j2) {
            const uint32_t wv = crow[j2];
            if (wv) {
              crow[j2] = 0u;
#pragma unroll
              for (int h = 0; h < 2; ++h) {
                uint32_t kk = (wv >> (16 * h)) & 0xFFFFu;
                const int j = k * PPC + 2 * j2 + h;
                if (kk && dedup) {  // LineWriter + PixelCache: one write per pixel per item
                  const uint64_t bit = 1ull << j;
                  if (seen & bit) kk = 0;
                  else {
                    seen |= bit;
                    kk = 1;
                  }
                }
                if (kk) {
                  burns += kk;
                  fold<T, FUN>(acc[j], v, bg, kk);
                }
              }
            }
          }
        }
        wave_sync();
      }
    };

    if (binned) {
      for (uint32_t i = pb; i < pe; ++i) process_item(a.pair_desc[i]);
    } else {
      // scan mode: test every item's box in order (small item counts)
      for (int64_t base = 0; base < a.n_items; base += 64) {
        const int64_t it = base + lane;
        bool hit = false;
        if (it < a.n_items) {
          const ItemDesc d = a.desc[it];
          hit = (d.kind_layer >> 2) == layer && d.r0 < tile_r1 && d.r1 > R0 && d.c0 < tile_c1 &&
                d.c1 > C0 && d.r0 < d.r1 && d.c0 < d.c1;
        }
        uint64_t m = __ballot(hit);
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          process_item(a.desc[base + bit]);
        }
      }
    }

    // fetch the next tile's item range before this tile's stores
    const uint32_t tn = t + slots;
    if (tn < hi && binned) {
      pb = a.tile_off[tn];
      pe = a.tile_off[tn + 1];
    }

    // ---- write the tile once (background + burns): each store instruction
    // covers RSTEP whole rows of the tile
    const int32_t c = C0 + ccol;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int32_t row = Rw + srow + k * RSTEP;
      if (row < a.nrows) {
        T* orow = (T*)a.out + ((uint64_t)layer * (uint64_t)a.nrows + (uint64_t)row) * (uint64_t)a.ncols;
        if (c + PPC <= a.ncols && ((uintptr_t)(orow + c) & 15u) == 0) {
          u32x4 vec;
          __builtin_memcpy(&vec, &acc[k * PPC], 16);
          __builtin_nontemporal_store(vec, (u32x4*)(orow + c));
        } else {
#pragma unroll
          for (int j = 0; j < PPC; ++j)
            if (c + j < a.ncols) orow[c + j] = acc[k * PPC + j];
        }
      }
    }
  }
  if (a.burns) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) burns += __shfl_xor(burns, off, 64);
    if (lane == 0 && burns) atomicAdd(&a.burns[blockIdx.x % kBurnSlots], burns);
  }
}

}  // namespace rzk
