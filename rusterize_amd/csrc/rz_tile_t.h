// rz_tile_t.h -- the tile kernel templates (k_tile, k_tile_heavy): scanline
// coverage, burn folds and the write-once store of every output tile.
// Instantiated per output dtype in rz_tile_inst.hip.
#pragma once

#include "rz_device.h"

namespace rzk {

// Persistent work queue of the tile kernels.  A unit of work is one 16-row
// strip of one output tile, and wave w of every workgroup takes strip w, so
// the waves working on one strip index walk neighbouring tiles together
// (their stores fill neighbouring lines).  Workgroup b runs on XCD b % 8;
// each XCD owns a contiguous eighth of the launch's tiles, so neighbouring
// tiles share cache lines and items in that XCD's L2.  A wave's first tile
// is its workgroup's slot in that range (no atomic); after the first `slots`
// tiles, queue (x, w) hands out blocks of a.chunk consecutive tiles.  A grab
// is issued one block ahead, so its latency overlaps the current block.  A
// wave whose own XCD range is used up reads all eight counters of its strip
// with one relaxed load (counters only grow: a range that reads as used up
// is used up) and continues on the next XCD range with work left; jobs of
// very uneven size still end together, and every wave leaves when all eight
// ranges of its strip are used up.  The waves of a workgroup never
// synchronise with each other.
struct TileQueue {
  uint32_t n_loc, xcd, slots, B, cur_x, pending, t_cur, t_end;
  uint32_t* qbase;  // queue (x, w) at qbase + 4x * kQueueStride
  int lane;

  __device__ TileQueue(const TileArgs& a, int wave, int lane_) : lane(lane_) {
    n_loc = a.t_hi - a.t_lo;
    xcd = blockIdx.x % NUM_XCD;
    const uint32_t slot = blockIdx.x / NUM_XCD;
    slots = gridDim.x / NUM_XCD;
    B = a.chunk;
    qbase = a.queue + (uint32_t)wave * kQueueStride;
    cur_x = xcd;
    pending = 0;  // lane 0: result of the grab in flight on cur_x
    t_cur = xcd_unit_lo(xcd, n_loc) + slot;  // current block [t_cur, t_end)
    t_end = t_cur + 1u;
  }
  __device__ void issue(uint32_t x) {
    if (lane == 0) pending = atomicAdd(qbase + x * 4u * kQueueStride, 1u);
  }
  // first tile handed out by queue x's grab number c
  __device__ uint64_t grab_tile(uint32_t x, uint32_t c) const {
    return (uint64_t)xcd_unit_lo(x, n_loc) + slots + (uint64_t)c * B;
  }
  __device__ bool take(uint32_t x) {
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)pending, 0);
    const uint64_t s = grab_tile(x, c), xhi = xcd_unit_lo(x + 1u, n_loc);
    if (s >= xhi) return false;
    t_cur = (uint32_t)s;
    t_end = (uint32_t)min(s + (uint64_t)B, xhi);
    return true;
  }
  // (a grab on cur_x is always in flight when next_block runs)
  __device__ bool next_block() {
    if (take(cur_x)) return true;
    for (;;) {
      bool avail = false;
      if (lane < NUM_XCD) {
        const uint32_t seen = __hip_atomic_load(qbase + (uint32_t)lane * 4u * kQueueStride,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        avail = grab_tile((uint32_t)lane, seen) < (uint64_t)xcd_unit_lo((uint32_t)lane + 1u, n_loc);
      }
      const uint32_t m = (uint32_t)__ballot(avail) & 0xFFu;
      if (m == 0u) return false;
      const uint32_t rot = ((m >> xcd) | (m << (NUM_XCD - xcd))) & 0xFFu;
      const uint32_t x = (xcd + (uint32_t)__builtin_ctz(rot)) % NUM_XCD;
      issue(x);
      if (take(x)) {
        cur_x = x;
        return true;
      }
    }
  }
  // Whether the wave has a first tile (t_cur).
  __device__ bool begin() {
    bool have = t_cur < xcd_unit_lo(xcd + 1u, n_loc);
    issue(xcd);  // the grab after the first tile (or the first block, if there is no first tile)
    if (!have) {
      have = next_block();
      if (have) issue(cur_x);
    }
    return have;
  }
  // Local index of the current tile; moves to the next one.
  __device__ uint32_t pop() { return t_cur++; }
  // Whether there is a next tile (t_cur), taking the next block when needed.
  __device__ bool advance() {
    if (t_cur != t_end) return true;
    const bool have = next_block();
    if (have) issue(cur_x);
    return have;
  }
};

// Geometry of one unit: job, band, tile and strip window.
struct StripUnit {
  const RZ_CONST JobDev* J;
  uint32_t layer;
  int32_t nrows, ncols, C0, Rw, strip_r1, tile_c1;
};

__device__ inline StripUnit strip_unit(const TileArgs& a, uint32_t t, int32_t strip, int tw) {
  StripUnit u;
  u.J = &as_const(a.jobs)[find_job(as_const(a.job_tend), a.n_jobs, (int64_t)t)];
  u.nrows = u.J->nrows;
  u.ncols = u.J->ncols;
  const uint32_t tj = t - u.J->tile_base;
  u.layer = tj / u.J->tiles_per_layer;
  const uint32_t tl = tj - u.layer * u.J->tiles_per_layer;
  const int32_t ty = (int32_t)(tl / (uint32_t)u.J->ntx), tx = (int32_t)(tl % (uint32_t)u.J->ntx);
  u.C0 = tx * tw;
  u.Rw = ty * TILE_H + strip * WAVE_ROWS;  // first row of the strip
  u.strip_r1 = min(u.Rw + WAVE_ROWS, u.nrows);
  u.tile_c1 = min(u.C0 + tw, u.ncols);
  return u;
}

// A polygon item's coverage mask word for this lane's row / quarter (lane
// rl = lane / 4, q = lane % 4): precomputed (wide item, heavy tile) or
// computed from its edges (banded: only the strip's band).
template <int NWORD>
__device__ inline uint64_t polygon_mask(const TileArgs& a, const ItemDesc& d, const StripUnit& u,
                                        bool hb_ok, uint64_t hbase, uint32_t pb, uint32_t pe, uint32_t pi,
                                        int32_t strip, PolyEdge* stage, int lane) {
  const int rl = lane >> 2, q = lane & 3;
  uint64_t cov = 0;
  if (d.flags & kItemWide) {  // the wide pre-pass wrote the item's masks
    const int32_t nw = wide_row_words(d.c0, d.c1), wi = (u.C0 >> 6) + q - (d.c0 >> 6);
    if (q < NWORD && wi >= 0 && wi < nw)
      cov = as_const(a.wide_masks)[(uint64_t)d.wide * 16u +
                                   (uint64_t)(((u.Rw >> 4) - (d.r0 >> 4)) * 16 + rl) * (uint64_t)nw + (uint64_t)wi];
  } else if (hb_ok) {  // heavy tile: the coverage pre-pass wrote this mask
    if (q < NWORD)
      cov = as_const(a.masks)[(hbase * 4u + (uint64_t)strip * (pe - pb) + (pi - pb)) * a.slot_words +
                              (uint32_t)(rl * NWORD + q)];
  } else {
    // the item's edges, or for a banded item only those of this strip
    const PolyEdge* ebase = a.edges + d.vbeg;
    uint32_t n_e = (uint32_t)d.nv;
    if (d.band) {
      const uint32_t bi = d.band - 1u + (uint32_t)(u.Rw / kBandRows - d.r0 / kBandRows);
      const uint64_t b0 = as_const(a.band_off)[bi];
      ebase = a.band_edges + b0;
      n_e = (uint32_t)(as_const(a.band_off)[bi + 1u] - b0);
    }
    if constexpr (NWORD == 1) {
      cov = strip_coverage<1>(ebase, n_e, u.Rw, u.strip_r1, u.C0, (double)u.ncols, stage, lane);
    } else {
      // only the tile words the item's columns [c0, c1) meet can be
      // covered: the crossings are toggled and the parity scanned over
      // those (1, 2 or all 4 words; crossings left of them flip the left
      // parity, as left of a tile)
      const int32_t wlo = max(0, (d.c0 - u.C0) >> 6), whi = min(NWORD - 1, (d.c1 - 1 - u.C0) >> 6);
      const int32_t C0e = u.C0 + 64 * wlo;
      const int wq = (lane & 3) - wlo;  // this lane's word in the narrowed window (negative: none)
      if (whi == wlo)
        cov = strip_coverage<1>(ebase, n_e, u.Rw, u.strip_r1, C0e, (double)u.ncols, stage, lane, wq == 0 ? 0 : 1);
      else if (whi == wlo + 1)
        cov = strip_coverage<2>(ebase, n_e, u.Rw, u.strip_r1, C0e, (double)u.ncols, stage, lane, wq < 0 ? 2 : wq);
      else
        cov = strip_coverage<NWORD>(ebase, n_e, u.Rw, u.strip_r1, u.C0, (double)u.ncols, stage, lane);
    }
  }
  return cov;
}

// Writes a strip once: lane l's chunk of PPC pixels in rows
// srow + k * RSTEP (k < RPL) starting at column C0 + ccol, from `chunk(k)`
// (16 bytes); each store instruction covers RSTEP whole tile rows.
template <typename T, int PPC, int RPL, int RSTEP, typename F>
__device__ inline void store_strip(const StripUnit& u, int srow, int ccol, F&& chunk) {
  const int32_t c = u.C0 + ccol;
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int32_t row = u.Rw + srow + k * RSTEP;
    if (row < u.nrows) {
      T* orow = (T*)u.J->out + ((uint64_t)u.layer * (uint64_t)u.nrows + (uint64_t)row) * (uint64_t)u.ncols;
      const u32x4 vec = chunk(k);
      if (c + PPC <= u.ncols && ((uintptr_t)(orow + c) & 15u) == 0) {
        __builtin_nontemporal_store(vec, (u32x4*)(orow + c));
      } else {
        T px[PPC];
        __builtin_memcpy(px, &vec, 16);
#pragma unroll
        for (int j = 0; j < PPC; ++j)
          if (c + j < u.ncols) orow[c + j] = px[j];
      }
    }
  }
}

// Waves per SIMD the register allocation of k_tile targets: 7 for 4-byte
// outputs, 6 for 1- and 2-byte outputs (measured: the latency-bound strip
// loop gains more from the extra wave than the few spilled registers cost),
// 5 for 8-byte outputs.
template <typename T>
struct TileOccupancy {
  static constexpr int value = sizeof(T) == 4 ? 7 : (sizeof(T) == 8 ? 5 : 6);
};

// Tile kernel of polygon-only groups, one launch per pixel-function group of
// a batch (a single call is a batch of one job).
//
// Per unit (16 rows x TW columns of the job's output): the background fill
// of RasterInfo::build_raster (rust/src/geo/raster.rs:23-28) and every write
// of every item meeting the strip, folded in geometry order in registers,
// then written to HBM once.
//
// Two lane layouts per wave:
//  * coverage: lane (rl = lane/4, q = lane%4) takes every 4th active edge of
//    row rl; the four lanes of a row XOR-combine their crossing toggles and
//    publish the row's coverage mask in LDS;
//  * fold / store: lane l owns the 16-byte chunk l % LPR of rows
//    l / LPR + k * RSTEP, so every store instruction writes whole tile rows
//    (full 128-byte lines, contiguous across lanes).
// The polygon path is branch-free per lane (predicated crossings and folds).
template <typename T, int FUN>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(TileOccupancy<T>::value))) void k_tile(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int NWORD = TW / 64;                    // u64 words of a row mask
  constexpr int PPC = 16 / (int)sizeof(T);          // pixels per 16-byte chunk
  constexpr int LPR = TW / PPC;                     // lanes per row (store layout)
  constexpr int RSTEP = 64 / LPR;                   // rows per store instruction
  constexpr int RPL = WAVE_ROWS / RSTEP;            // rows per lane
  constexpr int NPIX = RPL * PPC;                   // pixels per lane
  // 1- and 2-byte outputs keep their pixels packed, four (two) per 32-bit
  // word, and fold them lane-wise (fold_word)
  constexpr bool PACKED = sizeof(T) <= 2;
  constexpr int NW = PACKED ? NPIX * (int)sizeof(T) / 4 : 1;  // words per lane
  constexpr int PPW = 4 / (int)sizeof(T);                     // pixels per word (packed)
  __shared__ PolyEdge s_edge[4][64];
  __shared__ uint64_t s_mask[4][2][WAVE_ROWS][NWORD];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;            // coverage layout
  const int srow = lane / LPR, ch = lane % LPR;      // store layout
  const int ccol = ch * PPC;                         // first local column of the chunk
  PolyEdge* stage = s_edge[wave];
  unsigned long long burns = 0;
  int mbuf = 0;  // alternating row-mask buffer: one wave_sync per polygon item

  TileQueue tq(a, wave, lane);
  const RZ_CONST uint32_t* toff_c = as_const(a.tile_off);
  const RZ_CONST ItemDesc* pdesc_c = as_const(a.pair_desc);
  const RZ_CONST ItemDesc* desc_c = as_const(a.desc);
  uint32_t pb = 0, pe = 0;
  auto fetch_range = [&](uint32_t tl) {
    if (a.tile_off) {
      pb = toff_c[a.t_lo + tl];
      pe = toff_c[a.t_lo + tl + 1u];
    }
  };
  bool have = tq.begin();
  if (have) fetch_range(tq.t_cur);
  const int32_t strip = wave;

  while (have) {
    const uint32_t t = a.t_lo + tq.pop();
    const StripUnit u = strip_unit(a, t, strip, TW);
    const RZ_CONST JobDev& J = *u.J;
    const T bg = from_bits<T>(J.bg_bits);
    // heavy tile (many items): every polygon mask comes from the pre-pass
    // (a tile is heavy iff its pair range holds at least heavy_min items, so
    // the slot base is only loaded for heavy tiles)
    uint64_t hbase = 0;
    const bool hb_ok = a.heavy_base && !J.scan && pe - pb >= a.heavy_min;
    if (hb_ok) hbase = as_const(a.heavy_base)[t];

    T acc[PACKED ? 1 : NPIX];  // acc[k * PPC + j]: row srow + k*RSTEP, column ccol + j
    uint32_t accw[NW];         // packed: pixel k * PPC + j in word (k * PPC + j) / PPW
    const uint32_t bgw = PACKED ? lane_splat<T>(J.bg_bits) : 0u;
    if constexpr (PACKED) {
#pragma unroll
      for (int i = 0; i < NW; ++i) accw[i] = bgw;
    } else {
#pragma unroll
      for (int j = 0; j < NPIX; ++j) acc[j] = bg;
    }

    auto process_item = [&](const ItemDesc d, uint32_t pi) {
      // wave-uniform skip: the item box misses this strip
      if (d.r1 <= u.Rw || d.r0 >= u.strip_r1 || d.c1 <= u.C0 || d.c0 >= u.tile_c1) return;
      const T v = from_bits<T>(d.value);
      const uint64_t vbits = d.value;
      const uint64_t cov = polygon_mask<NWORD>(a, d, u, hb_ok, hbase, pb, pe, pi, strip, stage, lane);
      if (q < NWORD) s_mask[wave][mbuf][rl][q] = cov;
      wave_sync();
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const uint64_t row_word = s_mask[wave][mbuf][srow + k * RSTEP][ccol >> 6];
        const uint32_t bits = (uint32_t)(row_word >> (ccol & 63)) & ((1u << PPC) - 1u);
        if (__any(bits != 0u)) {  // wave-uniform skip of uncovered row groups
          burns += (unsigned long long)__popc(bits);
          if constexpr (PACKED) {
            const uint32_t vw = lane_splat<T>(vbits);
#pragma unroll
            for (int i = 0; i < PPC / PPW; ++i)
              accw[k * (PPC / PPW) + i] = fold_word<T, FUN>(
                  accw[k * (PPC / PPW) + i], vw, bgw,
                  lane_mask<(int)sizeof(T)>((bits >> (i * PPW)) & ((1u << PPW) - 1u)));
          } else {
#pragma unroll
            for (int j = 0; j < PPC; ++j)
              acc[k * PPC + j] = fold_select<T, FUN>(acc[k * PPC + j], v, bg, (bits >> j) & 1u);
          }
        }
      }
      mbuf ^= 1;  // the next polygon writes the other buffer
    };

    if (!J.scan) {
      // (a heavy tile of a polygon-only group is folded by k_tile_heavy,
      // which runs after this kernel and stores over this strip)
      const uint32_t pend = (hb_ok && a.heavy_split) ? pb : pe;
      for (uint32_t i = pb; i < pend; ++i) process_item(load_desc(pdesc_c + i), i);
    } else if (J.scan_compact && a.scan_box) {
      // scan job of one band: the 8-byte item boxes in strip and tile units,
      // 64 at a time, the next 64 loaded while this batch's hits are
      // processed
      const int64_t ib = J.i_base, ie = J.i_base + J.ni;
      const uint32_t sidx = (uint32_t)u.Rw / WAVE_ROWS, tx = (uint32_t)u.C0 / (uint32_t)TW;
      const uint2* sb = a.scan_box;
      uint2 nbx = ib + lane < ie ? sb[ib + lane] : make_uint2(0xFFFFu, 0u);
      for (int64_t base = ib; base < ie; base += 64) {
        const uint2 bx = nbx;
        const bool valid = base + lane < ie;
        if (base + 64 + lane < ie) nbx = sb[base + 64 + lane];
        uint64_t m = __ballot(valid && scan_box_meets(bx, sidx, tx));
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          process_item(load_desc(desc_c + base + bit), 0u);
        }
      }
    } else {
      // scan job: test every item's box in order (small item counts), 64
      // boxes at a time, the next 64 loaded while this batch's hits are
      // processed
      const int64_t ib = J.i_base, ie = J.i_base + J.ni;
      BoxLayer nbx = ib + lane < ie ? load_box(desc_c + ib + lane) : BoxLayer{0, 0, 0, 0, 0u};
      for (int64_t base = ib; base < ie; base += 64) {
        const BoxLayer bx = nbx;
        const bool valid = base + lane < ie;
        if (base + 64 + lane < ie) nbx = load_box(desc_c + base + 64 + lane);
        const bool hit = valid && (bx.kind_layer >> 2) == u.layer && bx.r0 < u.strip_r1 && bx.r1 > u.Rw &&
                         bx.c0 < u.tile_c1 && bx.c1 > u.C0 && bx.r0 < bx.r1 && bx.c0 < bx.c1;
        uint64_t m = __ballot(hit);
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          process_item(load_desc(desc_c + base + bit), 0u);
        }
      }
    }

    // take the next tile (and its item range) before this strip's stores
    have = tq.advance();
    if (have) fetch_range(tq.t_cur);

    // ---- write the strip once (background + burns)
    store_strip<T, PPC, RPL, RSTEP>(u, srow, ccol, [&](int k) {
      u32x4 vec;
      if constexpr (PACKED) vec = u32x4{accw[k * 4], accw[k * 4 + 1], accw[k * 4 + 2], accw[k * 4 + 3]};
      else __builtin_memcpy(&vec, &acc[k * PPC], 16);
      return vec;
    });
  }
  if (a.burns) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) burns += __shfl_xor(burns, off, 64);
    if (lane == 0 && burns) atomicAdd(&a.burns[blockIdx.x % kBurnSlots], burns);
  }
}

// One line item of a group walked together in k_tile_lp: its records, the
// vertex range its end cell is tested against, kind | banded << 2.
struct GroupItem {
  const uint4* recs;
  int64_t vbeg;
  uint64_t last;
  uint32_t kb;
  uint32_t pad;
};

// Tile kernel of groups with lines or points (polygon items may be mixed
// in), one launch per pixel-function group.  Same work queue, strips, item
// order, background fill and write-once store as k_tile, but the strip's
// accumulator lives in wave-private LDS (16 x TW pixels), so a line or
// point write lands where it belongs instead of going through a per-item
// pass over the whole strip:
//  * first / last / min / max / any give the same result for one or several
//    writes of one value, so a line's cells set LDS mark bits (atomic OR,
//    not waited for) and after the line each lane folds its marked pixels
//    once.  A run of consecutive line items (up to 8 in a 64-wide strip,
//    with the one-vertex points between them) is walked together, its
//    records spread over the lanes and each line marking its own array; one
//    apply then folds each pixel's marks line by line, in item order (a line
//    meets a strip in a few short segments, so one item alone leaves most
//    lanes idle), and the points follow (a point whose cell a line of the
//    group also writes folds between the lines before and after it);
//  * sum and count depend on the number of writes: a line's cells are
//    counted in 16-bit LDS counters, and after the line each lane folds its
//    pixels k times.  Two lines of at most 255 segments each in the strip
//    (with the one-vertex points between them) are walked together, each
//    counting in its own 8-bit counters (a segment writes a pixel once),
//    and fold line 0, the points, then line 1.  With xres != yres
//    (LineWriter + PixelCache, burn_geometry.rs:179-185, 202-208) or
//    all-touched a pixel is written once per line: a mark bit, folded once;
//  * the unordered float sum adds every write with an LDS atomic;
//  * runs of one-vertex point items are folded in geometry order by one
//    lane each, from descriptors loaded 64 at a time (the point's cell is its
//    box, so its vertex is not read);
//  * polygon items fold their coverage masks into the LDS pixels; an
//    all-touched polygon folds its outline cells and its fill once each.
// Segments come as 16-byte records with floored end points (k_prep); a line
// item of many segments over several strips reads only its band of the strip.
// (Folding idempotent writes directly, lane by lane, measured 13 % slower
// than the marks: the read-modify-write per cell serialises the walk.)
template <typename T, int FUN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_tile_lp(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int NWORD = TW / 64;
  constexpr int PPC = 16 / (int)sizeof(T);
  constexpr int LPR = TW / PPC;
  constexpr int RSTEP = 64 / LPR;
  constexpr int RPL = WAVE_ROWS / RSTEP;
  constexpr int NPIXS = WAVE_ROWS * TW;           // pixels of a strip
  constexpr bool COUNTED = FUN == 0 || FUN == 5;  // sum, count: the number of writes matters
  constexpr int PPL = NPIXS / 64;                      // pixels per lane in an apply (16 or 64)
  constexpr int MW = NPIXS / 32;                       // words of one mark array
  // line items walked together with marked writes (fewer in sum and count,
  // which group only their dedup lines this way: there the registers of
  // larger groups spill; 2, 4, 8 and 16 measured)
  constexpr int KGRP = TW == 64 ? (COUNTED ? 4 : 8) : 2;
  constexpr int CNT_WORDS = COUNTED ? NPIXS / 2 : 1;   // u16 counters of one item, 2 per word
  // counted line items walked together: CB-bit counters (NPIXS * CB / 32
  // words per item) in the counter words, for items of at most 2^CB - 1
  // records in the strip (a record writes a pixel at most once).  More
  // items measured slower, whether with more LDS, with the edge stage as
  // counters or with 4-bit counters (4 items): the kernel spills more.
  constexpr int CB = 8;
  constexpr int CW = NPIXS * CB / 32;                  // words of one item's counters
  constexpr int KCNT = COUNTED ? (CNT_WORDS / CW < KGRP ? CNT_WORDS / CW : KGRP) : 1;
  __shared__ PolyEdge s_edge[4][64];
  __shared__ uint64_t s_mask[4][2][WAVE_ROWS][NWORD];
  __shared__ __align__(16) T s_acc[4][NPIXS];
  __shared__ __align__(16) uint32_t s_cnt[4][CNT_WORDS];
  __shared__ uint32_t s_mark[4][KGRP * MW];
  __shared__ GroupItem s_grp[4][KGRP];
  // unordered float sum: the strip holds plain sums and s_wr which pixels have any
  constexpr bool FSUM = FUN == 0 && (sizeof(T) == 4 || sizeof(T) == 8) && (T)0.5 != (T)0;
  __shared__ uint32_t s_wr[4][FSUM ? NPIXS / 32 : 1];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;
  const int srow = lane / LPR, ch = lane % LPR;
  const int ccol = ch * PPC;
  PolyEdge* stage = s_edge[wave];
  T* acc = s_acc[wave];
  uint32_t* cnt = s_cnt[wave];
  uint32_t* mark = s_mark[wave];
  uint32_t* wr = s_wr[wave];
  // (every apply leaves its counters and marks at zero)
  if (COUNTED)
    for (int w = lane; w < CNT_WORDS; w += 64) cnt[w] = 0u;
  for (int w = lane; w < KGRP * MW; w += 64) mark[w] = 0u;
  unsigned long long burns = 0;
  int mbuf = 0;

  TileQueue tq(a, wave, lane);
  const RZ_CONST uint32_t* toff_c = as_const(a.tile_off);
  const RZ_CONST ItemDesc* pdesc_c = as_const(a.pair_desc);
  const RZ_CONST ItemDesc* desc_c = as_const(a.desc);
  uint32_t pb = 0, pe = 0;
  auto fetch_range = [&](uint32_t tl) {
    if (a.tile_off) {
      pb = toff_c[a.t_lo + tl];
      pe = toff_c[a.t_lo + tl + 1u];
    }
  };
  bool have = tq.begin();
  if (have) fetch_range(tq.t_cur);
  const int32_t strip = wave;

  while (have) {
    const uint32_t t = a.t_lo + tq.pop();
    const StripUnit u = strip_unit(a, t, strip, TW);
    const RZ_CONST JobDev& J = *u.J;
    const T bg = from_bits<T>(J.bg_bits);
    uint64_t hbase = 0;
    const bool hb_ok = a.heavy_base && !J.scan && pe - pb >= a.heavy_min;
    if (hb_ok) hbase = as_const(a.heavy_base)[t];
    const bool dedup_job = J.dedup_lines != 0;
    const bool at_job = J.all_touched != 0;

    // unordered float sum (no dedup: a line's pixel counts once there):
    // lines and points add as they land (RZ_SUM_UNORDERED)
    const bool unord = FSUM && J.sum_unordered && !dedup_job && !at_job;
    // the strip starts as the background (unordered: as empty sums)
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      T* p = acc + (srow + k * RSTEP) * TW + ccol;
#pragma unroll
      for (int j = 0; j < PPC; ++j) p[j] = unord ? (T)0 : bg;
    }
    if (FSUM && unord)
      for (int w = lane; w < NPIXS / 32; w += 64) wr[w] = 0u;
    wave_sync();
    // an unordered write of v at strip pixel idx (any lane, any order)
    auto add1 = [&](uint32_t idx, T v) {
      if constexpr (FSUM) {
        atomicAdd(&acc[idx], v);
        atomicOr(&wr[idx >> 5], 1u << (idx & 31u));
      }
    };

    // one write of value v at strip pixel idx, by this lane alone
    auto fold1 = [&](uint32_t idx, T v) { fold<T, FUN>(acc[idx], v, bg, 1u); };

    // each lane folds its marked pixels once and clears their marks
    auto apply_marks = [&](T v) {
      wave_sync();
#pragma unroll
      for (int c0 = 0; c0 < PPL; c0 += 16) {
        const int base = lane * PPL + c0;
        const uint32_t bits = (mark[base >> 5] >> (base & 31)) & 0xFFFFu;
        if (bits) {
          burns += (unsigned long long)__popc(bits);
          T px[16];
          __builtin_memcpy(px, acc + base, sizeof(px));
#pragma unroll
          for (int j = 0; j < 16; ++j) px[j] = fold_select<T, FUN>(px[j], v, bg, (bits >> j) & 1u);
          __builtin_memcpy(acc + base, px, sizeof(px));
        }
      }
      wave_sync();
      for (int w = lane; w < NPIXS / 32; w += 64) mark[w] = 0u;
      wave_sync();
    };

    // all-touched: the cells of the kept segment from local vertex lv that
    // fall in this strip, as marks in mk (the walker stops past the strip)
    auto walk_at = [&](int64_t lv, uint32_t* mk) {
      double x0, y0, x1, y1;
      if (!kept_segment(J, lv, x0, y0, x1, y1)) return;
      // (cells lie within a pixel of the floors of the end points)
      if (floor(fmax(x0, x1)) + 1.0 < (double)u.C0 || floor(fmin(x0, x1)) - 1.0 >= (double)u.tile_c1 ||
          floor(fmax(y0, y1)) + 1.0 < (double)u.Rw || floor(fmin(y0, y1)) - 1.0 >= (double)u.strip_r1)
        return;
      const bool ok = walk_all_touched(x0, y0, x1, y1, J.ncols, J.nrows, u.C0, u.tile_c1, u.Rw, u.strip_r1,
                                       [&](int64_t x, int64_t y) {
                                         const uint32_t idx = (uint32_t)(y - u.Rw) * TW + (uint32_t)(x - u.C0);
                                         atomicOr(&mk[idx >> 5], 1u << (idx & 31u));
                                       });
      if (!ok) atomicOr(a.status, 64u);
    };

    // An all-touched polygon (burn_geometry.rs:219-238): its outline cells
    // and its fill, each pixel folded once.  Idempotent functions fold the
    // union; sum and count (LineWriter + FillWriter over the outline's
    // PixelCache) fold the outline cells once and the fill pixels the cache
    // does not contain: inside the cache window that is every fill pixel off
    // the outline, again the union; a fill pixel outside the window is
    // skipped when the cache bit its wrapped index lands on is set.
    auto polygon_at = [&](const ItemDesc& d, uint32_t pi) {
      const T v = from_bits<T>(d.value);
      for (int64_t e = lane; e < d.nv; e += 64) walk_at(d.vbeg + e - J.v_base, mark);
      const uint64_t cov = polygon_mask<NWORD>(a, d, u, hb_ok, hbase, pb, pe, pi, strip, stage, lane);
      wave_sync();
      if (q < NWORD) {
        uint64_t m = cov;
        if (COUNTED && (d.flags & kItemAtCache) && m && a.at_bits && a.at_base[d.last + 1] <= a.at_cap) {
          CacheWin cw;
          item_cache(a.item_win, (int64_t)d.last, cw);
          const uint64_t length = cw.cap / cw.width;
          const unsigned long long* bits = a.at_bits + a.at_base[d.last];
          const int64_t y = (int64_t)u.Rw + rl, xb = (int64_t)u.C0 + q * 64;
          for (uint64_t mm = m; mm; mm &= mm - 1) {
            const int b = __builtin_ctzll(mm);
            const int64_t x = xb + b;
            if ((uint64_t)x - (uint64_t)cw.x0 < cw.width && (uint64_t)y - (uint64_t)cw.y0 < length) continue;
            const uint64_t ci = cache_index(cw, x, y);
            if (ci < cw.cap && ((bits[ci >> 6] >> (ci & 63)) & 1ull)) m &= ~(1ull << b);
          }
        }
        const int idx = rl * TW + q * 64;
        mark[idx >> 5] |= (uint32_t)m;
        mark[(idx >> 5) + 1] |= (uint32_t)(m >> 32);
      }
      apply_marks(v);
    };

    auto polygon = [&](const ItemDesc& d, uint32_t pi) {
      const T v = from_bits<T>(d.value);
      const uint64_t cov = polygon_mask<NWORD>(a, d, u, hb_ok, hbase, pb, pe, pi, strip, stage, lane);
      if (q < NWORD) s_mask[wave][mbuf][rl][q] = cov;
      wave_sync();
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int r = srow + k * RSTEP;
        const uint64_t row_word = s_mask[wave][mbuf][r][ccol >> 6];
        const uint32_t bits = (uint32_t)(row_word >> (ccol & 63)) & ((1u << PPC) - 1u);
        if (__any(bits != 0u)) {
          burns += (unsigned long long)__popc(bits);
          T* p = acc + r * TW + ccol;
          T px[PPC];
          __builtin_memcpy(px, p, 16);
          if (FSUM && unord) {
#pragma unroll
            for (int j = 0; j < PPC; ++j) px[j] = ((bits >> j) & 1u) ? px[j] + v : px[j];
            const int idx = r * TW + ccol;
            if (bits) atomicOr(&wr[idx >> 5], bits << (idx & 31));
          } else {
#pragma unroll
            for (int j = 0; j < PPC; ++j) px[j] = fold_select<T, FUN>(px[j], v, bg, (bits >> j) & 1u);
          }
          __builtin_memcpy(p, px, 16);
        }
      }
      wave_sync();
      mbuf ^= 1;
    };

    // a line or a multi-point item: its segment / cell records, lanes over
    // records (chunks of 65,535 records keep a pixel's count in 16 bits)
    // a line or a multi-point item: its records [recs, recs + n_r) (a line's
    // band of this strip, or the item's vertices), lanes over records; `pre`
    // is this lane's record of the first 64, already loaded (chunks of
    // 65,535 records keep a pixel's count in 16 bits)
    auto lines = [&](const ItemDesc& d, const uint4* recs, uint32_t n_r, bool banded, uint4 pre) {
      const int32_t kind = (int32_t)(d.kind_layer & 3u);
      const T v = from_bits<T>(d.value);
      const bool dedup = kind == 1 && (dedup_job || at_job);
      // a write is a mark (one fold per pixel: idempotent functions, or a
      // line's first write per pixel with xres != yres or all-touched) or a
      // count; both are LDS atomics whose result is not waited for
      const bool marked = !COUNTED || dedup;
      // a write at strip pixel idx (emit_idx: the clipped walk's cells,
      // inside by construction) or at a cell
      auto emit_idx = [&](uint32_t idx) {
        if (FSUM && unord) {
          ++burns;
          add1(idx, v);
        } else if (marked) {
          atomicOr(&mark[idx >> 5], 1u << (idx & 31u));
        } else {
          atomicAdd(&cnt[idx >> 1], 1u << ((idx & 1u) * 16u));
        }
      };
      auto emit = [&](auto cx, auto cy) {
        if (cx < u.C0 || cx >= u.tile_c1 || cy < u.Rw || cy >= u.strip_r1) return;
        emit_idx((uint32_t)(cy - u.Rw) * TW + (uint32_t)(cx - u.C0));
      };
      // end cell of the last kept edge of an open line (burners.rs:86-89)
      auto end_cell = [&](const uint4& r, uint32_t e) {
        return banded ? (r.w & kLsFlag) != 0
                      : ((uint64_t)(d.vbeg + (int64_t)e + 1) == d.last && !(r.w & kLsFlag));
      };
      for (uint32_t cs = 0; cs < n_r; cs += 65535u) {
        const uint32_t ce = min(n_r, cs + 65535u);
        bool far = false;
        for (uint32_t e = cs + (uint32_t)lane; e < ce; e += 64u) {
          const uint4 r = e < 64u ? pre : recs[e];
          if (r.x == kLsNone) continue;
          if (kind == 2) {  // a vertex of a multi-point: its cell
            emit((int32_t)r.x, (int32_t)r.y);
            continue;
          }
          if (at_job) {  // the all-touched walker, from the world coordinates
            walk_at(d.vbeg + (int64_t)e - J.v_base, mark);
            continue;
          }
          if (r.x == kLsFar) {
            far = true;
            continue;
          }
          const int32_t ix0 = (int32_t)((int64_t)r.x - kLsBias), iy0 = (int32_t)((int64_t)r.y - kLsBias);
          const int32_t ix1 = (int32_t)((int64_t)r.z - kLsBias);
          const int32_t iy1 = (int32_t)((int64_t)(r.w & ~kLsFlag) - kLsBias);
          walk_window32_idx(ix0, iy0, ix1, iy1, u.C0, u.tile_c1, u.Rw, u.strip_r1, TW, emit_idx);
          if (end_cell(r, e)) emit(ix1, iy1);
        }
        if (__any(far)) {
          // segments beyond +-2^30 pixels (never banded): the end points again
          // from the world coordinates (edges.rs:119-122), walked in 128 bits
          for (uint32_t e = cs + (uint32_t)lane; e < ce; e += 64u) {
            const uint4 r = recs[e];
            if (r.x != kLsFar) continue;
            const int64_t lv = d.vbeg + (int64_t)e - J.v_base;
            auto px_x = [&](double X) { return isnan(X) ? X : (X - J.xmin) / J.xres; };
            auto px_y = [&](double Y) { return isnan(Y) ? Y : (J.ymax - Y) / J.yres; };
            const int64_t ix0 = sat_isize(floor(px_x(J.xy[2 * lv]))), iy0 = sat_isize(floor(px_y(J.xy[2 * lv + 1])));
            const int64_t ix1 = sat_isize(floor(px_x(J.xy[2 * lv + 2]))), iy1 = sat_isize(floor(px_y(J.xy[2 * lv + 3])));
            const int64_t lim = (int64_t)1 << 60;
            if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim &&
                iy1 > -lim && iy1 < lim)
              walk_window<__int128>(ix0, iy0, ix1, iy1, u.C0, u.tile_c1, u.Rw, u.strip_r1, emit);
            else
              atomicOr(a.status, 1u);  // pixel coordinates beyond +-2^60: unsupported
            if (end_cell(r, e)) emit(ix1, iy1);
          }
        }
        if (COUNTED && !marked && !unord) {
          // each lane folds its PPL pixels k times each and clears their counters
          wave_sync();
#pragma unroll
          for (int c0 = 0; c0 < PPL; c0 += 16) {
            const int base = lane * PPL + c0;
            uint32_t w[8];
            __builtin_memcpy(w, cnt + base / 2, 32);
            if (w[0] | w[1] | w[2] | w[3] | w[4] | w[5] | w[6] | w[7]) {
              T px[16];
              __builtin_memcpy(px, acc + base, sizeof(px));
#pragma unroll
              for (int j = 0; j < 16; ++j) {
                const uint32_t k = (w[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
                burns += k;
                fold<T, FUN>(px[j], v, bg, k);
              }
              __builtin_memcpy(acc + base, px, sizeof(px));
              const uint32_t z[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
              __builtin_memcpy(cnt + base / 2, z, 32);
            }
          }
          wave_sync();
        }
      }
      if (marked && !unord) apply_marks(v);
    };

    // the records of a line or multi-point item in this strip: base, count,
    // whether they are the line's band (each lane may ask for its own item)
    auto records = [&](const ItemDesc& d, const uint4*& recs, uint32_t& n_r) -> bool {
      if ((d.kind_layer & 3u) == 1u && d.band) {
        const uint32_t bi = d.band - 1u + (uint32_t)(u.Rw / kBandRows - d.r0 / kBandRows);
        const uint64_t b0 = a.band_off[bi];
        recs = a.band_segs + b0;
        n_r = (uint32_t)(a.band_off[bi + 1u] - b0);
        return true;
      }
      recs = a.lseg + d.vbeg;
      n_r = (uint32_t)d.nv;
      return false;
    };
    auto first_chunk = [&](const uint4* recs, uint32_t n_r) {
      return (uint32_t)lane < n_r ? recs[lane] : make_uint4(kLsNone, 0u, 0u, 0u);
    };

    // The items of lanes G, consecutive in item order: the line items L (two
    // or more, at most 65,535 records each) and the one-vertex points
    // between them (cell pidx, value bits pv in their lanes).  Each lane's
    // own item is dl with records [rp, rp + rn).  The lanes take the lines'
    // records in turn, line k writing mark array k (cmode false) or CB-bit
    // counter array k (cmode: sum and count without dedup, at most
    // 2^CB - 1 records per line, and a record writes a pixel at most once);
    // then each lane folds its pixels' writes line by line, and the points
    // are folded in order after them, except that a point whose cell a line
    // of the group also writes is folded between the lines before it and
    // those after it.  A counted group (two lines) always folds line 0, the
    // points between them, then line 1.
    auto line_group = [&](uint64_t G, uint64_t L, const ItemDesc& dl, const uint4* rp, uint32_t rn, bool bnd,
                          bool cmode, uint32_t pidx, uint64_t pv) {
      GroupItem* gi = s_grp[wave];
      const int ng = __popcll(L);
      const uint64_t P = G & ~L;
      // the points of lanes pm, one at a time in item order
      auto fold_points = [&](uint64_t pm) {
        for (uint64_t m = pm; m; m &= m - 1ull) {
          const int b = __builtin_ctzll(m);
          const uint32_t idx = (uint32_t)__builtin_amdgcn_readlane((int)pidx, b);
          const uint64_t vb = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pv >> 32), b) << 32) |
                              (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pv, b);
          if (lane == 0) {
            ++burns;
            fold1(idx, from_bits<T>(vb));
          }
        }
      };
      if ((L >> lane) & 1ull) {
        GroupItem it;
        it.recs = rp;
        it.vbeg = dl.vbeg;
        it.last = dl.last;
        it.kb = (dl.kind_layer & 3u) | (bnd ? 4u : 0u);
        it.pad = 0u;
        gi[__popcll(L & ((1ull << lane) - 1ull))] = it;
      }
      // record offsets and values of the lines (uniform)
      uint32_t pre[KGRP];
      T vk[KGRP];
      uint32_t tot = 0;
      uint64_t g = L;
#pragma unroll
      for (int k = 0; k < KGRP; ++k) {
        pre[k] = tot;
        vk[k] = bg;
        if (g) {
          const int pk = __builtin_ctzll(g);
          g &= g - 1ull;
          tot += (uint32_t)__builtin_amdgcn_readlane((int)rn, pk);
          const uint64_t vb = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(dl.value >> 32), pk) << 32) |
                              (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)dl.value, pk);
          vk[k] = from_bits<T>(vb);
        }
      }
      // counter array of item k
      auto carr = [&](int k) { return cnt + k * CW; };
      wave_sync();
      // record e of the group: its item k and index ei in that item
      auto locate = [&](uint32_t e, int& k, uint32_t& ei) {
        uint32_t b = 0;
        k = 0;
#pragma unroll
        for (int j = 1; j < KGRP; ++j)
          if (e >= pre[j]) {
            k = j;
            b = pre[j];
          }
        ei = e - b;
      };
      bool far = false;
      for (uint32_t e = (uint32_t)lane; e < tot; e += 64u) {
        int k;
        uint32_t ei;
        locate(e, k, ei);
        const GroupItem it = gi[k];
        uint32_t* mk = mark + k * MW;
        uint32_t* ck = carr(k);
        auto emit_idx = [&](uint32_t idx) {
          if (COUNTED && cmode) atomicAdd(&ck[idx / (32 / CB)], 1u << ((idx % (32 / CB)) * CB));
          else atomicOr(&mk[idx >> 5], 1u << (idx & 31u));
        };
        auto emit = [&](auto cx, auto cy) {
          if (cx < u.C0 || cx >= u.tile_c1 || cy < u.Rw || cy >= u.strip_r1) return;
          emit_idx((uint32_t)(cy - u.Rw) * TW + (uint32_t)(cx - u.C0));
        };
        const uint4 r = it.recs[ei];
        if (r.x == kLsNone) continue;
        if ((it.kb & 3u) == 2u) {  // a vertex of a multi-point: its cell
          emit((int32_t)r.x, (int32_t)r.y);
          continue;
        }
        if (at_job) {
          walk_at(it.vbeg + (int64_t)ei - J.v_base, mk);
          continue;
        }
        if (r.x == kLsFar) {
          far = true;
          continue;
        }
        const int32_t ix0 = (int32_t)((int64_t)r.x - kLsBias), iy0 = (int32_t)((int64_t)r.y - kLsBias);
        const int32_t ix1 = (int32_t)((int64_t)r.z - kLsBias);
        const int32_t iy1 = (int32_t)((int64_t)(r.w & ~kLsFlag) - kLsBias);
        walk_window32_idx(ix0, iy0, ix1, iy1, u.C0, u.tile_c1, u.Rw, u.strip_r1, TW, emit_idx);
        const bool endc = (it.kb & 4u) ? (r.w & kLsFlag) != 0
                                       : ((uint64_t)(it.vbeg + (int64_t)ei + 1) == it.last && !(r.w & kLsFlag));
        if (endc) emit(ix1, iy1);
      }
      if (__any(far)) {
        for (uint32_t e = (uint32_t)lane; e < tot; e += 64u) {
          int k;
          uint32_t ei;
          locate(e, k, ei);
          const GroupItem it = gi[k];
          const uint4 r = it.recs[ei];
          if (r.x != kLsFar) continue;
          uint32_t* mk = mark + k * MW;
          uint32_t* ck = carr(k);
          auto emit = [&](auto cx, auto cy) {
            if (cx < u.C0 || cx >= u.tile_c1 || cy < u.Rw || cy >= u.strip_r1) return;
            const uint32_t idx = (uint32_t)(cy - u.Rw) * TW + (uint32_t)(cx - u.C0);
            if (COUNTED && cmode) atomicAdd(&ck[idx / (32 / CB)], 1u << ((idx % (32 / CB)) * CB));
            else atomicOr(&mk[idx >> 5], 1u << (idx & 31u));
          };
          const int64_t lv = it.vbeg + (int64_t)ei - J.v_base;
          auto px_x = [&](double X) { return isnan(X) ? X : (X - J.xmin) / J.xres; };
          auto px_y = [&](double Y) { return isnan(Y) ? Y : (J.ymax - Y) / J.yres; };
          const int64_t ix0 = sat_isize(floor(px_x(J.xy[2 * lv]))), iy0 = sat_isize(floor(px_y(J.xy[2 * lv + 1])));
          const int64_t ix1 = sat_isize(floor(px_x(J.xy[2 * lv + 2]))), iy1 = sat_isize(floor(px_y(J.xy[2 * lv + 3])));
          const int64_t lim = (int64_t)1 << 60;
          if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim && iy1 > -lim &&
              iy1 < lim)
            walk_window<__int128>(ix0, iy0, ix1, iy1, u.C0, u.tile_c1, u.Rw, u.strip_r1, emit);
          else
            atomicOr(a.status, 1u);  // pixel coordinates beyond +-2^60: unsupported
          const bool endc = (it.kb & 4u) ? (r.w & kLsFlag) != 0
                                         : ((uint64_t)(it.vbeg + (int64_t)ei + 1) == it.last && !(r.w & kLsFlag));
          if (endc) emit(ix1, iy1);
        }
      }
      wave_sync();
      // the writes of line k (uniform) at this lane's pixels, folded
      auto apply_one = [&](int k) {
        T v = bg;
#pragma unroll
        for (int j = 0; j < KGRP; ++j)
          if (j == k) v = vk[j];
#pragma unroll
        for (int c0 = 0; c0 < PPL; c0 += 16) {
          const int base = lane * PPL + c0;
          if (COUNTED && cmode) {
            constexpr int NW = 16 * CB / 32;
            uint32_t w[NW];
            __builtin_memcpy(w, carr(k) + base / (32 / CB), sizeof(w));
            uint32_t any = 0u;
#pragma unroll
            for (int x = 0; x < NW; ++x) any |= w[x];
            if (any) {
              T px[16];
              __builtin_memcpy(px, acc + base, sizeof(px));
#pragma unroll
              for (int j = 0; j < 16; ++j) {
                const uint32_t c = (w[j / (32 / CB)] >> ((j % (32 / CB)) * CB)) & ((1u << CB) - 1u);
                burns += c;
                fold<T, FUN>(px[j], v, bg, c);
              }
              __builtin_memcpy(acc + base, px, sizeof(px));
            }
          } else {
            const uint32_t b = (mark[k * MW + (base >> 5)] >> (base & 31)) & 0xFFFFu;
            if (b) {
              T px[16];
              __builtin_memcpy(px, acc + base, sizeof(px));
              burns += (unsigned long long)__popc(b);
#pragma unroll
              for (int j = 0; j < 16; ++j) px[j] = fold_select<T, FUN>(px[j], v, bg, (b >> j) & 1u);
              __builtin_memcpy(acc + base, px, sizeof(px));
            }
          }
        }
      };
      if (COUNTED && cmode) {
        // two lines with the points between them: line 0, the points, line 1
        apply_one(0);
        wave_sync();
        fold_points(P);
        wave_sync();
        apply_one(1);
        wave_sync();
        for (int w = lane; w < ng * CW; w += 64) cnt[w] = 0u;
        wave_sync();
        return;
      }
      // the points whose cells a line of the group also writes (C): each is
      // folded after the lines before it and before those after it; the
      // others share no cell with a line and are folded last
      bool hit = false;
      if ((P >> lane) & 1ull) {
#pragma unroll
        for (int k = 0; k < KGRP; ++k) {
          if (k >= ng) break;
          hit |= ((mark[k * MW + (pidx >> 5)] >> (pidx & 31u)) & 1u) != 0u;
        }
      }
      const uint64_t C = __ballot(hit);
      // each lane folds its pixels' marks of lines [lo, hi), line by line (a
      // wave-wide loop over the lines would fold every line in every lane)
      auto apply_range = [&](int lo, int hi) {
#pragma unroll
        for (int c0 = 0; c0 < PPL; c0 += 16) {
          const int base = lane * PPL + c0;
          uint32_t bk[KGRP];
          uint32_t kmask = 0u;  // the lines with marks in this lane's pixels
#pragma unroll
          for (int k = 0; k < KGRP; ++k) {
            bk[k] = (k >= lo && k < hi) ? (mark[k * MW + (base >> 5)] >> (base & 31)) & 0xFFFFu : 0u;
            kmask |= bk[k] ? 1u << k : 0u;
          }
          if (kmask) {
            T px[16];
            __builtin_memcpy(px, acc + base, sizeof(px));
            while (kmask) {
              const int k = __builtin_ctz(kmask);
              kmask &= kmask - 1u;
              uint32_t b = 0u;
              T v = bg;
#pragma unroll
              for (int j = 0; j < KGRP; ++j)
                if (j == k) {
                  b = bk[j];
                  v = vk[j];
                }
              burns += (unsigned long long)__popc(b);
#pragma unroll
              for (int j = 0; j < 16; ++j) px[j] = fold_select<T, FUN>(px[j], v, bg, (b >> j) & 1u);
            }
            __builtin_memcpy(acc + base, px, sizeof(px));
          }
        }
      };
      int lo = 0;
      for (uint64_t m = C; m; m &= m - 1ull) {
        const int b = __builtin_ctzll(m);
        const int r = __popcll(L & ((1ull << b) - 1ull));  // the lines before the point
        if (r > lo) {
          apply_range(lo, r);
          wave_sync();
          lo = r;
        }
        fold_points(1ull << b);
        wave_sync();
      }
      apply_range(lo, ng);
      wave_sync();
      for (int w = lane; w < ng * MW; w += 64) mark[w] = 0u;
      wave_sync();
      fold_points(P & ~C);
    };

    auto process_item = [&](const ItemDesc& d, uint32_t pi) {
      if (d.r1 <= u.Rw || d.r0 >= u.strip_r1 || d.c1 <= u.C0 || d.c0 >= u.tile_c1) return;
      if ((d.kind_layer & 3u) == 0u) {
        if (at_job) polygon_at(d, pi);
        else polygon(d, pi);
      } else {
        const uint4* recs;
        uint32_t n_r;
        const bool banded = records(d, recs, n_r);
        lines(d, recs, n_r, banded, first_chunk(recs, n_r));
      }
    };

    if (!J.scan) {
      // descriptors 64 at a time, one per lane: which items meet the strip,
      // the cells and values of one-vertex points (their box), and the
      // record range of every line (its band bounds, loaded by all lanes at
      // once).  Items are then taken in order: a run of points is folded by
      // one lane from the lanes' copies, a polygon or line is reloaded
      // uniformly and processed by the wave, and the next line's first
      // records are loaded while the current one is walked.
      for (uint32_t i0 = pb; i0 < pe; i0 += 64u) {
        const uint32_t nb = min(64u, pe - i0);
        bool pt = false, ln = false, po = false, bnd = false, grp = false;
        uint32_t pidx = 0, rn = 0;
        uint64_t pv = 0;
        const uint4* rp = nullptr;
        ItemDesc dl;
        __builtin_memset(&dl, 0, sizeof(dl));
        if ((uint32_t)lane < nb) {
          dl = load_desc(pdesc_c + i0 + (uint32_t)lane);
          const uint32_t kind = dl.kind_layer & 3u;
          const bool vis = !(dl.r1 <= u.Rw || dl.r0 >= u.strip_r1 || dl.c1 <= u.C0 || dl.c0 >= u.tile_c1);
          pt = vis && kind == 2u && dl.nv == 1;
          po = vis && kind == 0u;
          ln = vis && !pt && !po;
          if (pt) {
            pidx = (uint32_t)(dl.r0 - u.Rw) * TW + (uint32_t)(dl.c0 - u.C0);
            pv = dl.value;
          }
          if (ln) bnd = records(dl, rp, rn);
          // lines walked in groups: those whose writes are marks, or counted
          // ones (sum, count without dedup) of at most 2^CB - 1 records here
          grp = ln && rn <= 0xFFFFu &&
                (!COUNTED ? true : (dedup_job || at_job) ? kind == 1u : (!unord && rn < (1u << CB)));
        }
        const uint64_t ptm = __ballot(pt), lnm = __ballot(ln), grm = __ballot(grp);
        uint64_t todo = ptm | lnm | __ballot(po);
        // the first records of line item q (a uniform lane index)
        auto prefetch = [&](int q) {
          const uint4* r = (const uint4*)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)((uint64_t)rp >> 32), q) << 32) |
                                          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)rp, q));
          return first_chunk(r, (uint32_t)__builtin_amdgcn_readlane((int)rn, q));
        };
        uint64_t lrest = lnm;
        uint4 pre = make_uint4(kLsNone, 0u, 0u, 0u);
        if (lrest) pre = prefetch(__builtin_ctzll(lrest));
        while (todo) {
          const int p = __builtin_ctzll(todo);
          if ((ptm >> p) & 1ull) {
            // the run of points up to the next line or polygon, in order
            const uint64_t other = todo & ~ptm & (~0ull << p);
            const uint64_t run = todo & ptm & (other ? ((1ull << __builtin_ctzll(other)) - 1ull) : ~0ull);
            if (FSUM && unord) {
              if ((run >> lane) & 1ull) {
                ++burns;
                add1(pidx, from_bits<T>(pv));
              }
              todo &= ~run;
              continue;
            }
            uint64_t m = run;
            while (m) {
              const int b = __builtin_ctzll(m);
              m &= m - 1;
              const uint32_t idx = (uint32_t)__builtin_amdgcn_readlane((int)pidx, b);
              const uint64_t vb = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(pv >> 32), b) << 32) |
                                  (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pv, b);
              if (lane == 0) {
                ++burns;
                fold1(idx, from_bits<T>(vb));
              }
            }
            todo &= ~run;
            continue;
          }
          if ((grm >> p) & 1ull) {
            // the run of groupable lines and one-vertex points (not in dedup
            // groups of sum and count) up to the next other item, to its
            // KGRP-th line
            const uint64_t mem = (COUNTED && (dedup_job || at_job)) ? grm : (grm | ptm);
            const uint64_t other = todo & ~mem & (~0ull << p);
            const uint64_t run = todo & mem & (other ? ((1ull << __builtin_ctzll(other)) - 1ull) : ~0ull);
            const bool cmode = COUNTED && !dedup_job && !at_job;
            uint64_t lr = run & grm, L = 0;
#pragma unroll
            for (int k = 0; k < KGRP; ++k) {
              if (!cmode || k < KCNT) {
                L |= lr & (0ull - lr);
                lr &= lr - 1ull;
              }
            }
            if (L & (L - 1ull)) {
              const int hi = 63 - __builtin_clzll(L);
              const uint64_t G = run & (hi == 63 ? ~0ull : ((2ull << hi) - 1ull));
              line_group(G, L, dl, rp, rn, bnd, cmode, pidx, pv);
              todo &= ~G;
              lrest &= ~G;
              if (lrest) pre = prefetch(__builtin_ctzll(lrest));
              continue;
            }
          }
          todo &= todo - 1;
          // (reloaded with one uniform load: broadcasting lane p's copy kept
          // every descriptor word live across the loop, measured slower)
          const ItemDesc d = load_desc(pdesc_c + i0 + (uint32_t)p);
          if ((lnm >> p) & 1ull) {
            const uint4 cur = pre;
            lrest &= lrest - 1;
            if (lrest) pre = prefetch(__builtin_ctzll(lrest));
            const uint4* recs = (const uint4*)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)((uint64_t)rp >> 32), p) << 32) |
                                               (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)rp, p));
            const uint32_t n_r = (uint32_t)__builtin_amdgcn_readlane((int)rn, p);
            const bool banded = __builtin_amdgcn_readlane((int)bnd, p) != 0;
            lines(d, recs, n_r, banded, cur);
          } else if (at_job) {
            polygon_at(d, i0 + (uint32_t)p);
          } else {
            polygon(d, i0 + (uint32_t)p);
          }
        }
      }
    } else if (J.scan_compact && a.scan_box) {
      // scan job of one band: the 8-byte item boxes in strip and tile units,
      // 64 at a time, the next 64 loaded while this batch's hits are
      // processed
      const int64_t ib = J.i_base, ie = J.i_base + J.ni;
      const uint32_t sidx = (uint32_t)u.Rw / WAVE_ROWS, tx = (uint32_t)u.C0 / (uint32_t)TW;
      const uint2* sb = a.scan_box;
      uint2 nbx = ib + lane < ie ? sb[ib + lane] : make_uint2(0xFFFFu, 0u);
      for (int64_t base = ib; base < ie; base += 64) {
        const uint2 bx = nbx;
        const bool valid = base + lane < ie;
        if (base + 64 + lane < ie) nbx = sb[base + 64 + lane];
        uint64_t m = __ballot(valid && scan_box_meets(bx, sidx, tx));
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          process_item(load_desc(desc_c + base + bit), 0u);
        }
      }
    } else {
      // scan job: test every item's box in order (small item counts), 64
      // boxes at a time, the next 64 loaded while this batch's hits are
      // processed
      const int64_t ib = J.i_base, ie = J.i_base + J.ni;
      BoxLayer nbx = ib + lane < ie ? load_box(desc_c + ib + lane) : BoxLayer{0, 0, 0, 0, 0u};
      for (int64_t base = ib; base < ie; base += 64) {
        const BoxLayer bx = nbx;
        const bool valid = base + lane < ie;
        if (base + 64 + lane < ie) nbx = load_box(desc_c + base + 64 + lane);
        const bool hit = valid && (bx.kind_layer >> 2) == u.layer && bx.r0 < u.strip_r1 && bx.r1 > u.Rw &&
                         bx.c0 < u.tile_c1 && bx.c1 > u.C0 && bx.r0 < bx.r1 && bx.c0 < bx.c1;
        uint64_t m = __ballot(hit);
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          process_item(load_desc(desc_c + base + bit), 0u);
        }
      }
    }

    have = tq.advance();
    if (have) fetch_range(tq.t_cur);

    // ---- write the strip once from LDS
    wave_sync();
    store_strip<T, PPC, RPL, RSTEP>(u, srow, ccol, [&](int k) {
      const int idx = (srow + k * RSTEP) * TW + ccol;
      T px[PPC];
      __builtin_memcpy(px, acc + idx, 16);
      if (FSUM && unord) {  // pixels with no write keep the background
        const uint32_t w = wr[idx >> 5] >> (idx & 31);
#pragma unroll
        for (int j = 0; j < PPC; ++j) px[j] = ((w >> j) & 1u) ? px[j] : bg;
      }
      u32x4 vec;
      __builtin_memcpy(&vec, px, 16);
      return vec;
    });
    wave_sync();  // the next strip's background overwrites these pixels
  }
  if (a.burns) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) burns += __shfl_xor(burns, off, 64);
    if (lane == 0 && burns) atomicAdd(&a.burns[blockIdx.x % kBurnSlots], burns);
  }
}

// Heavy tiles of a polygon-only group (k_tile skips them): one wave per
// (heavy tile, strip) unit, taken from a work counter.  Every item's strip
// mask was written by the pre-pass (k_cover), strip-major per tile.  Masks
// and descriptors go through LDS kBlk items at a time: the next block is
// loaded into registers while the current block is folded from LDS, then
// stored to LDS, so the per-item loads of a tile with thousands of items
// overlap instead of serialising.  Same folds, geometry order and stores as
// k_tile.
template <typename T, int FUN>
__global__ __launch_bounds__(256) void k_tile_heavy(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int NWORD = TW / 64;
  constexpr int PPC = 16 / (int)sizeof(T);
  constexpr int LPR = TW / PPC;
  constexpr int RSTEP = 64 / LPR;
  constexpr int RPL = WAVE_ROWS / RSTEP;
  constexpr int NPIX = RPL * PPC;
  constexpr bool PACKED = sizeof(T) <= 2;  // pixels packed in 32-bit words (fold_word)
  constexpr int NW = PACKED ? NPIX * (int)sizeof(T) / 4 : 1;
  constexpr int PPW = 4 / (int)sizeof(T);
  constexpr int SW = 16 * NWORD;          // words of one item's strip mask
  constexpr uint32_t kBlk = 256 / SW;     // items per block (16 or 4)
  __shared__ uint64_t s_mk[4][256];
  __shared__ ItemDesc s_ds[4][16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int srow = lane / LPR, ch = lane % LPR, ccol = ch * PPC;
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST int64_t* tend_c = as_const(a.job_tend);
  // this launch's heavy tiles: heavy_list[h0 .. h1), four units (strips) each
  const uint32_t h0 = a.heavy_rank[a.t_lo], units = 4u * (a.heavy_rank[a.t_hi] - h0);
  // units per counter grab: a few per wave of the grid, at most 16
  const uint32_t grab = max(1u, min(16u, units / (gridDim.x * 4u * 4u)));
  uint64_t* mk = s_mk[wave];
  ItemDesc* ds = s_ds[wave];
  unsigned long long burns = 0;
  uint32_t u = 0, u_end = 0;
  for (;;) {
    if (u == u_end) {
      uint32_t g = 0;
      if (lane == 0) g = atomicAdd(a.heavy_ctr, grab);
      u = (uint32_t)__builtin_amdgcn_readfirstlane((int)g);
      u_end = min(u + grab, units);
      if (u >= units) break;
    }
    const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.heavy_list[h0 + (u >> 2)]);
    const int32_t strip = (int32_t)(u & 3u);
    ++u;
    const RZ_CONST JobDev& J = jobs_c[find_job(tend_c, a.n_jobs, (int64_t)t)];
    const int32_t nrows = J.nrows, ncols = J.ncols;
    const T bg = from_bits<T>(J.bg_bits);
    const uint32_t tj = t - J.tile_base;
    const uint32_t layer = tj / J.tiles_per_layer;
    const uint32_t tl = tj - layer * J.tiles_per_layer;
    const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
    const int32_t C0 = tx * TW, Rw = ty * TILE_H + strip * WAVE_ROWS;
    const int32_t strip_r1 = min(Rw + WAVE_ROWS, nrows), tile_c1 = min(C0 + TW, ncols);
    const uint32_t pb = a.tile_off[t], cnt = a.tile_off[t + 1] - pb;
    const uint64_t* sbase = a.masks + (a.heavy_base[t] * 4u + (uint64_t)strip * cnt) * a.slot_words;
    const ItemDesc* dbase = a.pair_desc + pb;

    T acc[PACKED ? 1 : NPIX];
    uint32_t accw[NW];
    const uint32_t bgw = PACKED ? lane_splat<T>(J.bg_bits) : 0u;
    if constexpr (PACKED) {
#pragma unroll
      for (int i = 0; i < NW; ++i) accw[i] = bgw;
    } else {
#pragma unroll
      for (int j = 0; j < NPIX; ++j) acc[j] = bg;
    }

    // block j0 into registers: masks (lane L: words 4L .. 4L + 3 of the
    // block) and descriptors (lane L: bytes 16 (L % 4) .. of descriptor L / 4)
    const uint32_t jm = (uint32_t)(lane * 4) / SW, om = (uint32_t)(lane * 4) % SW, jd = (uint32_t)lane >> 2;
    uint4 m_lo = make_uint4(0u, 0u, 0u, 0u), m_hi = m_lo, d_q = m_lo;
    auto load_block = [&](uint32_t j0) {
      if (j0 + jm < cnt) {
        const uint4* src = reinterpret_cast<const uint4*>(sbase + (uint64_t)(j0 + jm) * a.slot_words + om);
        m_lo = src[0];
        m_hi = src[1];
      }
      if (jd < kBlk && j0 + jd < cnt)
        d_q = reinterpret_cast<const uint4*>(dbase + j0 + jd)[lane & 3];
    };
    auto stage_block = [&]() {
      wave_sync();  // the previous block's reads are done
      uint4* dm = reinterpret_cast<uint4*>(mk + lane * 4);
      dm[0] = m_lo;
      dm[1] = m_hi;
      if (jd < kBlk) reinterpret_cast<uint4*>(ds + jd)[lane & 3] = d_q;
      wave_sync();
    };
    load_block(0u);
    for (uint32_t j0 = 0; j0 < cnt; j0 += kBlk) {
      stage_block();
      if (j0 + kBlk < cnt) load_block(j0 + kBlk);  // in flight during the folds below
      const uint32_t gn = min(kBlk, cnt - j0);
      for (uint32_t g = 0; g < gn; ++g) {
        const ItemDesc& d = ds[g];
        if (d.r1 <= Rw || d.r0 >= strip_r1 || d.c1 <= C0 || d.c0 >= tile_c1) continue;
        const T v = from_bits<T>(d.value);
        const uint64_t vbits = d.value;
        const uint64_t* m = mk + g * SW;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const uint64_t row_word = m[(srow + k * RSTEP) * NWORD + (ccol >> 6)];
          const uint32_t bits = (uint32_t)(row_word >> (ccol & 63)) & ((1u << PPC) - 1u);
          if (__any(bits != 0u)) {
            burns += (unsigned long long)__popc(bits);
            if constexpr (PACKED) {
              const uint32_t vw = lane_splat<T>(vbits);
#pragma unroll
              for (int i = 0; i < PPC / PPW; ++i)
                accw[k * (PPC / PPW) + i] = fold_word<T, FUN>(
                    accw[k * (PPC / PPW) + i], vw, bgw,
                    lane_mask<(int)sizeof(T)>((bits >> (i * PPW)) & ((1u << PPW) - 1u)));
            } else {
#pragma unroll
              for (int j = 0; j < PPC; ++j)
                acc[k * PPC + j] = fold_select<T, FUN>(acc[k * PPC + j], v, bg, (bits >> j) & 1u);
            }
          }
        }
      }
    }

    // write the strip once, as k_tile does
    const int32_t c = C0 + ccol;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int32_t row = Rw + srow + k * RSTEP;
      if (row < nrows) {
        T* orow = (T*)J.out + ((uint64_t)layer * (uint64_t)nrows + (uint64_t)row) * (uint64_t)ncols;
        if (c + PPC <= ncols && ((uintptr_t)(orow + c) & 15u) == 0) {
          u32x4 vec;
          if constexpr (PACKED) vec = u32x4{accw[k * 4], accw[k * 4 + 1], accw[k * 4 + 2], accw[k * 4 + 3]};
          else __builtin_memcpy(&vec, &acc[k * PPC], 16);
          __builtin_nontemporal_store(vec, (u32x4*)(orow + c));
        } else {
#pragma unroll
          for (int j = 0; j < PPC; ++j)
            if (c + j < ncols) {
              if constexpr (PACKED) {
                T x;
                __builtin_memcpy(&x, reinterpret_cast<const char*>(accw) + (k * PPC + j) * sizeof(T), sizeof(T));
                orow[c + j] = x;
              } else {
                orow[c + j] = acc[k * PPC + j];
              }
            }
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
