// rz_runtime.hip -- device runtime of librusterize_hip: contexts, resident
// geometry, workspace and the launch sequence of one rasterise call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "rz_internal.h"
#include "rz_kernels.h"
#include "rz_synth_job.h"

// The set-up, binning and pre-pass kernels are compiled in this translation
// unit; the tile kernel instantiations come from rz_tile_inst.hip (one
// translation unit per dtype, declared in rz_launch.h).
#include "rz_launch.h"
#include "rz_kernels.hip"
#include "rz_sort.hip"
#include "rz_tile.hip"

using namespace rzk;

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return rz::fail(RZ_DEVICE_ERROR, std::string("HIP error '") + hipGetErrorString(_e) + \
                                           "' at " #expr);                                 \
  } while (0)

namespace {

constexpr int kStages = 5;
const char* kStageNames[kStages] = {"prep", "bin", "sort", "cover", "tile"};

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

// Grows a workspace buffer to at least `bytes`, in stream order on s (the
// old buffer is freed and the new one allocated after the work already
// queued on s, from the stream-ordered pool: no device-wide synchronisation,
// so other contexts' streams keep running).
hipError_t ensure(hipStream_t s, Buf& b, size_t bytes) {
  if (bytes <= b.cap && b.p) return hipSuccess;
  const size_t b_old_cap = b.cap;
  if (b.p) {
    hipError_t e = hipFreeAsync(b.p, s);
    if (e != hipSuccess) return e;
    b.p = nullptr;
    b.cap = 0;
  }
  // grow by at least half the old capacity: a workspace that keeps growing
  // (batches of varying size) is reallocated a logarithmic number of times
  size_t cap = std::max<size_t>(std::max<size_t>(bytes, 256), b_old_cap + b_old_cap / 2);
  hipError_t e = hipMallocAsync(&b.p, cap, s);
  if (e == hipSuccess) b.cap = cap;
  return e;
}

inline unsigned blocks_for(int64_t n, int bs = 256) {
  return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs);
}

}  // namespace

struct EventSet {
  hipEvent_t ev[kStages + 1] = {};
  bool tile_only = false;  // only the tile stage's two events were recorded
};

struct rz_context {
  int device = 0;
  hipStream_t stream = nullptr;
  // workspace
  Buf edges, lseg, item_box, item_last, desc, pair_desc, counts, offsets, keys_a, vals_a, keys_b,
      vals_b, tile_off, temp, misc, jobtab, queue, band_n, band_base, band_cnt, band_off, band_edges,
      heavy_cnt, heavy_base, masks, scan_tmp, scan_tmp2, scan_tmp3, scan_tmp4,
      sort_hist, sort_off, sort_tmp, cover_list, wide_units, wide_words, wide_ubase, wide_wbase,
      wide_item, wide_masks, scan_tmp5, scan_tmp6, band_items, heavy_list, scan_tmp7, wide_j, band_segs,
      item_win, at_words, at_base, at_bits, scan_tmp8, scan_box;
  // pinned: [0] pairs, [1] status, [2] band entries, [3] banded edges, [4] wide units, [5] wide words,
  // [6] all-touched PixelCache words
  uint64_t* host_misc = nullptr;
  // job table: built in tab_build, staged in pinned host_tab (tab_ev marks
  // the end of its last upload); tab_dev / tab_dev_bytes say which device
  // buffer holds the staged bytes
  std::vector<unsigned char> tab_build;
  void* host_tab = nullptr;
  size_t host_tab_cap = 0;
  hipEvent_t tab_ev = nullptr;
  bool tab_pending = false;
  const void* tab_dev = nullptr;
  size_t tab_dev_bytes = 0;
  // profiling: 0 off, 1 stage events, 2 stage events + exact burn counts,
  // 3 the tile stage's two events only (the least added to a timed call)
  int profiling = 0;
  hipEvent_t ev[kStages + 1] = {};  // the current call's events (taken from the pool)
  std::vector<EventSet> ev_pool, ev_pending;
  EventSet cur;
  double acc_ms[kStages] = {};
  uint64_t last_pairs = 0, last_tiles = 0, last_band_entries = 0, last_band_edges = 0;
  uint64_t last_wide_units = 0, last_wide_words = 0, last_at_words = 0;
  bool burns_pending = false;
  std::vector<Buf*> all() {
    return {&edges,  &lseg,   &item_box, &item_last, &desc, &counts,    &offsets, &keys_a,
            &vals_a, &keys_b, &vals_b, &tile_off, &temp,      &misc, &pair_desc, &jobtab,  &queue,
            &band_n, &band_base, &band_cnt, &band_off, &band_edges, &heavy_cnt, &heavy_base,
            &masks, &scan_tmp, &scan_tmp2, &scan_tmp3, &scan_tmp4,
            &sort_hist, &sort_off, &sort_tmp, &cover_list, &wide_units, &wide_words, &wide_ubase,
            &wide_wbase, &wide_item, &wide_masks, &scan_tmp5, &scan_tmp6, &band_items, &heavy_list, &scan_tmp7,
            &wide_j, &band_segs, &item_win, &at_words, &at_base, &at_bits, &scan_tmp8, &scan_box};
  }
  // Takes an event set for the next call (events are recorded on the stream
  // and resolved later, so timing adds no synchronisation to the call).
  hipError_t take_events() {
    if (ev_pool.empty()) {
      EventSet s;
      for (auto& e : s.ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
      }
      ev_pool.push_back(s);
    }
    cur = ev_pool.back();
    ev_pool.pop_back();
    cur.tile_only = profiling == 3;
    for (int i = 0; i <= kStages; ++i) ev[i] = cur.ev[i];
    return hipSuccess;
  }
  void accumulate_times() { ev_pending.push_back(cur); }
  void resolve_times() {
    for (auto& s : ev_pending) {
      if (hipEventSynchronize(s.ev[kStages]) == hipSuccess)
        for (int i = s.tile_only ? kStages - 1 : 0; i < kStages; ++i) {
          float t = 0.f;
          if (hipEventElapsedTime(&t, s.ev[i], s.ev[i + 1]) == hipSuccess) acc_ms[i] += (double)t;
        }
      ev_pool.push_back(s);
    }
    ev_pending.clear();
  }
};

struct rz_device_geometry {
  rz_context* ctx = nullptr;
  int64_t G = 0, I = 0, R = 0, V = 0;
  int64_t* geom_item_off = nullptr;
  int32_t* item_kind = nullptr;
  int64_t* item_ring_off = nullptr;
  int64_t* ring_voff = nullptr;
  double* xy = nullptr;
  int32_t* item_geom = nullptr;
  int32_t* ring_item = nullptr;
  uint64_t* vert_info = nullptr;
  int64_t* item_vbeg = nullptr;
  bool has_poly = false, has_line = false, has_point = false;
  int64_t max_poly_nv = 0;  // most vertices of one polygon item
  int64_t max_line_nv = 0;  // most vertices of one line item
};

// ------------------------------------------------ allocation-free workspace --
// Device memory of rz_rasterize_device_noalloc: a fixed array in the code
// object, so that entry point never allocates.  Every device the library
// runs on holds its own copy until the process ends.  In host code the name
// `scratch` is only a stand-in; its device address comes from
// hipGetSymbolAddress, taken once per device when the device's first
// context is created (rz_context_create) and kept in g_scratch.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;              // guards g_scratch
static void* g_scratch[kMaxDevices] = {};    // device address of `scratch` per device
static std::mutex g_noalloc_mu[kMaxDevices]; // one allocation-free call per device at a time

// Resolves `scratch` on `device` (the current device); once per device.
static rz_status scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return RZ_OK;  // that device has no allocation-free path
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return RZ_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return rz::fail(RZ_DEVICE_ERROR, std::string("HIP error '") + hipGetErrorString(e) +
                                         "' resolving the allocation-free workspace");
  g_scratch[device] = p;
  return RZ_OK;
}

// ------------------------------------------------------------- contexts --

static rz_status set_device(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return rz::fail(RZ_DEVICE_ERROR,
                    "no HIP device available (rusterize_amd needs an MI355X / gfx950 GPU)");
  if (device < 0 || device >= n) return rz::fail(RZ_DEVICE_ERROR, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return rz::fail(RZ_DEVICE_ERROR, std::string("device is ") + prop.gcnArchName +
                                         "; librusterize_hip is built for gfx950 only");
  return RZ_OK;
}

extern "C" rz_status rz_context_create(int device, rz_context** out) {
  rz_status st = set_device(device);
  if (st) return st;
  auto* c = new (std::nothrow) rz_context();
  if (!c) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  c->device = device;
  st = scratch_setup(device);
  if (st) {
    delete c;
    return st;
  }
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipHostMalloc((void**)&c->host_misc, 8 * sizeof(uint64_t), hipHostMallocDefault));
  HIP_TRY(hipEventCreateWithFlags(&c->tab_ev, hipEventDisableTiming));
  *out = c;
  return RZ_OK;
}

extern "C" rz_status rz_context_destroy(rz_context* c) {
  if (!c) return RZ_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (Buf* b : c->all())
    if (b->p) (void)hipFree(b->p);
  c->resolve_times();
  for (auto& s : c->ev_pool)
    for (auto& e : s.ev)
      if (e) (void)hipEventDestroy(e);
  if (c->host_misc) (void)hipHostFree(c->host_misc);
  if (c->host_tab) (void)hipHostFree(c->host_tab);
  if (c->tab_ev) (void)hipEventDestroy(c->tab_ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RZ_OK;
}

extern "C" size_t rz_context_workspace_bytes(const rz_context* c) {
  size_t s = 0;
  for (Buf* b : const_cast<rz_context*>(c)->all()) s += b->cap;
  return s;
}

// Reads the error flags the context's kernels raised (the caller has
// synchronised the stream) and clears them, so each failure is reported once.
static rz_status check_status_word(rz_context* c) {
  uint32_t* st = (uint32_t*)((char*)c->misc.p);
  uint32_t flags = 0;
  HIP_TRY(hipMemcpy(&flags, st, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (flags) {
    const uint32_t zero = 0;
    HIP_TRY(hipMemcpy(st, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  if (flags & 1u)
    return rz::fail(RZ_CAPACITY_ERROR,
                    "line coordinates beyond +-2^60 pixels are outside the supported range");
  if (flags & 16u) return rz::fail(RZ_VALUE_ERROR, "layer index out of range");
  if (flags & 32u)
    return rz::fail(RZ_CAPACITY_ERROR,
                    "an all-touched polygon's pixel cache would exceed 2^34 bits (its outline box is too large)");
  if (flags & 64u)
    return rz::fail(RZ_VALUE_ERROR, "an all-touched line walk did not end (non-finite coordinates)");
  if (flags & (4u | 8u))
    return rz::fail(RZ_CAPACITY_ERROR,
                    "the batch outgrew the capacities it was captured with; capture it again");
  return RZ_OK;
}

extern "C" rz_status rz_context_synchronize(rz_context* c) {
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->misc.p) return check_status_word(c);
  return RZ_OK;
}

extern "C" rz_status rz_context_set_profiling(rz_context* c, int level) {
  c->profiling = level < 0 ? 0 : (level > 3 ? 3 : level);
  return RZ_OK;
}

extern "C" const char* rz_stage_name(int i) { return (i >= 0 && i < kStages) ? kStageNames[i] : ""; }

extern "C" int rz_context_stage_times(rz_context* c, double* ms, int n) {
  (void)hipSetDevice(c->device);
  c->resolve_times();
  for (int i = 0; i < kStages && i < n; ++i) ms[i] = c->acc_ms[i];
  return kStages;
}

extern "C" void rz_context_reset_stage_times(rz_context* c) {
  c->resolve_times();
  for (double& m : c->acc_ms) m = 0.0;
}

extern "C" rz_status rz_context_last_counters(rz_context* c, uint64_t* burns, uint64_t* pairs,
                                              uint64_t* tiles) {
  HIP_TRY(hipSetDevice(c->device));
  if (burns) {
    *burns = 0;
    if (c->burns_pending && c->misc.p) {
      unsigned long long slots[kBurnSlots];
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipMemcpy(slots, (char*)c->misc.p + 64, sizeof(slots), hipMemcpyDeviceToHost));
      for (unsigned long long v : slots) *burns += v;
    }
  }
  if (pairs) *pairs = c->last_pairs;
  if (tiles) *tiles = c->last_tiles;
  return RZ_OK;
}

extern "C" rz_status rz_device_malloc(rz_context* c, size_t bytes, void** out) {
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMalloc(out, std::max<size_t>(bytes, 1)));
  return RZ_OK;
}
extern "C" rz_status rz_device_free(rz_context* c, void* p) {
  HIP_TRY(hipSetDevice(c->device));
  if (p) HIP_TRY(hipFree(p));
  return RZ_OK;
}
extern "C" rz_status rz_memcpy_h2d(rz_context* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipSetDevice(c->device));
  if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RZ_OK;
}
extern "C" rz_status rz_memcpy_d2d(rz_context* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipSetDevice(c->device));
  if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return RZ_OK;
}
extern "C" rz_status rz_memcpy_d2h(rz_context* c, void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipSetDevice(c->device));
  if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RZ_OK;
}

// ----------------------------------------------------- resident geometry --

template <typename T>
static hipError_t upload(rz_context* c, T** dst, const T* src, int64_t n) {
  hipError_t e = hipMalloc((void**)dst, std::max<size_t>(1, (size_t)n * sizeof(T)));
  if (e != hipSuccess) return e;
  if (n > 0) e = hipMemcpyAsync(*dst, src, (size_t)n * sizeof(T), hipMemcpyHostToDevice, c->stream);
  return e;
}

extern "C" rz_status rz_device_geometry_free(rz_device_geometry* dg) {
  if (!dg) return RZ_OK;
  (void)hipSetDevice(dg->ctx->device);
  (void)hipStreamSynchronize(dg->ctx->stream);
  void* ptrs[] = {dg->geom_item_off, dg->item_kind, dg->item_ring_off, dg->ring_voff, dg->xy,
                  dg->item_geom,     dg->ring_item, dg->vert_info,     dg->item_vbeg};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete dg;
  return RZ_OK;
}

extern "C" int64_t rz_device_geometry_count(const rz_device_geometry* dg) { return dg ? dg->G : 0; }

// New coordinates for a resident set of the same topology (its derived maps
// depend on the ring and item structure only), copied on the stream of the
// context that uploaded it and waited for; captured batches over the set
// read them on their next replay.
extern "C" rz_status rz_device_geometry_set_xy(rz_device_geometry* dg, const double* xy, int64_t n_vertices) {
  if (!dg || (!xy && n_vertices > 0)) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_device_geometry_set_xy");
  if (n_vertices != dg->V) return rz::fail(RZ_VALUE_ERROR, "vertex count differs from the resident geometry's");
  if (n_vertices == 0) return RZ_OK;
  HIP_TRY(hipSetDevice(dg->ctx->device));
  HIP_TRY(hipMemcpyAsync(dg->xy, xy, (size_t)n_vertices * 16, hipMemcpyHostToDevice, dg->ctx->stream));
  HIP_TRY(hipStreamSynchronize(dg->ctx->stream));
  return RZ_OK;
}

static rz_status finish_geometry(rz_context* c, rz_device_geometry* dg, rz_device_geometry** out);

extern "C" rz_status rz_device_geometry_upload(rz_context* c, const rz_geometry* g,
                                               rz_device_geometry** out) {
  rz_status st = rz::check_geometry(g);
  if (st) return st;
  if (g->n_items >= INT32_MAX || g->n_rings >= INT32_MAX || g->n_geoms >= INT32_MAX ||
      g->n_vertices >= ((int64_t)1 << 40))
    return rz::fail(RZ_CAPACITY_ERROR, "geometry set exceeds 2^31 items / rings / geometries");
  HIP_TRY(hipSetDevice(c->device));
  auto* dg = new (std::nothrow) rz_device_geometry();
  if (!dg) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  dg->ctx = c;
  dg->G = g->n_geoms;
  dg->I = g->n_items;
  dg->R = g->n_rings;
  dg->V = g->n_vertices;
  for (int64_t i = 0; i < g->n_items; ++i) {
    int k = g->item_kind[i];
    if (g->item_ring_offsets[i + 1] == g->item_ring_offsets[i]) continue;
    const int64_t nv = g->ring_vertex_offsets[g->item_ring_offsets[i + 1]] -
                       g->ring_vertex_offsets[g->item_ring_offsets[i]];
    if (k == RZ_ITEM_POLYGONS) dg->max_poly_nv = std::max<int64_t>(dg->max_poly_nv, nv);
    if (k == RZ_ITEM_LINES) dg->max_line_nv = std::max<int64_t>(dg->max_line_nv, nv);
    dg->has_poly |= k == RZ_ITEM_POLYGONS;
    dg->has_line |= k == RZ_ITEM_LINES;
    dg->has_point |= k == RZ_ITEM_POINTS;
  }
  std::vector<int64_t> zero1(1, 0);
  hipError_t e = upload(c, &dg->geom_item_off, g->geom_item_offsets, g->n_geoms + 1);
  if (e == hipSuccess) e = upload(c, &dg->item_kind, g->item_kind, g->n_items);
  if (e == hipSuccess)
    e = upload(c, &dg->item_ring_off, g->n_items ? g->item_ring_offsets : zero1.data(),
               g->n_items + 1);
  if (e == hipSuccess)
    e = upload(c, &dg->ring_voff, g->n_rings ? g->ring_vertex_offsets : zero1.data(),
               g->n_rings + 1);
  if (e == hipSuccess) e = upload(c, &dg->xy, g->xy, 2 * g->n_vertices);
  if (e == hipSuccess) e = hipMalloc((void**)&dg->item_geom, std::max<size_t>(1, 4 * g->n_items));
  if (e == hipSuccess) e = hipMalloc((void**)&dg->ring_item, std::max<size_t>(1, 4 * g->n_rings));
  if (e == hipSuccess)
    e = hipMalloc((void**)&dg->vert_info, std::max<size_t>(8, 8 * g->n_vertices));
  if (e == hipSuccess)
    e = hipMalloc((void**)&dg->item_vbeg, std::max<size_t>(8, 8 * (g->n_items + 1)));
  if (e != hipSuccess) {
    rz_device_geometry_free(dg);
    return rz::fail(RZ_DEVICE_ERROR, std::string("geometry upload failed: ") + hipGetErrorString(e));
  }
  return finish_geometry(c, dg, out);
}

// Derived maps of a resident geometry set whose offsets and coordinates are
// on the device (or queued on the context's stream): item -> geometry,
// ring -> item, item -> first vertex, vertex -> item / kind / ring flags.
// Frees dg on failure.
static rz_status finish_geometry(rz_context* c, rz_device_geometry* dg, rz_device_geometry** out) {
  hipError_t e = hipSuccess;
  if (dg->G)
    hipLaunchKernelGGL(k_item_geom, dim3(blocks_for(dg->G)), dim3(256), 0, c->stream,
                       dg->geom_item_off, dg->G, dg->item_geom);
  if (dg->I)
    hipLaunchKernelGGL(k_ring_item, dim3(blocks_for(dg->I)), dim3(256), 0, c->stream,
                       dg->item_ring_off, dg->ring_voff, dg->I, dg->ring_item, dg->item_vbeg);
  else
    (void)hipMemsetAsync(dg->item_vbeg, 0, 8, c->stream);
  if (dg->R)
    hipLaunchKernelGGL(k_vert_info, dim3(blocks_for(dg->R)), dim3(256), 0, c->stream, dg->ring_voff,
                       dg->ring_item, dg->item_kind, dg->xy, dg->R, dg->vert_info);
  e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    rz_device_geometry_free(dg);
    return rz::fail(RZ_DEVICE_ERROR, std::string("geometry upload failed: ") + hipGetErrorString(e));
  }
  *out = dg;
  return RZ_OK;
}

// One thread per vertex of job k of the batched-workload generator: the
// coordinates of rz_synth_job.h, bit-identical to the host generator.
__global__ void k_synth_job_xy(uint64_t seed, uint64_t k, int32_t side, const int64_t* ring_voff,
                               int64_t n_rings, int64_t V, double* xy) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  int64_t lo = 0, hi = n_rings - 1;  // last ring with ring_voff[r] <= v
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (ring_voff[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  const rzsynth::PolySpec ps = rzsynth::poly_spec(seed, k, (uint64_t)lo, side);
  double x, y;
  rzsynth::vertex(seed, k, (uint64_t)lo, ps, (int32_t)(v - ring_voff[lo]), &x, &y);
  xy[2 * v] = x;
  xy[2 * v + 1] = y;
}

extern "C" rz_status rz_device_synth_job(rz_context* c, uint64_t seed, int64_t k,
                                         rz_device_geometry** out) {
  if (!c || !out || k < 0) return rz::fail(RZ_VALUE_ERROR, "bad argument to rz_device_synth_job");
  HIP_TRY(hipSetDevice(c->device));
  const rzsynth::JobSpec js = rzsynth::job_spec(seed, (uint64_t)k);
  const int64_t G = js.n_poly;
  std::vector<int64_t> seq((size_t)G + 1), voff((size_t)G + 1, 0);
  std::vector<int32_t> kinds((size_t)G, RZ_ITEM_POLYGONS);
  for (int64_t p = 0; p <= G; ++p) seq[(size_t)p] = p;
  int64_t max_nv = 0;
  for (int64_t p = 0; p < G; ++p) {
    const int64_t nv = rzsynth::poly_spec(seed, (uint64_t)k, (uint64_t)p, js.side).n_vert + 1;
    voff[(size_t)p + 1] = voff[(size_t)p] + nv;
    max_nv = std::max(max_nv, nv);
  }
  auto* dg = new (std::nothrow) rz_device_geometry();
  if (!dg) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  dg->ctx = c;
  dg->G = dg->I = dg->R = G;
  dg->V = voff[(size_t)G];
  dg->has_poly = G > 0;
  dg->max_poly_nv = max_nv;
  hipError_t e = upload(c, &dg->geom_item_off, seq.data(), G + 1);
  if (e == hipSuccess) e = upload(c, &dg->item_kind, kinds.data(), G);
  if (e == hipSuccess) e = upload(c, &dg->item_ring_off, seq.data(), G + 1);
  if (e == hipSuccess) e = upload(c, &dg->ring_voff, voff.data(), G + 1);
  if (e == hipSuccess) e = hipMalloc((void**)&dg->xy, std::max<size_t>(16, 16 * (size_t)dg->V));
  if (e == hipSuccess) e = hipMalloc((void**)&dg->item_geom, std::max<size_t>(4, 4 * (size_t)G));
  if (e == hipSuccess) e = hipMalloc((void**)&dg->ring_item, std::max<size_t>(4, 4 * (size_t)G));
  if (e == hipSuccess) e = hipMalloc((void**)&dg->vert_info, std::max<size_t>(8, 8 * (size_t)dg->V));
  if (e == hipSuccess) e = hipMalloc((void**)&dg->item_vbeg, 8 * ((size_t)G + 1));
  if (e != hipSuccess) {
    rz_device_geometry_free(dg);
    return rz::fail(RZ_DEVICE_ERROR, std::string("job geometry allocation failed: ") + hipGetErrorString(e));
  }
  if (dg->V > 0)
    hipLaunchKernelGGL(k_synth_job_xy, dim3(blocks_for(dg->V)), dim3(256), 0, c->stream, seed,
                       (uint64_t)k, js.side, (const int64_t*)dg->ring_voff, G, dg->V, dg->xy);
  // the host vectors above must outlive the asynchronous uploads
  rz_status st = finish_geometry(c, dg, out);
  return st;
}

// ------------------------------------------------------------ rasterise --

namespace {

// ---- autotuned launch knobs
//
// Per size class (tile width, launch tiles, items per tile) the autotuner
// (rz_autotune) stores the tiles per queue grab and the item count from which
// a tile's coverage is precomputed by the k_cover pre-pass.  Classes without
// an entry use the defaults below.
struct Tuning {
  uint32_t chunk = 0;      // 0: default
  uint32_t heavy_min = 0;  // 0: default (kHeavyItems)
};

struct TuneState {
  std::mutex mu;
  std::unordered_map<uint32_t, Tuning> table;
  bool force = false;  // the autotuner's candidate under measurement
  Tuning forced;
};

TuneState& tune_state() {
  static TuneState s;
  return s;
}

inline int floor_log2_u64(uint64_t v) {
  int l = 0;
  while (l < 62 && ((uint64_t)2 << l) <= v) ++l;
  return l;
}

// Size class: bit 8 tile width (0: 64 px, 1: 256 px), bits 4-7 log4 of the
// launch's tiles, bits 0-3 log2 of (items per tile + 1).
uint32_t size_class(int tile_w, uint64_t tiles, uint64_t pairs) {
  const uint64_t t = std::max<uint64_t>(tiles, 1);
  const uint32_t tb = (uint32_t)std::min(15, floor_log2_u64(t) / 2);
  const uint32_t db = (uint32_t)std::min(15, floor_log2_u64(pairs / t + 1));
  return (tile_w == 256 ? 1u << 8 : 0u) | tb << 4 | db;
}

Tuning tuning_for(uint32_t cls) {
  TuneState& ts = tune_state();
  std::lock_guard<std::mutex> lk(ts.mu);
  if (ts.force) return ts.forced;
  auto it = ts.table.find(cls);
  return it == ts.table.end() ? Tuning{} : it->second;
}

void launch_heavy_dispatch(rz_dtype dt, int fun, const TileArgs& a, hipStream_t s) {
  switch (dt) {
    case RZ_U8: launch_heavy_u8(fun, a, s); break;
    case RZ_U16: launch_heavy_u16(fun, a, s); break;
    case RZ_U32: launch_heavy_u32(fun, a, s); break;
    case RZ_U64: launch_heavy_u64(fun, a, s); break;
    case RZ_I8: launch_heavy_i8(fun, a, s); break;
    case RZ_I16: launch_heavy_i16(fun, a, s); break;
    case RZ_I32: launch_heavy_i32(fun, a, s); break;
    case RZ_I64: launch_heavy_i64(fun, a, s); break;
    case RZ_F32: launch_heavy_f32(fun, a, s); break;
    default: launch_heavy_f64(fun, a, s); break;
  }
}

void launch_tile_dispatch(rz_dtype dt, int fun, const TileArgs& a, bool lines, hipStream_t s) {
  switch (dt) {
    case RZ_U8: launch_tile_u8(fun, a, lines, s); break;
    case RZ_U16: launch_tile_u16(fun, a, lines, s); break;
    case RZ_U32: launch_tile_u32(fun, a, lines, s); break;
    case RZ_U64: launch_tile_u64(fun, a, lines, s); break;
    case RZ_I8: launch_tile_i8(fun, a, lines, s); break;
    case RZ_I16: launch_tile_i16(fun, a, lines, s); break;
    case RZ_I32: launch_tile_i32(fun, a, lines, s); break;
    case RZ_I64: launch_tile_i64(fun, a, lines, s); break;
    case RZ_F32: launch_tile_f32(fun, a, lines, s); break;
    default: launch_tile_f64(fun, a, lines, s); break;
  }
}

void zero_words(void* p, uint64_t n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for((int64_t)n)), dim3(256), 0, s, (uint32_t*)p, n);
}

// A tile launch: jobs [j0, j1) of one (dtype, function, lines) group.
struct Group {
  rz_dtype dtype;
  int fun;
  bool lines;
  uint32_t t_lo, t_hi;
};

rz_status check_job_common(const rz_job& j, bool need_out) {
  if (!j.geometry || (need_out && !j.out_dev))
    return rz::fail(RZ_VALUE_ERROR, "NULL geometry or output of a job");
  if ((unsigned)j.dtype > 9) return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
  if ((unsigned)j.fun > 6) return rz::fail(RZ_VALUE_ERROR, "Unknown pixel function");
  if ((unsigned)j.sum_mode > 1) return rz::fail(RZ_VALUE_ERROR, "unknown sum mode");
  const uint32_t nl = j.layer_dev ? j.n_layers : 1u;
  if (nl == 0 || nl >= (1u << 29)) return rz::fail(RZ_VALUE_ERROR, "n_layers must be in [1, 2^29)");
  if (j.info.nrows >= (uint64_t)INT32_MAX - TILE_H || j.info.ncols >= (uint64_t)INT32_MAX - 256)
    return rz::fail(RZ_CAPACITY_ERROR, "raster dimensions exceed 2^31 rows or columns");
  return RZ_OK;
}

}  // namespace

// The pipeline of a batch.  Jobs are ordered by (dtype, function, lines) so
// every group's tiles form one contiguous range; vertex, item and tile
// indices are global over the batch.
// Fixed capacities of a captured batch (see rz_batch_graph_create).
struct BatchCaps {
  uint64_t pairs, band_entries, band_edges, wide_units, wide_words, at_words;
};

// The device record of job j, whose vertex / item / tile ranges start at
// v_base / i_base / t_base of its batch (J.scan is left 0); returns the
// job's tile count (tiles per band x bands).
static uint64_t job_record(const rz_job& j, JobDev& J, int64_t v_base, int64_t i_base, uint64_t t_base) {
  const rz_device_geometry* dg = j.geometry;
  const int esz = (int)rz_dtype_size(j.dtype);
  const int TW = tile_w(esz);
  const uint32_t nl = j.layer_dev ? j.n_layers : 1u;
  const int32_t nrows = (int32_t)j.info.nrows, ncols = (int32_t)j.info.ncols;
  const int32_t ntx = (ncols + TW - 1) / TW, nty = (nrows + TILE_H - 1) / TILE_H;
  const uint64_t tpl = (uint64_t)ntx * (uint64_t)nty, jt = tpl * nl;
  std::memset(&J, 0, sizeof(J));
  J.xy = dg->xy;
  J.vert_info = dg->vert_info;
  J.ring_voff = dg->ring_voff;
  J.ring_item = dg->ring_item;
  J.item_kind = dg->item_kind;
  J.item_geom = dg->item_geom;
  J.item_vbeg = dg->item_vbeg;
  J.v_base = v_base;
  J.i_base = i_base;
  J.nv = dg->V;
  J.ni = dg->I;
  J.xmin = j.info.xmin;
  J.ymax = j.info.ymax;
  J.xres = j.info.xres;
  J.yres = j.info.yres;
  J.nrows = nrows;
  J.ncols = ncols;
  J.ntx = ntx;
  J.nty = nty;
  J.esz = esz;
  J.tile_w = TW;
  J.tiles_per_layer = (uint32_t)std::min<uint64_t>(tpl, UINT32_MAX);
  J.tile_base = (uint32_t)std::min<uint64_t>(t_base, UINT32_MAX);
  J.n_tiles = (uint32_t)std::min<uint64_t>(jt, UINT32_MAX);
  J.fun = (int32_t)j.fun;
  J.dedup_lines = j.info.xres != j.info.yres;
  J.sum_unordered = j.fun == RZ_SUM && j.sum_mode == RZ_SUM_UNORDERED && (j.dtype == RZ_F32 || j.dtype == RZ_F64);
  J.all_touched = j.all_touched ? 1 : 0;
  J.field = j.field_dev;
  J.scalar_bits = j.scalar_bits;
  J.bg_bits = j.background_bits;
  J.layer_of_geom = j.layer_dev;
  J.n_layers = nl;
  J.out = j.out_dev;
  return jt;
}

// masks_only: the set-up and the wide pre-pass only, for every polygon item
// and with no cap on the masks (the sparse encoding reads them); no bins, no
// tile launches.
static rz_status run_batch(rz_context* c, const rz_job* jobs_in, int64_t n_in,
                           const BatchCaps* caps = nullptr, bool masks_only = false) {
  const bool capture = caps != nullptr;
  if (n_in <= 0) return RZ_OK;
  if (n_in >= (int64_t)1 << 24) return rz::fail(RZ_CAPACITY_ERROR, "more than 2^24 jobs in one batch");
  for (int64_t k = 0; k < n_in; ++k) {
    rz_status st = check_job_common(jobs_in[k], !masks_only);
    if (st) return st;
  }
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // RUSTERIZE_BINNING=always|never overrides the size heuristic (tests use it
  // to run both item-list forms on the same inputs).
  int binning = 0;
  if (const char* bm = std::getenv("RUSTERIZE_BINNING")) {
    if (std::strcmp(bm, "always") == 0) binning = 1;
    else if (std::strcmp(bm, "never") == 0) binning = -1;
  }
  // RUSTERIZE_COVERAGE=rows|edges forces one coverage layout in the wide
  // pre-pass (tests run both on the same inputs).
  int cov_mode = 0;
  if (const char* cm = std::getenv("RUSTERIZE_COVERAGE")) {
    if (std::strcmp(cm, "rows") == 0) cov_mode = 1;
    else if (std::strcmp(cm, "edges") == 0) cov_mode = 2;
  }
  // RUSTERIZE_WIDE=never|always: no item / every polygon item takes the wide
  // coverage pre-pass (tests run both on the same inputs).
  int wide_mode = 0;
  if (const char* wm = std::getenv("RUSTERIZE_WIDE")) {
    if (std::strcmp(wm, "never") == 0) wide_mode = 1;
    else if (std::strcmp(wm, "always") == 0) wide_mode = 2;
  }
  if (masks_only) wide_mode = 2;
  // (all-touched jobs walk their polygons' outlines too: the lines kernel)
  auto has_lines = [](const rz_job& j) { return j.geometry->has_line || j.geometry->has_point || j.all_touched; };
  std::vector<int64_t> order((size_t)n_in);
  for (int64_t k = 0; k < n_in; ++k) order[(size_t)k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    const rz_job &x = jobs_in[a], &y = jobs_in[b];
    if (x.dtype != y.dtype) return x.dtype < y.dtype;
    if (x.fun != y.fun) return x.fun < y.fun;
    return has_lines(x) < has_lines(y);
  });

  // job table: JobDev[n] then the vertex / item / tile prefix ends (int64 x n each)
  const size_t n = (size_t)n_in;
  const size_t tab_bytes = n * sizeof(JobDev) + 3 * n * sizeof(int64_t);
  c->tab_build.assign(tab_bytes, 0);
  JobDev* tab = (JobDev*)c->tab_build.data();
  int64_t* vend = (int64_t*)(tab + n);
  int64_t* iend = vend + n;
  int64_t* tend = iend + n;
  std::vector<Group> groups;
  int64_t V = 0, I = 0;
  uint64_t T = 0, out_px = 0;
  bool any_binned = false, any_poly = false, any_lp = false, any_band = false;
  bool any_poly_band = false, any_line_band = false;
  bool any_tw64 = false, any_tw256 = false;
  bool any_at_cache = false;  // all-touched sum / count polygons: PixelCache windows
  bool any_wide_job = false;  // jobs whose polygons may take the wide pre-pass
  bool any_scan_compact = false;
  int max_tile_w = 64;
  for (size_t k = 0; k < n; ++k) {
    const rz_job& j = jobs_in[order[k]];
    const rz_device_geometry* dg = j.geometry;
    JobDev& J = tab[k];
    const uint64_t jt = job_record(j, J, V, I, T);
    const int TW = J.tile_w;
    const uint32_t nl = J.n_layers;
    const int32_t nrows = J.nrows, ncols = J.ncols;
    const uint64_t ni = (uint64_t)dg->I;
    // scan: every strip tests every item box; otherwise the items are binned
    // into per-tile lists.  Binning a lone call costs a host round trip and
    // about eight small launches, so a single call scans up to ~4M
    // item-tile tests; in a batch that cost is shared by every job.
    const uint64_t scan_limit = n_in == 1 ? ((uint64_t)1 << 22) : ((uint64_t)1 << 16);
    bool scan = ni <= 64 || (ni <= 2048 && ni * std::max<uint64_t>(jt, 1) <= scan_limit);
    if (masks_only) scan = true;
    else if (binning == 1) scan = ni == 0;
    else if (binning == -1) scan = true;
    J.scan = scan ? 1 : 0;
    J.scan_compact = scan && nl == 1 && (nrows + WAVE_ROWS - 1) / WAVE_ROWS < 0xFFFF && J.ntx < 0xFFFF;
    any_scan_compact |= J.scan_compact != 0;
    any_binned |= !scan && ni > 0;
    const bool lines = has_lines(j);
    any_tw64 |= TW == 64;
    any_tw256 |= TW == 256;
    max_tile_w = std::max(max_tile_w, TW);
    any_poly |= dg->has_poly;
    // A small (scanned) job whose rings fit two 64-edge stages gains nothing
    // from the wide pre-pass: each of its strips loads a ring's edges once
    // either way, and the pass would add three launches to the call.
    any_wide_job |= dg->has_poly && !(scan && dg->max_poly_nv <= kBandMinEdges);
    any_poly_band |= dg->has_poly && dg->max_poly_nv > kBandMinEdges;
    any_line_band |= dg->has_line && dg->max_line_nv > kLineBandMinSegs;
    any_band = any_poly_band || any_line_band;
    any_lp |= lines;
    any_at_cache |= j.all_touched && dg->has_poly && (j.fun == RZ_SUM || j.fun == RZ_COUNT);
    if (groups.empty() || groups.back().dtype != j.dtype || groups.back().fun != (int)j.fun ||
        groups.back().lines != lines)
      groups.push_back(Group{j.dtype, (int)j.fun, lines, (uint32_t)T, (uint32_t)T});
    V += dg->V;
    I += dg->I;
    T += jt;
    out_px += (uint64_t)nrows * (uint64_t)ncols * nl;
    if (T >= ((uint64_t)1 << 29))
      return rz::fail(RZ_CAPACITY_ERROR, "more than 2^29 output tiles in one call or batch");
    groups.back().t_hi = (uint32_t)T;
    vend[k] = V;
    iend[k] = I;
    tend[k] = (int64_t)T;
  }
  if (I >= (int64_t)UINT32_MAX) return rz::fail(RZ_CAPACITY_ERROR, "more than 2^32 items in one batch");
  const uint32_t n_tiles = (uint32_t)T;

  // workspace
  HIP_TRY(ensure(c->stream, c->jobtab, tab_bytes));
  const size_t misc_bytes = 64 + kBurnSlots * sizeof(unsigned long long);
  HIP_TRY(ensure(c->stream, c->misc, misc_bytes));
  // per group: the tile queue, then one word: the k_tile_heavy work counter
  const size_t queue_words = groups.size() * (kQueueWords + 1);
  HIP_TRY(ensure(c->stream, c->queue, queue_words * 4));
  uint32_t* d_status = (uint32_t*)c->misc.p;
  unsigned long long* d_burns = (unsigned long long*)((char*)c->misc.p + 64);
  if (any_poly) HIP_TRY(ensure(c->stream, c->edges, (size_t)std::max<int64_t>(V, 1) * sizeof(PolyEdge)));
  if (any_lp) HIP_TRY(ensure(c->stream, c->lseg, (size_t)std::max<int64_t>(V, 1) * sizeof(uint4)));
  HIP_TRY(ensure(c->stream, c->item_box, (size_t)std::max<int64_t>(I, 1) * sizeof(ItemBox)));
  HIP_TRY(ensure(c->stream, c->item_last, (size_t)std::max<int64_t>(I, 1) * 8));
  HIP_TRY(ensure(c->stream, c->desc, (size_t)std::max<int64_t>(I, 1) * sizeof(ItemDesc)));
  if (any_band) {
    HIP_TRY(ensure(c->stream, c->band_n, (size_t)(I + 1) * 4));
    HIP_TRY(ensure(c->stream, c->band_base, (size_t)(I + 1) * 8));
  }
  if (any_binned) {
    HIP_TRY(ensure(c->stream, c->counts, (size_t)(I + 1) * 4));
    HIP_TRY(ensure(c->stream, c->offsets, (size_t)(I + 1) * 8));
  }
  if (any_scan_compact) HIP_TRY(ensure(c->stream, c->scan_box, (size_t)std::max<int64_t>(I, 1) * 8));
  if (any_at_cache) {
    HIP_TRY(ensure(c->stream, c->item_win, (size_t)std::max<int64_t>(I, 1) * 64));
    HIP_TRY(ensure(c->stream, c->at_words, (size_t)(I + 1) * 4));
    HIP_TRY(ensure(c->stream, c->at_base, (size_t)(I + 1) * 8));
  }
  const bool any_wide = any_poly && wide_mode != 1 && I > 0 && (wide_mode == 2 || any_wide_job);
  // A small eager call whose only pre-pass is the wide one sizes the wide
  // buffers for the most they can hold (2 bytes per output pixel, every
  // unit at least 16 words) instead of reading the totals back: no host
  // round trip.
  const bool wide_fixed =
      !capture && !masks_only && any_wide && !any_binned && !any_band && out_px <= ((uint64_t)1 << 26);
  if (any_wide) {
    HIP_TRY(ensure(c->stream, c->wide_units, (size_t)(I + 1) * 4));
    HIP_TRY(ensure(c->stream, c->wide_words, (size_t)(I + 1) * 4));
    HIP_TRY(ensure(c->stream, c->wide_ubase, (size_t)(I + 1) * 8));
    HIP_TRY(ensure(c->stream, c->wide_wbase, (size_t)(I + 1) * 8));
  }
  const JobDev* d_jobs = (const JobDev*)c->jobtab.p;
  const int64_t* d_vend = (const int64_t*)(d_jobs + n);
  const int64_t* d_iend = d_vend + n;
  const int64_t* d_tend = d_iend + n;

  // (stage events: every stage at levels 1-2, only the tile stage at 3)
  const bool stage_ev = c->profiling == 1 || c->profiling == 2;
  if (c->profiling) HIP_TRY(c->take_events());
  if (stage_ev) HIP_TRY(hipEventRecord(c->ev[0], s));
  // upload the job table unless the device already holds these exact bytes
  // (repeated calls with the same arguments skip the copy)
  if (!(c->tab_dev == c->jobtab.p && c->tab_dev_bytes == tab_bytes &&
        std::memcmp(c->host_tab, c->tab_build.data(), tab_bytes) == 0)) {
    if (c->tab_pending) HIP_TRY(hipEventSynchronize(c->tab_ev));  // staging still being read
    c->tab_pending = false;
    if (c->host_tab_cap < tab_bytes) {
      if (c->host_tab) HIP_TRY(hipHostFree(c->host_tab));
      c->host_tab = nullptr;
      c->host_tab_cap = 0;
      HIP_TRY(hipHostMalloc(&c->host_tab, tab_bytes, hipHostMallocDefault));
      c->host_tab_cap = tab_bytes;
    }
    std::memcpy(c->host_tab, c->tab_build.data(), tab_bytes);
    c->tab_dev = nullptr;
    HIP_TRY(hipMemcpyAsync(c->jobtab.p, c->host_tab, tab_bytes, hipMemcpyHostToDevice, s));
    if (!capture) {
      HIP_TRY(hipEventRecord(c->tab_ev, s));
      c->tab_pending = true;
    }
    c->tab_dev = c->jobtab.p;
    c->tab_dev_bytes = tab_bytes;
  }
  // 1. item boxes reset and the call's counters cleared (one launch), edges /
  // pixel coordinates / item boxes, then the item descriptors
  {
    InitArgs ia{};
    std::memset(&ia, 0, sizeof(ia));
    ia.box = (ItemBox*)c->item_box.p;
    ia.last = (unsigned long long*)c->item_last.p;
    ia.n_items = I;
    ia.win = any_at_cache ? (unsigned long long*)c->item_win.p : nullptr;
    ia.clear(c->misc.p, (c->profiling == 2 ? misc_bytes : 64) / 4);
    if (any_at_cache) ia.clear((uint32_t*)c->at_words.p + I, 1);
    ia.clear(c->queue.p, queue_words);
    if (any_binned) ia.clear((uint32_t*)c->counts.p + I, 1);
    if (any_band) ia.clear((uint32_t*)c->band_n.p + I, 1);
    if (any_wide) {
      ia.clear((uint32_t*)c->wide_units.p + I, 1);
      ia.clear((uint32_t*)c->wide_words.p + I, 1);
    }
    const uint64_t nz = std::max<uint64_t>((uint64_t)I, ia.words());
    hipLaunchKernelGGL(k_init, dim3(blocks_for((int64_t)nz)), dim3(256), 0, s, ia);
  }
  if (V > 0) {
    PrepArgs pa{};
    pa.jobs = d_jobs;
    pa.job_vend = d_vend;
    pa.n_jobs = (int)n;
    pa.n_vertices = V;
    pa.edges = (PolyEdge*)c->edges.p;
    pa.lseg = (uint4*)c->lseg.p;
    pa.item_box = (ItemBox*)c->item_box.p;
    pa.item_last = (unsigned long long*)c->item_last.p;
    pa.item_win = any_at_cache ? (unsigned long long*)c->item_win.p : nullptr;
    hipLaunchKernelGGL(k_prep, dim3(blocks_for(V)), dim3(256), 0, s, pa);
  }
  if (I > 0) {
    DescArgs da{};
    da.jobs = d_jobs;
    da.job_iend = d_iend;
    da.n_jobs = (int)n;
    da.n_items = I;
    da.item_box = (const ItemBox*)c->item_box.p;
    da.item_last = (const unsigned long long*)c->item_last.p;
    da.desc = (ItemDesc*)c->desc.p;
    da.counts = any_binned ? (uint32_t*)c->counts.p : nullptr;
    da.band_n = any_band ? (uint32_t*)c->band_n.p : nullptr;
    da.wide_mode = wide_mode;
    da.wide_units = any_wide ? (uint32_t*)c->wide_units.p : nullptr;
    da.wide_words = any_wide ? (uint32_t*)c->wide_words.p : nullptr;
    da.status = d_status;
    da.item_win = any_at_cache ? (const unsigned long long*)c->item_win.p : nullptr;
    da.at_words = any_at_cache ? (uint32_t*)c->at_words.p : nullptr;
    da.scan_box = any_scan_compact ? (uint2*)c->scan_box.p : nullptr;
    hipLaunchKernelGGL(k_item_desc, dim3(blocks_for(I)), dim3(256), 0, s, da);
  }
  TileArgs ta{};
  ta.scan_box = any_scan_compact ? (const uint2*)c->scan_box.p : nullptr;
  // all-touched sum / count polygons whose fill reaches past their
  // PixelCache window: their caches in global memory (an eager call reads
  // their size back; a captured batch uses its fixed capacity, and items
  // past it fail the replay)
  if (any_at_cache && I > 0) {
    HIP_TRY(ensure(c->stream, c->scan_tmp8, scan_blocks((uint64_t)I + 1) * 8));
    exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->at_words.p, (uint64_t*)c->at_base.p,
                                       (uint64_t)I + 1, (uint64_t*)c->scan_tmp8.p, s);
    uint64_t words = 0;
    if (capture) {
      words = caps->at_words;
    } else {
      HIP_TRY(hipMemcpyAsync(&c->host_misc[6], (const uint64_t*)c->at_base.p + I, 8, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      words = c->host_misc[6];
      c->last_at_words = words;
    }
    if (words > 0 || capture) {
      HIP_TRY(ensure(c->stream, c->at_bits, std::max<uint64_t>(words, 1) * 8));
      zero_words(c->at_bits.p, std::max<uint64_t>(words, 1) * 2, s);
      AtCacheArgs aa{};
      aa.cap_words = words;
      aa.jobs = d_jobs;
      aa.desc = (const ItemDesc*)c->desc.p;
      aa.n_items = I;
      aa.item_win = (const unsigned long long*)c->item_win.p;
      aa.at_base = (const uint64_t*)c->at_base.p;
      aa.at_bits = (unsigned long long*)c->at_bits.p;
      aa.status = d_status;
      hipLaunchKernelGGL(k_at_cache, dim3((unsigned)((I + 3) / 4)), dim3(256), 0, s, aa);
    }
    ta.item_win = (const unsigned long long*)c->item_win.p;
    ta.at_base = (const uint64_t*)c->at_base.p;
    ta.at_bits = (words > 0 || capture) ? (const unsigned long long*)c->at_bits.p : nullptr;
    ta.at_cap = words;
  }
  if (stage_ev) HIP_TRY(hipEventRecord(c->ev[1], s));

  // 2. binning: stable sort-middle item lists per tile for binned jobs.
  // Eager calls read the pair, band-entry and band-edge totals back (one or
  // two host round trips) and size the buffers to them; a captured batch
  // (caps != null) uses fixed capacities instead and never leaves the device.
  uint64_t n_pairs = 0;
  if (any_binned) {
    HIP_TRY(ensure(c->stream, c->scan_tmp, scan_blocks((uint64_t)I + 1) * 8));
    exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->counts.p, (uint64_t*)c->offsets.p,
                                       (uint64_t)I + 1, (uint64_t*)c->scan_tmp.p, s);
    if (!capture)
      HIP_TRY(hipMemcpyAsync(&c->host_misc[0], (const uint64_t*)c->offsets.p + I, 8,
                             hipMemcpyDeviceToHost, s));
  }
  if (any_band) {
    HIP_TRY(ensure(c->stream, c->scan_tmp2, scan_blocks((uint64_t)I + 1) * 8));
    exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->band_n.p, (uint64_t*)c->band_base.p,
                                       (uint64_t)I + 1, (uint64_t*)c->scan_tmp2.p, s);
    if (!capture)
      HIP_TRY(hipMemcpyAsync(&c->host_misc[2], (const uint64_t*)c->band_base.p + I, 8,
                             hipMemcpyDeviceToHost, s));
  }
  if (any_wide) {
    HIP_TRY(ensure(c->stream, c->scan_tmp5, scan_blocks((uint64_t)I + 1) * 8));
    HIP_TRY(ensure(c->stream, c->scan_tmp6, scan_blocks((uint64_t)I + 1) * 8));
    exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->wide_units.p, (uint64_t*)c->wide_ubase.p,
                                       (uint64_t)I + 1, (uint64_t*)c->scan_tmp5.p, s);
    exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->wide_words.p, (uint64_t*)c->wide_wbase.p,
                                       (uint64_t)I + 1, (uint64_t*)c->scan_tmp6.p, s);
    if (!capture && !wide_fixed) {
      HIP_TRY(hipMemcpyAsync(&c->host_misc[4], (const uint64_t*)c->wide_ubase.p + I, 8,
                             hipMemcpyDeviceToHost, s));
      HIP_TRY(hipMemcpyAsync(&c->host_misc[5], (const uint64_t*)c->wide_wbase.p + I, 8,
                             hipMemcpyDeviceToHost, s));
    }
  }
  if (!capture && (any_binned || any_band || (any_wide && !wide_fixed))) HIP_TRY(hipStreamSynchronize(s));
  // wide pre-pass: unit list, flags and mask bases (before the descriptors
  // are copied per pair).  The masks take at most 2 bytes per output pixel
  // (16 overlapping wide items per pixel at one bit); items past that stay
  // per tile.
  uint64_t wide_units = 0, wide_words = 0;
  c->last_wide_units = c->last_wide_words = 0;
  if (any_wide) {
    // (wide_fixed: the most the masks may hold, from the output size; the
    // actual totals stay on the device, where k_wide_assign and k_wide_cover
    // read them)
    const uint64_t words_cap = std::max<uint64_t>(out_px / 4, (uint64_t)1 << 20);
    wide_units = capture ? caps->wide_units : (wide_fixed ? words_cap / 16 : c->host_misc[4]);
    wide_words = capture ? caps->wide_words : (wide_fixed ? words_cap : c->host_misc[5]);
    c->last_wide_units = wide_units;
    c->last_wide_words = wide_words;
    if (!masks_only) wide_words = std::min<uint64_t>(wide_words, words_cap);
    wide_words = std::min<uint64_t>(wide_words, (uint64_t)1 << 36);  // ItemDesc::wide holds words / 16
    if (wide_units > 0 && wide_words > 0) {
      HIP_TRY(ensure(c->stream, c->wide_item, wide_units * 4));
      HIP_TRY(ensure(c->stream, c->wide_masks, wide_words * 8));
      WideAssignArgs wa{};
      wa.desc = (ItemDesc*)c->desc.p;
      wa.n_items = I;
      wa.wide_units = (const uint32_t*)c->wide_units.p;
      wa.unit_base = (const uint64_t*)c->wide_ubase.p;
      wa.word_base = (const uint64_t*)c->wide_wbase.p;
      wa.unit_cap = wide_units;
      wa.word_cap = wide_words;
      wa.unit_item = (uint32_t*)c->wide_item.p;
      hipLaunchKernelGGL(k_wide_assign, dim3(blocks_for(I)), dim3(256), 0, s, wa);
    } else {
      wide_units = 0;
    }
  }
  c->last_band_entries = c->last_band_edges = 0;
  if (any_band) {
    // edge bands of long polygon rings (see kBandMinEdges)
    const uint64_t nb = capture ? caps->band_entries : c->host_misc[2];
    c->last_band_entries = nb;
    if (nb > 0) {
      HIP_TRY(ensure(c->stream, c->band_cnt, (nb + 1) * 4));
      HIP_TRY(ensure(c->stream, c->band_off, (nb + 1) * 8));
      HIP_TRY(ensure(c->stream, c->scan_tmp3, scan_blocks(nb + 1) * 8));
      zero_words(c->band_cnt.p, nb + 1, s);
      // list of the banded items (band_items), its length after them
      HIP_TRY(ensure(c->stream, c->band_items, ((size_t)I + 1) * 4));
      uint32_t* n_list = (uint32_t*)c->band_items.p + I;
      zero_words(n_list, 1, s);
      hipLaunchKernelGGL(k_band_assign, dim3(blocks_for(I, kBandAssignBlock)), dim3(kBandAssignBlock), 0, s,
                         (ItemDesc*)c->desc.p,
                         (const uint32_t*)c->band_n.p, (const uint64_t*)c->band_base.p, I, nb,
                         (uint32_t*)c->band_items.p, n_list);
      BandArgs bd{};
      bd.desc = (const ItemDesc*)c->desc.p;
      bd.band_n = (const uint32_t*)c->band_n.p;
      bd.items = (const uint32_t*)c->band_items.p;
      bd.n_list = n_list;
      bd.edges = (const PolyEdge*)c->edges.p;
      bd.lseg = (const uint4*)c->lseg.p;
      bd.band_cnt = (uint32_t*)c->band_cnt.p;
      bd.band_off = (const uint64_t*)c->band_off.p;
      bd.band_edges = nullptr;
      bd.band_segs = nullptr;
      bd.edge_cap = 0;
      bd.status = d_status;
      const unsigned bgrid = (unsigned)std::min<uint64_t>((nb + 3) / 4, 16384);
      hipLaunchKernelGGL(k_band_count, dim3(bgrid), dim3(256), 0, s, bd);
      exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->band_cnt.p, (uint64_t*)c->band_off.p,
                                         nb + 1, (uint64_t*)c->scan_tmp3.p, s);
      uint64_t ne = 0;
      if (capture) {
        ne = caps->band_edges;
      } else {
        HIP_TRY(hipMemcpyAsync(&c->host_misc[3], (const uint64_t*)c->band_off.p + nb, 8,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        ne = c->host_misc[3];
      }
      c->last_band_edges = ne;
      // polygon and line entries share one index space; each kind's
      // records go to an array of its own record size
      if (any_poly_band) {
        HIP_TRY(ensure(c->stream, c->band_edges, std::max<uint64_t>(ne, 1) * sizeof(PolyEdge)));
        bd.band_edges = (PolyEdge*)c->band_edges.p;
      }
      if (any_line_band) {
        HIP_TRY(ensure(c->stream, c->band_segs, std::max<uint64_t>(ne, 1) * sizeof(uint4)));
        bd.band_segs = (uint4*)c->band_segs.p;
      }
      bd.edge_cap = ne;
      if (ne > 0) hipLaunchKernelGGL(k_band_scatter, dim3(bgrid), dim3(256), 0, s, bd);
      ta.band_off = (const uint64_t*)c->band_off.p;
      ta.band_edges = bd.band_edges;
      ta.band_segs = bd.band_segs;
    }
  }
  // the wide pre-pass (after the bands it reads)
  // (with heavy tiles: their counts, slot bases, pair lists and slots, so the
  // wide items' masks also land in the heavy slots)
  auto launch_wide = [&](const uint32_t* h_cnt, const uint64_t* h_base, const uint32_t* w_j,
                         uint64_t* masks, uint32_t slot_words) {
    if (wide_units > 0) {
      WideCoverArgs wc{};
      wc.jobs = d_jobs;
      wc.desc = (const ItemDesc*)c->desc.p;
      wc.n_items = I;
      wc.unit_item = (const uint32_t*)c->wide_item.p;
      wc.unit_base = (const uint64_t*)c->wide_ubase.p;
      wc.n_units = (const uint64_t*)c->wide_ubase.p + I;
      wc.unit_cap = wide_units;
      wc.edges = (const PolyEdge*)c->edges.p;
      wc.band_off = ta.band_off;
      wc.band_edges = ta.band_edges;
      wc.wide_masks = (uint64_t*)c->wide_masks.p;
      wc.cov_mode = cov_mode;
      wc.heavy_cnt = h_cnt;
      wc.heavy_base = h_base;
      wc.pair_off = (const uint64_t*)c->offsets.p;
      wc.pair_cap = n_pairs;
      wc.wide_j = w_j;
      wc.masks = masks;
      wc.slot_words = slot_words;
      const uint32_t wgrid = (uint32_t)std::min<uint64_t>((wide_units + 3) / 4, wide_fixed ? 512 : 4096);
      hipLaunchKernelGGL(k_wide_cover, dim3(wgrid), dim3(256), 0, s, wc);
      ta.wide_masks = (const uint64_t*)c->wide_masks.p;
    }
  };
  if (any_binned) {
    const uint64_t* d_total = (const uint64_t*)c->offsets.p + I;
    n_pairs = capture ? caps->pairs : c->host_misc[0];
    if (n_pairs >= (uint64_t)UINT32_MAX)
      return rz::fail(RZ_CAPACITY_ERROR, "more than 2^32 tile/item pairs in one call");
    const uint64_t cap = std::max<uint64_t>(n_pairs, 1);
    HIP_TRY(ensure(c->stream, c->keys_a, cap * 4));
    HIP_TRY(ensure(c->stream, c->vals_a, cap * 4));
    HIP_TRY(ensure(c->stream, c->keys_b, cap * 4));
    HIP_TRY(ensure(c->stream, c->vals_b, cap * 4));
    HIP_TRY(ensure(c->stream, c->tile_off, ((size_t)n_tiles + 1) * 4));
    HIP_TRY(ensure(c->stream, c->pair_desc, cap * sizeof(ItemDesc)));
    BinArgs ba{};
    ba.jobs = d_jobs;
    ba.desc = (const ItemDesc*)c->desc.p;
    ba.n_items = I;
    ba.offsets = (const uint64_t*)c->offsets.p;
    ba.cap = n_pairs;
    ba.keys = (uint32_t*)c->keys_a.p;
    ba.vals = (uint32_t*)c->vals_a.p;
    ba.status = d_status;
    hipLaunchKernelGGL(k_bin_emit, dim3(blocks_for(I)), dim3(256), 0, s, ba);
    // a captured batch sorts its whole capacity: the slots past the actual
    // pairs get a key above every tile, so they sort last and are never read
    if (capture && n_pairs > 0)
      hipLaunchKernelGGL(k_fill_keys, dim3(blocks_for((int64_t)n_pairs)), dim3(256), 0, s,
                         (uint32_t*)c->keys_a.p, d_total, n_pairs);
    if (stage_ev) HIP_TRY(hipEventRecord(c->ev[2], s));
    // stable LSD radix sort by tile key keeps geometry order inside each tile
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= (uint64_t)n_tiles) ++end_bit;
    const uint32_t* sorted_keys = (const uint32_t*)c->keys_a.p;
    const uint32_t* sorted_vals = (const uint32_t*)c->vals_a.p;
    if (n_pairs > 0) {
      const uint64_t nbk = sort_blocks(n_pairs);
      HIP_TRY(ensure(c->stream, c->sort_hist, nbk * kRadix * 4));
      HIP_TRY(ensure(c->stream, c->sort_off, nbk * kRadix * 4));
      HIP_TRY(ensure(c->stream, c->sort_tmp, scan_blocks(nbk * kRadix) * 4));
      const bool in_b = radix_sort_pairs((uint32_t*)c->keys_a.p, (uint32_t*)c->vals_a.p,
                                         (uint32_t*)c->keys_b.p, (uint32_t*)c->vals_b.p, n_pairs, end_bit,
                                         (uint32_t*)c->sort_hist.p, (uint32_t*)c->sort_off.p,
                                         (uint32_t*)c->sort_tmp.p, s);
      if (in_b) {
        sorted_keys = (const uint32_t*)c->keys_b.p;
        sorted_vals = (const uint32_t*)c->vals_b.p;
      }
    }
    hipLaunchKernelGGL(k_tile_offsets, dim3(blocks_for((int64_t)n_tiles + 1)), dim3(256), 0, s,
                       sorted_keys, d_total, n_pairs, (const uint32_t*)d_status, n_tiles,
                       (uint32_t*)c->tile_off.p);
    if (n_pairs > 0)
      hipLaunchKernelGGL(k_gather_desc, dim3(blocks_for((int64_t)n_pairs)), dim3(256), 0, s,
                         (const ItemDesc*)c->desc.p, I, sorted_vals, d_total, n_pairs,
                         (ItemDesc*)c->pair_desc.p);
    ta.tile_off = (const uint32_t*)c->tile_off.p;
    ta.pair_desc = (const ItemDesc*)c->pair_desc.p;
    if (stage_ev) HIP_TRY(hipEventRecord(c->ev[3], s));

    // heavy tiles (kHeavyItems items or more): coverage masks precomputed in
    // parallel, one wave per (item, strip); the buffers are sized for every
    // pair so no host round trip is needed
    // the batch's class: that of its widest-tile / most-tiles group
    const Tuning bt = tuning_for(size_class(max_tile_w, n_tiles, n_pairs));
    const uint32_t heavy_min = bt.heavy_min ? bt.heavy_min : kHeavyItems;
    if (n_pairs >= heavy_min) {
      HIP_TRY(ensure(c->stream, c->heavy_cnt, ((size_t)n_tiles + 1) * 4));
      HIP_TRY(ensure(c->stream, c->heavy_base, ((size_t)n_tiles + 1) * 8));
      HIP_TRY(ensure(c->stream, c->scan_tmp4, scan_blocks((uint64_t)n_tiles + 1) * 8));
      // heavy flags, their ranks and the ordered list of heavy tiles
      HIP_TRY(ensure(c->stream, c->heavy_list, ((size_t)n_tiles + 1) * 4 * 3));
      uint32_t* h_flag = (uint32_t*)c->heavy_list.p;
      uint32_t* h_rank = h_flag + (n_tiles + 1);
      uint32_t* h_list = h_rank + (n_tiles + 1);
      HIP_TRY(ensure(c->stream, c->scan_tmp7, scan_blocks((uint64_t)n_tiles + 1) * 4));
      hipLaunchKernelGGL(k_heavy_count, dim3(blocks_for((int64_t)n_tiles + 1)), dim3(256), 0, s,
                         (const uint32_t*)c->tile_off.p, n_tiles, heavy_min,
                         (uint32_t*)c->heavy_cnt.p, h_flag);
      exclusive_scan<uint32_t, uint32_t>(h_flag, h_rank, (uint64_t)n_tiles + 1, (uint32_t*)c->scan_tmp7.p, s);
      hipLaunchKernelGGL(k_heavy_list, dim3(blocks_for((int64_t)n_tiles)), dim3(256), 0, s,
                         (const uint32_t*)h_flag, (const uint32_t*)h_rank, n_tiles, h_list);
      ta.heavy_list = h_list;
      ta.heavy_rank = h_rank;
      exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->heavy_cnt.p, (uint64_t*)c->heavy_base.p,
                                         (uint64_t)n_tiles + 1, (uint64_t*)c->scan_tmp4.p, s);
      const uint32_t slot_words = 16u * (uint32_t)(max_tile_w / 64);
      HIP_TRY(ensure(c->stream, c->cover_list, (size_t)n_pairs * 4 + 4));
      HIP_TRY(ensure(c->stream, c->wide_j, (size_t)n_pairs * 4));
      HIP_TRY(ensure(c->stream, c->masks, (size_t)n_pairs * 4 * slot_words * 8));
      uint32_t* cover_count = (uint32_t*)c->cover_list.p + n_pairs;
      zero_words(cover_count, 1, s);
      hipLaunchKernelGGL(k_heavy_pairs, dim3(blocks_for((int64_t)n_pairs, kHeavyPairsBlock)),
                         dim3(kHeavyPairsBlock), 0, s, d_jobs,
                         sorted_keys, sorted_vals, (const ItemDesc*)c->pair_desc.p, d_total, n_pairs,
                         (const uint32_t*)c->tile_off.p, (const uint64_t*)c->heavy_base.p,
                         (const uint64_t*)c->offsets.p, (uint32_t*)c->wide_j.p, (uint32_t*)c->cover_list.p,
                         cover_count, n_tiles);
      CoverArgs ca{};
      ca.jobs = d_jobs;
      ca.job_tend = d_tend;
      ca.n_jobs = (int)n;
      ca.pair_tile = sorted_keys;
      ca.pair_desc = (const ItemDesc*)c->pair_desc.p;
      ca.heavy_base = (const uint64_t*)c->heavy_base.p;
      ca.cover_list = (const uint32_t*)c->cover_list.p;
      ca.cover_count = cover_count;
      ca.n_tiles = n_tiles;
      ca.edges = (const PolyEdge*)c->edges.p;
      ca.band_off = ta.band_off;
      ca.band_edges = ta.band_edges;
      ca.tile_off = (const uint32_t*)c->tile_off.p;
      ca.masks = (uint64_t*)c->masks.p;
      ca.slot_words = slot_words;
      launch_wide((const uint32_t*)c->heavy_cnt.p, (const uint64_t*)c->heavy_base.p,
                  (const uint32_t*)c->wide_j.p, (uint64_t*)c->masks.p, slot_words);
      const uint32_t cgrid = (uint32_t)std::min<uint64_t>(n_pairs, 2048);
      if (any_tw64) hipLaunchKernelGGL(k_cover<1>, dim3(cgrid), dim3(256), 0, s, ca);
      if (any_tw256) hipLaunchKernelGGL(k_cover<4>, dim3(cgrid), dim3(256), 0, s, ca);
      ta.heavy_base = (const uint64_t*)c->heavy_base.p;
      ta.masks = (const uint64_t*)c->masks.p;
      ta.slot_words = slot_words;
      ta.heavy_min = heavy_min;
    } else {
      launch_wide(nullptr, nullptr, nullptr, nullptr, 0);
    }
  } else {
    if (stage_ev) {
      HIP_TRY(hipEventRecord(c->ev[2], s));
      HIP_TRY(hipEventRecord(c->ev[3], s));
    }
    launch_wide(nullptr, nullptr, nullptr, nullptr, 0);
  }
  if (masks_only) return RZ_OK;
  if (c->profiling) HIP_TRY(hipEventRecord(c->ev[4], s));

  // 3. tiles: one persistent work-queue launch per group (queues cleared by k_init)
  ta.jobs = d_jobs;
  ta.job_tend = d_tend;
  ta.n_jobs = (int)n;
  ta.desc = (const ItemDesc*)c->desc.p;
  ta.edges = (const PolyEdge*)c->edges.p;
  ta.lseg = (const uint4*)c->lseg.p;
  ta.burns = c->profiling == 2 ? d_burns : nullptr;
  ta.status = d_status;
  for (size_t g = 0; g < groups.size(); ++g) {
    if (groups[g].t_hi == groups[g].t_lo) continue;
    ta.t_lo = groups[g].t_lo;
    ta.t_hi = groups[g].t_hi;
    ta.queue = (uint32_t*)c->queue.p + g * kQueueWords;
    ta.heavy_ctr = (uint32_t*)c->queue.p + groups.size() * kQueueWords + g;
    // polygon-only groups leave their heavy tiles to k_tile_heavy
    ta.heavy_split = (ta.heavy_base && !groups[g].lines) ? 1 : 0;
    const uint64_t gt = groups[g].t_hi - groups[g].t_lo;
    ta.chunk = tuning_for(size_class(tile_w((int)rz_dtype_size(groups[g].dtype)), gt,
                                     n_tiles ? n_pairs * gt / n_tiles : 0)).chunk;
    launch_tile_dispatch(groups[g].dtype, groups[g].fun, ta, groups[g].lines, s);
    if (ta.heavy_split) launch_heavy_dispatch(groups[g].dtype, groups[g].fun, ta, s);
  }
  HIP_TRY(hipGetLastError());
  if (c->profiling) HIP_TRY(hipEventRecord(c->ev[5], s));
  if (c->profiling) c->accumulate_times();
  c->burns_pending = c->profiling == 2;
  c->last_pairs = n_pairs;
  c->last_tiles = n_tiles;
  return RZ_OK;
}

extern "C" rz_status rz_batch_rasterize(rz_context* c, const rz_job* jobs, int64_t n_jobs) {
  if (!c || (!jobs && n_jobs > 0)) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_batch_rasterize");
  return run_batch(c, jobs, n_jobs);
}

// ------------------------------------------------------- captured batches --

struct rz_batch_graph {
  rz_context* ctx = nullptr;  // the graph's own workspace and stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" rz_status rz_batch_graph_destroy(rz_batch_graph* g) {
  if (!g) return RZ_OK;
  if (g->ctx) {
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->ctx) rz_context_destroy(g->ctx);
  delete g;
  return RZ_OK;
}

extern "C" rz_status rz_batch_graph_create(int device, const rz_job* jobs, int64_t n_jobs,
                                           double headroom, rz_batch_graph** out) {
  if (!out || (!jobs && n_jobs > 0)) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_batch_graph_create");
  if (!(headroom >= 1.0) || headroom > 16.0) return rz::fail(RZ_VALUE_ERROR, "headroom must be in [1, 16]");
  auto* g = new (std::nothrow) rz_batch_graph();
  if (!g) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  rz_status st = rz_context_create(device, &g->ctx);
  if (st) {
    delete g;
    return st;
  }
  rz_context* c = g->ctx;
  // 1. one eager pass: its totals size the capacities
  st = run_batch(c, jobs, n_jobs);
  if (st == RZ_OK) st = rz_context_synchronize(c);
  if (st) {
    rz_batch_graph_destroy(g);
    return st;
  }
  auto grow = [&](uint64_t v) { return (uint64_t)((double)v * headroom) + (v ? 64 : 0); };
  const BatchCaps caps{grow(c->last_pairs),      grow(c->last_band_entries), grow(c->last_band_edges),
                       grow(c->last_wide_units), grow(c->last_wide_words),   grow(c->last_at_words)};
  // 2. a pass at those capacities allocates every buffer, so the captured
  // pass allocates nothing; 3. capture
  st = run_batch(c, jobs, n_jobs, &caps);
  if (st == RZ_OK) st = rz_context_synchronize(c);
  if (st) {
    rz_batch_graph_destroy(g);
    return st;
  }
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    st = run_batch(c, jobs, n_jobs, &caps);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g->graph);
    if (st == RZ_OK && e2 != hipSuccess) e = e2;
  }
  if (st == RZ_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (st || e != hipSuccess) {
    rz_batch_graph_destroy(g);
    return st ? st : rz::fail(RZ_DEVICE_ERROR, std::string("graph capture failed: ") + hipGetErrorString(e));
  }
  *out = g;
  return RZ_OK;
}

extern "C" rz_status rz_batch_graph_launch(rz_batch_graph* g) {
  if (!g || !g->exec) return rz::fail(RZ_VALUE_ERROR, "NULL graph");
  HIP_TRY(hipSetDevice(g->ctx->device));
  HIP_TRY(hipGraphLaunch(g->exec, g->ctx->stream));
  return RZ_OK;
}

extern "C" rz_status rz_batch_graph_synchronize(rz_batch_graph* g) {
  if (!g) return rz::fail(RZ_VALUE_ERROR, "NULL graph");
  return rz_context_synchronize(g->ctx);
}

// -------------------------------------------------------------- autotuner --

extern "C" rz_status rz_autotune(rz_context* c, const rz_job* jobs, int64_t n_jobs, int reps,
                                 int64_t* n_classes) {
  if (!c || (!jobs && n_jobs > 0) || reps < 1) return rz::fail(RZ_VALUE_ERROR, "bad argument to rz_autotune");
  HIP_TRY(hipSetDevice(c->device));
  static const uint32_t kChunks[] = {1, 2, 4, 8};
  static const uint32_t kHeavy[] = {16, 32, 64, 1u << 30};
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  TuneState& ts = tune_state();
  int64_t tuned = 0;
  rz_status st = RZ_OK;
  for (int64_t k = 0; k < n_jobs && st == RZ_OK; ++k) {
    Tuning best;
    float best_ms = 1e30f;
    uint32_t cls = 0;
    for (uint32_t ch : kChunks)
      for (uint32_t hv : kHeavy) {
        {
          std::lock_guard<std::mutex> lk(ts.mu);
          ts.force = true;
          ts.forced = Tuning{ch, hv};
        }
        st = run_batch(c, jobs + k, 1);  // warm-up (workspace)
        if (st == RZ_OK) st = rz_context_synchronize(c);
        float ms = 0.f;
        if (st == RZ_OK && hipEventRecord(e0, c->stream) != hipSuccess) st = RZ_DEVICE_ERROR;
        for (int r = 0; r < reps && st == RZ_OK; ++r) st = run_batch(c, jobs + k, 1);
        if (st == RZ_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
          st = rz::fail(RZ_DEVICE_ERROR, "autotune timing failed");
        if (st) break;
        if (ms < best_ms) {
          best_ms = ms;
          best = Tuning{ch, hv};
        }
        cls = size_class(tile_w((int)rz_dtype_size(jobs[k].dtype)), c->last_tiles, c->last_pairs);
      }
    if (st == RZ_OK) {
      std::lock_guard<std::mutex> lk(ts.mu);
      if (!ts.table.count(cls)) ++tuned;
      ts.table[cls] = best;
    }
  }
  {
    std::lock_guard<std::mutex> lk(ts.mu);
    ts.force = false;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (n_classes) *n_classes = tuned;
  return st;
}

extern "C" int64_t rz_autotune_table(uint32_t* cls, uint32_t* chunk, uint32_t* heavy_min, int64_t cap) {
  TuneState& ts = tune_state();
  std::lock_guard<std::mutex> lk(ts.mu);
  int64_t i = 0;
  for (const auto& kv : ts.table) {
    if (i < cap) {
      if (cls) cls[i] = kv.first;
      if (chunk) chunk[i] = kv.second.chunk;
      if (heavy_min) heavy_min[i] = kv.second.heavy_min;
    }
    ++i;
  }
  return i;
}

extern "C" rz_status rz_autotune_set(uint32_t cls, uint32_t chunk, uint32_t heavy_min) {
  if (chunk > 4096) return rz::fail(RZ_VALUE_ERROR, "chunk must be <= 4096");
  TuneState& ts = tune_state();
  std::lock_guard<std::mutex> lk(ts.mu);
  if (chunk == 0 && heavy_min == 0) ts.table.erase(cls);
  else ts.table[cls] = Tuning{chunk, heavy_min};
  return RZ_OK;
}

extern "C" rz_status rz_rasterize_device(rz_context* c, const rz_device_geometry* dg,
                                         const rz_raster_info* info, rz_dtype dtype, rz_fun fun,
                                         const void* field_dev, const void* scalar_host,
                                         const uint32_t* layer_dev, uint32_t n_layers,
                                         const void* bg_host, int all_touched, rz_sum_mode sum_mode,
                                         void* out_dev) {
  if (!c || !dg || !info || !bg_host || !out_dev)
    return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_rasterize_device");
  if ((unsigned)dtype > 9) return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
  if ((unsigned)fun > 6) return rz::fail(RZ_VALUE_ERROR, "Unknown pixel function");
  if (!field_dev && !scalar_host) return rz::fail(RZ_VALUE_ERROR, "no field and no scalar value");
  if ((unsigned)sum_mode > 1) return rz::fail(RZ_VALUE_ERROR, "unknown sum mode");
  if (info->nrows == 0 || info->ncols == 0) return RZ_OK;  // empty raster: nothing to write
  const size_t esz = rz_dtype_size(dtype);
  rz_job j;
  std::memset(&j, 0, sizeof(j));
  j.geometry = dg;
  j.info = *info;
  j.dtype = dtype;
  j.fun = fun;
  j.field_dev = field_dev;
  if (!field_dev) std::memcpy(&j.scalar_bits, scalar_host, esz);
  j.layer_dev = layer_dev;
  j.n_layers = layer_dev ? n_layers : 1u;
  std::memcpy(&j.background_bits, bg_host, esz);
  j.out_dev = out_dev;
  j.sum_mode = sum_mode;
  j.all_touched = all_touched ? 1 : 0;
  return run_batch(c, &j, 1);
}

// ------------------------------------------- allocation-free entry point --

// Writes the job record and its prefix ends (a batch of one) into the
// workspace; the record travels as a kernel argument, so no host copy and no
// staging buffer are involved.
__global__ void k_put_job(JobDev J, int64_t nv, int64_t ni, int64_t nt, JobDev* tab) {
  if (threadIdx.x == 0) {
    tab[0] = J;
    int64_t* ends = (int64_t*)(tab + 1);
    ends[0] = nv;
    ends[1] = ni;
    ends[2] = nt;
  }
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" rz_status rz_rasterize_device_noalloc(rz_context* c, const rz_device_geometry* dg,
                                                 const rz_raster_info* info, rz_dtype dtype, rz_fun fun,
                                                 const void* field_dev, const void* scalar_host,
                                                 const uint32_t* layer_dev, uint32_t n_layers,
                                                 const void* bg_host, int all_touched,
                                                 rz_sum_mode sum_mode, void* out_dev) {
  if (!c || !dg || !info || !bg_host || !out_dev)
    return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_rasterize_device_noalloc");
  if ((unsigned)dtype > 9) return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
  if ((unsigned)fun > 6) return rz::fail(RZ_VALUE_ERROR, "Unknown pixel function");
  if (!field_dev && !scalar_host) return rz::fail(RZ_VALUE_ERROR, "no field and no scalar value");
  if (all_touched)
    return rz::fail(RZ_VALUE_ERROR, "all_touched=True is not available on the allocation-free path");
  if ((unsigned)sum_mode > 1) return rz::fail(RZ_VALUE_ERROR, "unknown sum mode");
  if (info->nrows == 0 || info->ncols == 0) return RZ_OK;  // empty raster: nothing to write
  if (c->device < 0 || c->device >= kMaxDevices || !g_scratch[c->device])
    return rz::fail(RZ_DEVICE_ERROR, "no allocation-free workspace on this device");
  const size_t esz = rz_dtype_size(dtype);
  rz_job j;
  std::memset(&j, 0, sizeof(j));
  j.geometry = dg;
  j.info = *info;
  j.dtype = dtype;
  j.fun = fun;
  j.field_dev = field_dev;
  if (!field_dev) std::memcpy(&j.scalar_bits, scalar_host, esz);
  j.layer_dev = layer_dev;
  j.n_layers = layer_dev ? n_layers : 1u;
  std::memcpy(&j.background_bits, bg_host, esz);
  j.out_dev = out_dev;
  j.sum_mode = sum_mode;
  JobDev J;
  const uint64_t jt = job_record(j, J, 0, 0, 0);
  J.scan = 1;  // no bins: every strip tests the item boxes itself
  if (jt >= ((uint64_t)1 << 29))
    return rz::fail(RZ_CAPACITY_ERROR, "more than 2^29 output tiles in one call or batch");
  const int64_t V = dg->V, I = dg->I;
  const bool lines = dg->has_line || dg->has_point;
  // workspace layout in `scratch`: job record + ends, status / burn words,
  // tile queue, item boxes, item stats, descriptors, edges, line coordinates
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const size_t o_tab = take(sizeof(JobDev) + 3 * sizeof(int64_t));
  const size_t o_misc = take(64 + kBurnSlots * sizeof(unsigned long long));
  const size_t o_queue = take((kQueueWords + 1) * 4);
  const size_t o_box = take((size_t)std::max<int64_t>(I, 1) * sizeof(ItemBox));
  const size_t o_last = take((size_t)std::max<int64_t>(I, 1) * 8);
  const size_t o_desc = take((size_t)std::max<int64_t>(I, 1) * sizeof(ItemDesc));
  const size_t o_edges = dg->has_poly ? take((size_t)std::max<int64_t>(V, 1) * sizeof(PolyEdge)) : 0;
  const size_t o_lseg = lines ? take((size_t)std::max<int64_t>(V, 1) * sizeof(uint4)) : 0;
  if (off > kScratchBytes)
    return rz::fail(RZ_CAPACITY_ERROR,
                    "the geometry needs " + std::to_string(off) +
                        " bytes of workspace; the allocation-free entry point holds at most " +
                        std::to_string(kScratchBytes) + " (use rz_rasterize_device)");
  std::lock_guard<std::mutex> lk(g_noalloc_mu[c->device]);
  HIP_TRY(hipSetDevice(c->device));
  char* base = (char*)g_scratch[c->device];
  hipStream_t s = c->stream;
  JobDev* d_tab = (JobDev*)(base + o_tab);
  const int64_t* d_ends = (const int64_t*)(d_tab + 1);
  uint32_t* d_status = (uint32_t*)(base + o_misc);
  uint32_t* d_queue = (uint32_t*)(base + o_queue);
  ItemBox* d_box = (ItemBox*)(base + o_box);
  unsigned long long* d_last = (unsigned long long*)(base + o_last);
  ItemDesc* d_desc = (ItemDesc*)(base + o_desc);
  hipLaunchKernelGGL(k_put_job, dim3(1), dim3(64), 0, s, J, V, I, (int64_t)jt, d_tab);
  {
    // item boxes, the status word and the tile queue (the previous call's
    // contents stay in the array otherwise)
    InitArgs ia{};
    std::memset(&ia, 0, sizeof(ia));
    ia.box = d_box;
    ia.last = d_last;
    ia.n_items = I;
    ia.clear(d_status, 16);
    ia.clear(d_queue, kQueueWords + 1);
    const uint64_t nz = std::max<uint64_t>((uint64_t)I, ia.words());
    hipLaunchKernelGGL(k_init, dim3(blocks_for((int64_t)nz)), dim3(256), 0, s, ia);
  }
  if (V > 0) {
    PrepArgs pa{};
    pa.jobs = d_tab;
    pa.job_vend = d_ends;
    pa.n_jobs = 1;
    pa.n_vertices = V;
    pa.edges = dg->has_poly ? (PolyEdge*)(base + o_edges) : nullptr;
    pa.lseg = lines ? (uint4*)(base + o_lseg) : nullptr;
    pa.item_box = d_box;
    pa.item_last = d_last;
    hipLaunchKernelGGL(k_prep, dim3(blocks_for(V)), dim3(256), 0, s, pa);
  }
  if (I > 0) {
    DescArgs da{};
    std::memset(&da, 0, sizeof(da));
    da.jobs = d_tab;
    da.job_iend = d_ends + 1;
    da.n_jobs = 1;
    da.n_items = I;
    da.item_box = d_box;
    da.item_last = d_last;
    da.desc = d_desc;
    da.wide_mode = 1;
    da.status = d_status;
    hipLaunchKernelGGL(k_item_desc, dim3(blocks_for(I)), dim3(256), 0, s, da);
  }
  TileArgs ta{};
  std::memset(&ta, 0, sizeof(ta));
  ta.jobs = d_tab;
  ta.job_tend = d_ends + 2;
  ta.n_jobs = 1;
  ta.desc = d_desc;
  ta.edges = dg->has_poly ? (const PolyEdge*)(base + o_edges) : nullptr;
  ta.lseg = lines ? (const uint4*)(base + o_lseg) : nullptr;
  ta.status = d_status;
  ta.t_lo = 0;
  ta.t_hi = (uint32_t)jt;
  ta.queue = d_queue;
  ta.heavy_ctr = d_queue + kQueueWords;
  launch_tile_dispatch(dtype, (int)fun, ta, lines, s);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  uint32_t flags = 0;
  HIP_TRY(hipMemcpy(&flags, d_status, sizeof(flags), hipMemcpyDeviceToHost));
  if (flags & 1u)
    return rz::fail(RZ_CAPACITY_ERROR, "line coordinates beyond +-2^60 pixels are outside the supported range");
  if (flags & 16u) return rz::fail(RZ_VALUE_ERROR, "layer index out of range");
  return RZ_OK;
}

// --------------------------------------------------------------- serving --

#include "rz_server.hip"

// ------------------------------------------------- host-in / host-out ----

namespace {
std::mutex g_default_mu;
rz_context* g_default_ctx = nullptr;

rz_status default_context(rz_context** out) {
  if (!g_default_ctx) {
    int dev = 0;
    if (const char* e = std::getenv("RUSTERIZE_DEVICE")) dev = std::atoi(e);
    rz_status st = rz_context_create(dev, &g_default_ctx);
    if (st) {
      g_default_ctx = nullptr;
      return st;
    }
  }
  *out = g_default_ctx;
  return RZ_OK;
}
}  // namespace

extern "C" rz_status rz_rasterize(const rz_geometry* geoms, const rz_raster_info* info,
                                  rz_dtype dtype, rz_fun fun, const void* field,
                                  const void* scalar, const uint32_t* layer_of_geom,
                                  uint32_t n_layers, const void* background, int all_touched,
                                  rz_sum_mode sum_mode, void* out) {
  if (!geoms || !info || !background || !out)
    return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_rasterize");
  if ((unsigned)dtype > 9) return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
  if (!layer_of_geom) n_layers = 1;
  for (int64_t i = 0; layer_of_geom && i < geoms->n_geoms; ++i)
    if (layer_of_geom[i] >= n_layers) return rz::fail(RZ_VALUE_ERROR, "layer index out of range");
  std::lock_guard<std::mutex> lock(g_default_mu);
  rz_context* c = nullptr;
  rz_status st = default_context(&c);
  if (st) return st;
  HIP_TRY(hipSetDevice(c->device));
  const size_t esz = rz_dtype_size(dtype);
  const size_t out_bytes = (size_t)n_layers * info->nrows * info->ncols * esz;
  rz_device_geometry* dg = nullptr;
  st = rz_device_geometry_upload(c, geoms, &dg);
  if (st) return st;
  void *d_field = nullptr, *d_layer = nullptr, *d_out = nullptr;
  hipError_t e = hipSuccess;
  if (field) e = upload(c, (uint8_t**)&d_field, (const uint8_t*)field, (int64_t)(geoms->n_geoms * esz));
  if (e == hipSuccess && layer_of_geom) e = upload(c, (uint32_t**)&d_layer, layer_of_geom, geoms->n_geoms);
  if (e == hipSuccess) e = hipMalloc(&d_out, std::max<size_t>(out_bytes, 1));
  if (e == hipSuccess) {
    st = rz_rasterize_device(c, dg, info, dtype, fun, d_field, field ? nullptr : scalar,
                             (const uint32_t*)d_layer, n_layers, background, all_touched, sum_mode,
                             d_out);
    if (st == RZ_OK) {
      e = hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e == hipSuccess) st = check_status_word(c);
    }
  }
  if (d_field) (void)hipFree(d_field);
  if (d_layer) (void)hipFree(d_layer);
  if (d_out) (void)hipFree(d_out);
  rz_device_geometry_free(dg);
  if (e != hipSuccess) return rz::fail(RZ_DEVICE_ERROR, std::string("HIP error: ") + hipGetErrorString(e));
  return st;
}

// ------------------------------------------------------- sparse encoding --

#include "rz_sparse.hip"

// ------------------------------------------------------------ test hooks --

// Host evaluation of the Bresenham closed form used by the tile kernel
// (walk_window in rz_kernels.h), for brute-force checks on the CPU.  In the
// 32-bit domain the index-stepped form (walk_window32_idx) must give the
// same cells: -2 when it does not.
extern "C" int64_t rz_debug_walk_cells(int64_t ix0, int64_t iy0, int64_t ix1, int64_t iy1,
                                       int64_t X0, int64_t X1, int64_t Y0, int64_t Y1,
                                       int64_t* xs, int64_t* ys, int64_t cap) {
  int64_t n = 0;
  auto emit = [&](int64_t x, int64_t y) {
    if (n < cap) {
      xs[n] = x;
      ys[n] = y;
    }
    ++n;
  };
  const int64_t lim = (int64_t)1 << 30, wl = (int64_t)1 << 31;
  if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim &&
      iy1 > -lim && iy1 < lim && X0 > -wl && X1 < wl && Y0 > -wl && Y1 < wl && X0 <= X1 && Y0 <= Y1) {
    walk_window32((int32_t)ix0, (int32_t)iy0, (int32_t)ix1, (int32_t)iy1, (int32_t)X0, (int32_t)X1,
                  (int32_t)Y0, (int32_t)Y1, emit);
    const int64_t W = X1 - X0;
    if (W > 0 && W * (Y1 - Y0) < ((int64_t)1 << 32)) {
      int64_t k = 0;
      bool same = true;
      walk_window32_idx((int32_t)ix0, (int32_t)iy0, (int32_t)ix1, (int32_t)iy1, (int32_t)X0, (int32_t)X1,
                        (int32_t)Y0, (int32_t)Y1, (int32_t)W, [&](uint32_t idx) {
                          if (k < n && k < cap && (xs[k] != X0 + (int64_t)(idx % (uint64_t)W) ||
                                                   ys[k] != Y0 + (int64_t)(idx / (uint64_t)W)))
                            same = false;
                          ++k;
                        });
      if (!same || k != n) return -2;
    }
  } else if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim &&
           iy1 > -lim && iy1 < lim)
    walk_window<int64_t>(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1, emit);
  else
    walk_window<__int128>(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1, emit);
  return n;
}

// Host evaluation of the all-touched walker of the tile kernel
// (walk_all_touched in rz_kernels.h), for checks against the oracle.
extern "C" int64_t rz_debug_walk_all_touched(double x0, double y0, double x1, double y1, int64_t ncols,
                                             int64_t nrows, int64_t X0, int64_t X1, int64_t Y0, int64_t Y1,
                                             int64_t* xs, int64_t* ys, int64_t cap) {
  int64_t n = 0;
  auto emit = [&](int64_t x, int64_t y) {
    if (n < cap) {
      xs[n] = x;
      ys[n] = y;
    }
    ++n;
  };
  return walk_all_touched(x0, y0, x1, y1, ncols, nrows, X0, X1, Y0, Y1, emit) ? n : -1;
}

// Host evaluation of the polygon row coverage used by the tile kernel: the
// covered columns of [C0, C0+width) for crossings xs[0..n) of one row, as a
// 0/1 byte per column.  width must be 64, 128 or 256.
extern "C" int rz_debug_row_coverage(const double* xs, int64_t n, int64_t ncols, int32_t C0,
                                     int32_t width, uint8_t* out) {
  if (width != 64 && width != 128 && width != 256) return -1;
  uint64_t tog[4] = {0, 0, 0, 0};
  uint32_t lp = 0;
  int32_t amax = -1;
  for (int64_t i = 0; i < n; ++i)
    if (!is_pos_nan(xs[i])) toggle_crossing(span_col(xs[i], (double)ncols), C0, width, tog, lp, amax);
  for (int w = 0; w < width / 64; ++w) {
    uint64_t cov = width == 64    ? row_coverage_word<1>(tog, lp, amax, C0, w)
                   : width == 128 ? row_coverage_word<2>(tog, lp, amax, C0, w)
                                  : row_coverage_word<4>(tog, lp, amax, C0, w);
    for (int j = 0; j < 64; ++j) out[w * 64 + j] = (uint8_t)((cov >> j) & 1u);
  }
  return 0;
}
