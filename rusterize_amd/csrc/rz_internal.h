// rz_internal.h -- shared host-side declarations of librusterize_hip.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rusterize_hip.h"

struct rz_geometry_buffers {
  std::vector<int64_t> geom_item_offsets{0};
  std::vector<int32_t> item_kind;
  std::vector<int64_t> item_ring_offsets{0};
  std::vector<int64_t> ring_vertex_offsets{0};
  std::vector<double> xy;
  std::vector<uint8_t> ring_flags;

  // Builder helpers used by the parsers and generators.
  void begin_item(int32_t kind) {
    item_kind.push_back(kind);
    item_ring_offsets.push_back(item_ring_offsets.back());
  }
  // Appends a ring of n vertices to the current item.
  void add_ring(const double* pts, int64_t n, bool hole = false) {
    xy.insert(xy.end(), pts, pts + 2 * n);
    ring_flags.push_back(hole ? 1 : 0);
    ring_vertex_offsets.push_back(ring_vertex_offsets.back() + n);
    item_ring_offsets.back() += 1;
  }
  void end_geometry() { geom_item_offsets.push_back((int64_t)item_kind.size()); }
  int64_t n_geoms() const { return (int64_t)geom_item_offsets.size() - 1; }
};

namespace rz {

// Sets the thread-local error message and returns `st`.
rz_status fail(rz_status st, const std::string& msg);
rz_status ok();

// Validates the offsets of a geometry view (monotone, in range).
rz_status check_geometry(const rz_geometry* g);

}  // namespace rz
