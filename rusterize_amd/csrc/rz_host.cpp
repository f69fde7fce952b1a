// rz_host.cpp -- host-side parts of librusterize_hip: errors, enum parsing,
// the grid builder, geometry ingestion (WKT / WKB) and seeded generators.
//
// Built with g++ -O2 -ffp-contract=off so the grid arithmetic rounds exactly
// as the reference's f64 code (Rust never contracts a*b+c into an FMA).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>

#include "rz_internal.h"
#include "rz_synth_job.h"

// ------------------------------------------------------------------ errors --

static thread_local std::string g_last_error;

namespace rz {
rz_status fail(rz_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}
rz_status ok() { return RZ_OK; }

rz_status check_geometry(const rz_geometry* g) {
  if (!g) return fail(RZ_VALUE_ERROR, "geometry is NULL");
  if (g->n_geoms < 0 || g->n_items < 0 || g->n_rings < 0 || g->n_vertices < 0)
    return fail(RZ_VALUE_ERROR, "negative geometry sizes");
  if (!g->geom_item_offsets || (g->n_items && (!g->item_kind || !g->item_ring_offsets)) ||
      (g->n_rings && !g->ring_vertex_offsets) || (g->n_vertices && !g->xy))
    return fail(RZ_VALUE_ERROR, "geometry arrays missing");
  if (g->geom_item_offsets[0] != 0 || g->geom_item_offsets[g->n_geoms] != g->n_items)
    return fail(RZ_VALUE_ERROR, "geom_item_offsets do not span the items");
  // every ring belongs to an item (the per-vertex maps index items by ring)
  if (g->n_items == 0 && (g->n_rings != 0 || g->n_vertices != 0))
    return fail(RZ_VALUE_ERROR, "rings or vertices without items");
  for (int64_t i = 0; i < g->n_geoms; ++i)
    if (g->geom_item_offsets[i + 1] < g->geom_item_offsets[i])
      return fail(RZ_VALUE_ERROR, "geom_item_offsets not monotone");
  if (g->n_items) {
    if (g->item_ring_offsets[0] != 0 || g->item_ring_offsets[g->n_items] != g->n_rings)
      return fail(RZ_VALUE_ERROR, "item_ring_offsets do not span the rings");
    for (int64_t i = 0; i < g->n_items; ++i) {
      if (g->item_ring_offsets[i + 1] < g->item_ring_offsets[i])
        return fail(RZ_VALUE_ERROR, "item_ring_offsets not monotone");
      if (g->item_kind[i] < 0 || g->item_kind[i] > 2)
        return fail(RZ_VALUE_ERROR, "unknown item kind");
    }
  }
  if (g->n_rings) {
    if (g->ring_vertex_offsets[0] != 0 || g->ring_vertex_offsets[g->n_rings] != g->n_vertices)
      return fail(RZ_VALUE_ERROR, "ring_vertex_offsets do not span the vertices");
    for (int64_t r = 0; r < g->n_rings; ++r)
      if (g->ring_vertex_offsets[r + 1] < g->ring_vertex_offsets[r])
        return fail(RZ_VALUE_ERROR, "ring_vertex_offsets not monotone");
  } else if (g->n_vertices) {
    return fail(RZ_VALUE_ERROR, "vertices without rings");
  }
  return RZ_OK;
}
}  // namespace rz

extern "C" const char* rz_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* rz_version(void) { return "rusterize_amd 0.1.0 (gfx950)"; }

// ------------------------------------------------------------------- enums --

extern "C" rz_status rz_pixel_function_parse(const char* name, rz_fun* out) {
  // rust/src/rasterization/pixel_functions.rs:18-33
  static const char* names[7] = {"sum", "first", "last", "min", "max", "count", "any"};
  if (name)
    for (int i = 0; i < 7; ++i)
      if (std::strcmp(name, names[i]) == 0) {
        *out = (rz_fun)i;
        return RZ_OK;
      }
  return rz::fail(RZ_VALUE_ERROR, "Unknown pixel function");
}

extern "C" rz_status rz_dtype_parse(const char* name, rz_dtype* out) {
  // python/src/rusterize.rs:171-182
  static const char* names[10] = {"uint8", "uint16", "uint32", "uint64", "int8",
                                  "int16", "int32",  "int64",  "float32", "float64"};
  if (name)
    for (int i = 0; i < 10; ++i)
      if (std::strcmp(name, names[i]) == 0) {
        *out = (rz_dtype)i;
        return RZ_OK;
      }
  return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
}

extern "C" size_t rz_dtype_size(rz_dtype d) {
  static const size_t s[10] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8};
  return (unsigned)d < 10 ? s[d] : 0;
}

// ------------------------------------------------------------ grid builder --

// `v as usize` for f64 in Rust: NaN -> 0, saturating.
static uint64_t as_usize(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)v;
}

static bool is_pos_zero_total(double v) {
  // f64::total_cmp(&0.0) == Equal  <=>  v has the bit pattern of +0.0
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b == 0;
}

extern "C" rz_status rz_raster_info_build(const rz_geometry* g, int has_shape, uint64_t nrows,
                                          uint64_t ncols, int has_extent, const double* ext,
                                          int has_res, double xres, double yres, int tap,
                                          int32_t epsg, rz_raster_info* out) {
  // RasterInfoBuilder::build / build_with / finalize, rust/src/geo/raster.rs:50-156.
  if (!out) return rz::fail(RZ_VALUE_ERROR, "output is NULL");
  double xmin, ymin, xmax, ymax;
  bool inferred;
  if (has_extent) {
    if (is_pos_zero_total(ext[0]) && is_pos_zero_total(ext[1]) && is_pos_zero_total(ext[2]) &&
        is_pos_zero_total(ext[3]))
      return rz::fail(RZ_VALUE_ERROR, "Unspecified extent (all zeros).");
    xmin = ext[0];
    ymin = ext[1];
    xmax = ext[2];
    ymax = ext[3];
    inferred = false;
  } else {
    // geo::BoundingRect folded over every geometry == min/max over all
    // vertices; f64::min / f64::max ignore a NaN operand (std::fmin/fmax).
    bool have = false;
    double bx0 = 0, by0 = 0, bx1 = 0, by1 = 0;
    if (g)
      for (int64_t r = 0; r < g->n_rings; ++r) {
        if (g->ring_flags && (g->ring_flags[r] & 1)) continue;  // holes are not in the bbox
        for (int64_t v = g->ring_vertex_offsets[r]; v < g->ring_vertex_offsets[r + 1]; ++v) {
          double x = g->xy[2 * v], y = g->xy[2 * v + 1];
          if (!have) {
            bx0 = bx1 = x;
            by0 = by1 = y;
            have = true;
          } else {
            bx0 = std::fmin(bx0, x);
            by0 = std::fmin(by0, y);
            bx1 = std::fmax(bx1, x);
            by1 = std::fmax(by1, y);
          }
        }
      }
    if (!have) return rz::fail(RZ_RUNTIME_ERROR, "Cannot infer bounding box from geometry.");
    xmin = bx0;
    ymin = by0;
    xmax = bx1;
    ymax = by1;
    inferred = true;
  }
  if (!has_shape && !has_res)
    return rz::fail(RZ_VALUE_ERROR, "Must set at least one of `shape` or `resolution`");
  if (has_shape && has_res)
    return rz::fail(RZ_VALUE_ERROR, "Shape and resolution are mutually exclusive; provide only one");
  if (!has_shape) nrows = ncols = 0;
  if (!has_res) xres = yres = 0.0;
  if (has_shape && (nrows == 0 || ncols == 0))
    return rz::fail(RZ_VALUE_ERROR, "Shape values must be > 0.");
  if (has_res && (xres <= 0.0 || yres <= 0.0))
    return rz::fail(RZ_VALUE_ERROR, "Resolution values must be > 0.");
  if (inferred && !tap && has_res) {  // GDAL half-pixel buffer
    xmin -= xres / 2.0;
    xmax += xres / 2.0;
    ymin -= yres / 2.0;
    ymax += yres / 2.0;
  }
  if (!has_res) {
    xres = (xmax - xmin) / (double)ncols;
    yres = (ymax - ymin) / (double)nrows;
  } else if (tap) {
    xmin = std::floor(xmin / xres) * xres;
    xmax = std::ceil(xmax / xres) * xres;
    ymin = std::floor(ymin / yres) * yres;
    ymax = std::ceil(ymax / yres) * yres;
  }
  if (!has_shape) {
    nrows = as_usize(0.5 + (ymax - ymin) / yres);
    ncols = as_usize(0.5 + (xmax - xmin) / xres);
  }
  out->ncols = ncols;
  out->nrows = nrows;
  out->xmin = xmin;
  out->xmax = xmax;
  out->ymin = ymin;
  out->ymax = ymax;
  out->xres = xres;
  out->yres = yres;
  out->epsg = epsg;
  return RZ_OK;
}

// -------------------------------------------------------- geometry buffers --

extern "C" void rz_geometry_view(const rz_geometry_buffers* b, rz_geometry* v) {
  v->n_geoms = b->n_geoms();
  v->n_items = (int64_t)b->item_kind.size();
  v->n_rings = (int64_t)b->ring_vertex_offsets.size() - 1;
  v->n_vertices = (int64_t)b->xy.size() / 2;
  v->geom_item_offsets = b->geom_item_offsets.data();
  v->item_kind = b->item_kind.data();
  v->item_ring_offsets = b->item_ring_offsets.data();
  v->ring_vertex_offsets = b->ring_vertex_offsets.data();
  v->xy = b->xy.data();
  v->ring_flags = b->ring_flags.data();
}

extern "C" void rz_geometry_free(rz_geometry_buffers* b) { delete b; }

extern "C" rz_status rz_geometry_from_arrays(const rz_geometry* src, rz_geometry_buffers** out) {
  rz_status st = rz::check_geometry(src);
  if (st) return st;
  auto* b = new (std::nothrow) rz_geometry_buffers();
  if (!b) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  b->geom_item_offsets.assign(src->geom_item_offsets, src->geom_item_offsets + src->n_geoms + 1);
  b->item_kind.assign(src->item_kind, src->item_kind + src->n_items);
  b->item_ring_offsets.assign(src->item_ring_offsets, src->item_ring_offsets + src->n_items + 1);
  if (src->n_items == 0) b->item_ring_offsets.assign(1, 0);
  b->ring_vertex_offsets.assign(src->ring_vertex_offsets,
                                src->ring_vertex_offsets + src->n_rings + 1);
  if (src->n_rings == 0) b->ring_vertex_offsets.assign(1, 0);
  b->xy.assign(src->xy, src->xy + 2 * src->n_vertices);
  if (src->ring_flags) b->ring_flags.assign(src->ring_flags, src->ring_flags + src->n_rings);
  else b->ring_flags.assign((size_t)src->n_rings, 0);
  *out = b;
  return RZ_OK;
}

// geo-types LineString::close (used by Polygon::new): append the first
// coordinate unless the ring is empty or already closed.
static void close_ring(std::vector<double>& pts) {
  size_t n = pts.size() / 2;
  if (n == 0) return;
  if (!(pts[0] == pts[2 * n - 2] && pts[1] == pts[2 * n - 1])) {
    double x = pts[0], y = pts[1];
    pts.push_back(x);
    pts.push_back(y);
  }
}

// ---------------------------------------------------------------- WKT ------

namespace {

// Collections and multi geometries nest at most this deep in the parsers
// (deeper input is rejected instead of exhausting the stack).
constexpr int kMaxNesting = 64;

struct WktParser {
  const char* s;
  const char* e;
  rz_geometry_buffers* out;
  std::string err;
  int depth = 0;

  void ws() {
    while (s < e && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) ++s;
  }
  bool peek(char c) {
    ws();
    return s < e && *s == c;
  }
  bool expect(char c) {
    ws();
    if (s < e && *s == c) {
      ++s;
      return true;
    }
    err = std::string("expected '") + c + "'";
    return false;
  }
  std::string word() {
    ws();
    const char* b = s;
    while (s < e && ((*s >= 'A' && *s <= 'Z') || (*s >= 'a' && *s <= 'z'))) ++s;
    std::string w(b, s);
    for (auto& ch : w) ch = (char)std::toupper((unsigned char)ch);
    return w;
  }
  bool number(double* v) {
    ws();
    const char* b = s;
    while (s < e && ((*s >= '0' && *s <= '9') || *s == '.' || *s == '-' || *s == '+' ||
                     *s == 'e' || *s == 'E'))
      ++s;
    if (b == s) {
      err = "expected a number";
      return false;
    }
    std::string tok(b, s);
    char* endp = nullptr;
    errno = 0;
    double d = std::strtod(tok.c_str(), &endp);  // correctly rounded (glibc)
    if (endp != tok.c_str() + tok.size()) {
      err = "invalid number '" + tok + "'";
      return false;
    }
    *v = d;
    return true;
  }
  // Optional dimension tag after a keyword: Z, M, ZM. Returns coords/vertex.
  int dims(const std::string& kw, const std::string& base) {
    int nd = 2;
    std::string suffix = kw.substr(base.size());
    if (suffix == "Z") nd = 3;
    else if (suffix == "M") nd = 3;
    else if (suffix == "ZM") nd = 4;
    if (suffix.empty()) {
      const char* save = s;
      std::string w = word();
      if (w == "Z" || w == "M") nd = 3;
      else if (w == "ZM") nd = 4;
      else s = save;
    }
    return nd;
  }
  bool is_empty() {
    const char* save = s;
    std::string w = word();
    if (w == "EMPTY") return true;
    s = save;
    return false;
  }
  // Parses "x y [z [m]]"; keeps x, y.
  bool coord(int nd, std::vector<double>& pts) {
    double v[4];
    for (int i = 0; i < nd; ++i) {
      if (!number(&v[i])) return false;
    }
    // A WKT without a dimension tag may still carry a third / fourth value.
    for (int i = nd; i < 4; ++i) {
      ws();
      if (s < e && ((*s >= '0' && *s <= '9') || *s == '-' || *s == '+' || *s == '.')) {
        if (!number(&v[i])) return false;
      } else {
        break;
      }
    }
    pts.push_back(v[0]);
    pts.push_back(v[1]);
    return true;
  }
  bool coord_list(int nd, std::vector<double>& pts) {
    if (!expect('(')) return false;
    if (!coord(nd, pts)) return false;
    while (peek(',')) {
      ++s;
      if (!coord(nd, pts)) return false;
    }
    return expect(')');
  }
  bool polygon_body(int nd, std::vector<std::vector<double>>& rings) {
    if (!expect('(')) return false;
    do {
      std::vector<double> r;
      if (is_empty()) {
        rings.push_back(r);
      } else {
        if (!coord_list(nd, r)) return false;
        close_ring(r);
        rings.push_back(std::move(r));
      }
      if (!peek(',')) break;
      ++s;
    } while (true);
    return expect(')');
  }
  bool geometry() {
    if (depth >= kMaxNesting) {
      err = "geometry nested deeper than 64 levels";
      return false;
    }
    ++depth;
    const bool okg = geometry_body();
    --depth;
    return okg;
  }
  bool geometry_body() {
    std::string kw = word();
    auto starts = [&](const char* b) { return kw.rfind(b, 0) == 0; };
    if (starts("GEOMETRYCOLLECTION")) {
      dims(kw, "GEOMETRYCOLLECTION");
      if (is_empty()) return true;
      if (!expect('(')) return false;
      do {
        if (!geometry()) return false;
        if (!peek(',')) break;
        ++s;
      } while (true);
      return expect(')');
    }
    if (starts("MULTIPOLYGON")) {
      int nd = dims(kw, "MULTIPOLYGON");
      out->begin_item(RZ_ITEM_POLYGONS);
      if (is_empty()) return true;
      if (!expect('(')) return false;
      do {
        if (is_empty()) {
        } else {
          std::vector<std::vector<double>> rings;
          if (!polygon_body(nd, rings)) return false;
          for (size_t k = 0; k < rings.size(); ++k)
            out->add_ring(rings[k].data(), (int64_t)rings[k].size() / 2, k > 0);
        }
        if (!peek(',')) break;
        ++s;
      } while (true);
      return expect(')');
    }
    if (starts("MULTILINESTRING")) {
      int nd = dims(kw, "MULTILINESTRING");
      out->begin_item(RZ_ITEM_LINES);
      if (is_empty()) return true;
      if (!expect('(')) return false;
      do {
        std::vector<double> r;
        if (!is_empty()) {
          if (!coord_list(nd, r)) return false;
        }
        out->add_ring(r.data(), (int64_t)r.size() / 2);
        if (!peek(',')) break;
        ++s;
      } while (true);
      return expect(')');
    }
    if (starts("MULTIPOINT")) {
      int nd = dims(kw, "MULTIPOINT");
      out->begin_item(RZ_ITEM_POINTS);
      if (is_empty()) return true;
      if (!expect('(')) return false;
      std::vector<double> pts;
      do {
        if (is_empty()) {
        } else if (peek('(')) {
          ++s;
          if (!coord(nd, pts)) return false;
          if (!expect(')')) return false;
        } else {
          if (!coord(nd, pts)) return false;
        }
        if (!peek(',')) break;
        ++s;
      } while (true);
      out->add_ring(pts.data(), (int64_t)pts.size() / 2);
      return expect(')');
    }
    if (starts("POLYGON")) {
      int nd = dims(kw, "POLYGON");
      out->begin_item(RZ_ITEM_POLYGONS);
      if (is_empty()) return true;
      std::vector<std::vector<double>> rings;
      if (!polygon_body(nd, rings)) return false;
      for (size_t k = 0; k < rings.size(); ++k)
        out->add_ring(rings[k].data(), (int64_t)rings[k].size() / 2, k > 0);
      return true;
    }
    if (starts("LINESTRING")) {
      int nd = dims(kw, "LINESTRING");
      out->begin_item(RZ_ITEM_LINES);
      std::vector<double> r;
      if (!is_empty()) {
        if (!coord_list(nd, r)) return false;
      }
      out->add_ring(r.data(), (int64_t)r.size() / 2);
      return true;
    }
    if (starts("POINT")) {
      int nd = dims(kw, "POINT");
      if (is_empty()) return true;  // an empty point burns nothing
      std::vector<double> p;
      if (!expect('(')) return false;
      if (!coord(nd, p)) return false;
      if (!expect(')')) return false;
      out->begin_item(RZ_ITEM_POINTS);
      out->add_ring(p.data(), 1);
      return true;
    }
    err = "unknown geometry type '" + kw + "'";
    return false;
  }
};

// ---------------------------------------------------------------- WKB ------

struct WkbParser {
  const uint8_t* p;
  const uint8_t* e;
  rz_geometry_buffers* out;
  std::string err;
  int depth = 0;

  bool u8(uint8_t* v) {
    if (p + 1 > e) return fail_trunc();
    *v = *p++;
    return true;
  }
  bool fail_trunc() {
    err = "truncated WKB";
    return false;
  }
  bool u32(bool le, uint32_t* v) {
    if (p + 4 > e) return fail_trunc();
    uint32_t x = le ? (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24
                    : (uint32_t)p[3] | (uint32_t)p[2] << 8 | (uint32_t)p[1] << 16 | (uint32_t)p[0] << 24;
    p += 4;
    *v = x;
    return true;
  }
  bool f64(bool le, double* v) {
    if (p + 8 > e) return fail_trunc();
    uint64_t x = 0;
    for (int i = 0; i < 8; ++i) x |= (uint64_t)p[le ? i : 7 - i] << (8 * i);
    p += 8;
    std::memcpy(v, &x, 8);
    return true;
  }
  // Fails unless n records of at least `bytes` bytes each can follow.
  bool fits(uint32_t n, size_t bytes) {
    if ((uint64_t)n * bytes > (uint64_t)(e - p)) return fail_trunc();
    return true;
  }
  bool coords(bool le, int nd, uint32_t n, std::vector<double>& pts) {
    if (!fits(n, 8u * (size_t)nd)) return false;
    pts.reserve(pts.size() + 2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
      double v[4];
      for (int k = 0; k < nd; ++k)
        if (!f64(le, &v[k])) return false;
      pts.push_back(v[0]);
      pts.push_back(v[1]);
    }
    return true;
  }
  // Parses one WKB geometry. `top` marks the outermost level. `skip` is set
  // when the geometry is an empty point (geo cannot represent it).
  bool geometry(bool* skip) {
    if (depth >= kMaxNesting) {
      err = "geometry nested deeper than 64 levels";
      return false;
    }
    ++depth;
    const bool okg = geometry_body(skip);
    --depth;
    return okg;
  }
  bool geometry_body(bool* skip) {
    uint8_t bo;
    if (!u8(&bo)) return false;
    if (bo > 1) {
      err = "invalid WKB byte order";
      return false;
    }
    bool le = bo == 1;
    uint32_t t;
    if (!u32(le, &t)) return false;
    int nd = 2;
    if (t & 0x80000000u) nd++;
    if (t & 0x40000000u) nd++;
    if (t & 0x20000000u) {  // EWKB SRID
      uint32_t srid;
      if (!u32(le, &srid)) return false;
    }
    t &= 0x0FFFFFFFu;
    uint32_t base = t % 1000, iso = t / 1000;
    if (iso == 1 || iso == 2) nd = 3;
    else if (iso == 3) nd = 4;
    else if (iso != 0) {
      err = "unsupported WKB dimension";
      return false;
    }
    uint32_t n;
    switch (base) {
      case 1: {  // Point
        std::vector<double> pt;
        if (!coords(le, nd, 1, pt)) return false;
        if (std::isnan(pt[0]) && std::isnan(pt[1])) {  // POINT EMPTY
          if (skip) *skip = true;
          return true;
        }
        out->begin_item(RZ_ITEM_POINTS);
        out->add_ring(pt.data(), 1);
        return true;
      }
      case 2: {  // LineString
        std::vector<double> pts;
        if (!u32(le, &n) || !coords(le, nd, n, pts)) return false;
        out->begin_item(RZ_ITEM_LINES);
        out->add_ring(pts.data(), (int64_t)n);
        return true;
      }
      case 3: {  // Polygon
        if (!u32(le, &n) || !fits(n, 4)) return false;  // each ring holds a 4-byte count
        std::vector<std::vector<double>> rings(n);
        for (uint32_t r = 0; r < n; ++r) {
          uint32_t m;
          if (!u32(le, &m) || !coords(le, nd, m, rings[r])) return false;
          close_ring(rings[r]);
        }
        out->begin_item(RZ_ITEM_POLYGONS);
        for (size_t k = 0; k < rings.size(); ++k)
          out->add_ring(rings[k].data(), (int64_t)rings[k].size() / 2, k > 0);
        return true;
      }
      case 4: {  // MultiPoint
        if (!u32(le, &n) || !fits(n, 21)) return false;  // a point member is 21 bytes at least
        std::vector<double> pts;
        for (uint32_t i = 0; i < n; ++i) {
          uint8_t b2;
          uint32_t t2;
          if (!u8(&b2)) return false;
          bool le2 = b2 == 1;
          if (!u32(le2, &t2)) return false;
          int nd2 = 2 + ((t2 & 0x80000000u) ? 1 : 0) + ((t2 & 0x40000000u) ? 1 : 0);
          if (t2 & 0x20000000u) {
            uint32_t srid;
            if (!u32(le2, &srid)) return false;
          }
          t2 &= 0x0FFFFFFFu;
          if (t2 / 1000 == 1 || t2 / 1000 == 2) nd2 = 3;
          else if (t2 / 1000 == 3) nd2 = 4;
          if (t2 % 1000 != 1) {
            err = "MultiPoint member is not a Point";
            return false;
          }
          std::vector<double> pt;
          if (!coords(le2, nd2, 1, pt)) return false;
          if (std::isnan(pt[0]) && std::isnan(pt[1])) continue;
          pts.insert(pts.end(), pt.begin(), pt.end());
        }
        out->begin_item(RZ_ITEM_POINTS);
        out->add_ring(pts.data(), (int64_t)pts.size() / 2);
        return true;
      }
      case 5:    // MultiLineString
      case 6: {  // MultiPolygon
        if (!u32(le, &n) || !fits(n, 9)) return false;  // a member header is 9 bytes at least
        rz_geometry_buffers tmp;
        WkbParser sub{p, e, &tmp, {}, depth};
        for (uint32_t i = 0; i < n; ++i)
          if (!sub.geometry(nullptr)) {
            err = sub.err;
            return false;
          }
        p = sub.p;
        out->begin_item(base == 5 ? RZ_ITEM_LINES : RZ_ITEM_POLYGONS);
        for (size_t it = 0; it < tmp.item_kind.size(); ++it) {
          if (tmp.item_kind[it] != (base == 5 ? RZ_ITEM_LINES : RZ_ITEM_POLYGONS)) {
            err = "Multi geometry member of the wrong type";
            return false;
          }
          for (int64_t r = tmp.item_ring_offsets[it]; r < tmp.item_ring_offsets[it + 1]; ++r) {
            int64_t v0 = tmp.ring_vertex_offsets[r], v1 = tmp.ring_vertex_offsets[r + 1];
            out->add_ring(tmp.xy.data() + 2 * v0, v1 - v0, tmp.ring_flags[r] & 1);
          }
        }
        return true;
      }
      case 7: {  // GeometryCollection
        if (!u32(le, &n) || !fits(n, 9)) return false;
        for (uint32_t i = 0; i < n; ++i) {
          bool s2 = false;
          if (!geometry(&s2)) return false;
        }
        return true;
      }
      default:
        err = "unsupported WKB geometry type " + std::to_string(t);
        return false;
    }
  }
};

}  // namespace

// The parser bodies allocate with std::vector; an allocation failure (or any
// other exception) is mapped to a status here, so nothing propagates across
// the C boundary.
template <typename F>
static rz_status guarded(F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return rz::fail(RZ_CAPACITY_ERROR, "Cannot parse geometry: out of host memory");
  } catch (const std::exception& ex) {
    return rz::fail(RZ_PARSE_ERROR, std::string("Cannot parse geometry: ") + ex.what());
  }
}

static rz_status from_wkt(const char* const* texts, const int64_t* lengths, int64_t n,
                          rz_geometry_buffers** out) {
  if (n < 0 || (n > 0 && (!texts || !lengths)) || !out)
    return rz::fail(RZ_VALUE_ERROR, "bad argument to rz_geometry_from_wkt");
  auto* b = new (std::nothrow) rz_geometry_buffers();
  if (!b) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  *out = b;  // the caller frees it if a parser throws
  for (int64_t i = 0; i < n; ++i) {
    if (lengths[i] < 0 || (lengths[i] > 0 && !texts[i])) {
      delete b;
      *out = nullptr;
      return rz::fail(RZ_VALUE_ERROR, "bad WKT length");
    }
    WktParser ps{texts[i], texts[i] + lengths[i], b, {}};
    bool okp = ps.geometry();
    ps.ws();
    if (okp && ps.s != ps.e) {
      okp = false;
      ps.err = "trailing characters";
    }
    if (!okp) {
      delete b;
      *out = nullptr;
      return rz::fail(RZ_PARSE_ERROR, "Cannot parse geometry. Check that the WKT is valid. (" +
                                           ps.err + " in geometry " + std::to_string(i) + ")");
    }
    b->end_geometry();
  }
  *out = b;
  return RZ_OK;
}

extern "C" rz_status rz_geometry_from_wkt(const char* const* texts, const int64_t* lengths,
                                          int64_t n, rz_geometry_buffers** out) {
  rz_geometry_buffers* b = nullptr;
  rz_status st = guarded([&] { return from_wkt(texts, lengths, n, &b); });
  if (st != RZ_OK) {
    delete b;  // null unless a parser threw
    return st;
  }
  *out = b;
  return RZ_OK;
}

static rz_status from_wkb(const uint8_t* const* blobs, const int64_t* lengths, int64_t n,
                          rz_geometry_buffers** out) {
  if (n < 0 || (n > 0 && (!blobs || !lengths)) || !out)
    return rz::fail(RZ_VALUE_ERROR, "bad argument to rz_geometry_from_wkb");
  auto* b = new (std::nothrow) rz_geometry_buffers();
  if (!b) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  *out = b;  // the caller frees it if a parser throws
  for (int64_t i = 0; i < n; ++i) {
    size_t items_before = b->item_kind.size();
    size_t rings_before = b->ring_vertex_offsets.size();
    if (lengths[i] < 0 || (lengths[i] > 0 && !blobs[i])) {
      delete b;
      *out = nullptr;
      return rz::fail(RZ_VALUE_ERROR, "bad WKB length");
    }
    WkbParser ps{blobs[i], blobs[i] + lengths[i], b, {}};
    bool skip = false;
    if (!ps.geometry(&skip)) {
      delete b;
      *out = nullptr;
      return rz::fail(RZ_PARSE_ERROR,
                      "Cannot parse geometry. Check that the WKB bytes are valid. (" + ps.err +
                          " in geometry " + std::to_string(i) + ")");
    }
    if (skip) {  // parse_geometry.rs:113-116 drops geometries geo cannot build
      b->item_kind.resize(items_before);
      b->item_ring_offsets.resize(items_before + 1);
      b->ring_vertex_offsets.resize(rings_before);
      b->ring_flags.resize(rings_before - 1);
      b->xy.resize(2 * (size_t)b->ring_vertex_offsets.back());
      continue;
    }
    b->end_geometry();
  }
  *out = b;
  return RZ_OK;
}

extern "C" rz_status rz_geometry_from_wkb(const uint8_t* const* blobs, const int64_t* lengths,
                                          int64_t n, rz_geometry_buffers** out) {
  rz_geometry_buffers* b = nullptr;
  rz_status st = guarded([&] { return from_wkb(blobs, lengths, n, &b); });
  if (st != RZ_OK) {
    delete b;  // null unless a parser threw
    return st;
  }
  *out = b;
  return RZ_OK;
}

// --------------------------------------------------------- seeded inputs ---

namespace {
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1), 53 bits
  double uni(double a, double b) { return a + (b - a) * u01(); }
  int64_t irange(int64_t lo, int64_t hi) {  // inclusive
    return lo + (int64_t)(u01() * (double)(hi - lo + 1));
  }
  double loguni(double a, double b) { return std::exp(uni(std::log(a), std::log(b))); }
  double normal() {  // Box-Muller in f64
    double u1 = u01(), u2 = u01();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

const double kTwoPi = 6.283185307179586;

// Convex polygon inscribed in a circle: n sorted uniform angles.
void convex_ring(SplitMix64& rng, double cx, double cy, double radius, int n,
                 std::vector<double>& pts) {
  std::vector<double> ang(n);
  for (auto& a : ang) a = rng.uni(0.0, kTwoPi);
  std::sort(ang.begin(), ang.end());
  pts.clear();
  for (int k = 0; k < n; ++k) {
    pts.push_back(cx + radius * std::cos(ang[k]));
    pts.push_back(cy + radius * std::sin(ang[k]));
  }
  pts.push_back(pts[0]);
  pts.push_back(pts[1]);
}

// Star-shaped polygon: sorted uniform angles, radii in [rlo, 1] * R.
void star_ring(SplitMix64& rng, double cx, double cy, double R, int n, double rlo,
               std::vector<double>& pts) {
  std::vector<double> ang(n);
  for (auto& a : ang) a = rng.uni(0.0, kTwoPi);
  std::sort(ang.begin(), ang.end());
  pts.clear();
  for (int k = 0; k < n; ++k) {
    double r = rng.uni(rlo, 1.0) * R;
    pts.push_back(cx + r * std::cos(ang[k]));
    pts.push_back(cy + r * std::sin(ang[k]));
  }
  pts.push_back(pts[0]);
  pts.push_back(pts[1]);
}

void random_walk(SplitMix64& rng, double W, double H, int n, double slo, double shi,
                 std::vector<double>& pts) {
  pts.clear();
  double x = rng.uni(0.0, W), y = rng.uni(0.0, H);
  pts.push_back(x);
  pts.push_back(y);
  for (int k = 1; k < n; ++k) {
    double step = rng.uni(slo, shi), h = rng.uni(0.0, kTwoPi);
    x += step * std::cos(h);
    y += step * std::sin(h);
    pts.push_back(x);
    pts.push_back(y);
  }
}
}  // namespace

// kind 1: `count` star polygons, 8-64 vertices, R ~ U(p0, p1), radii
//         U(0.3, 1) * R, centres uniform over [0, grid_w) x [0, grid_h).
// kind 2: `count` convex polygons, 4-12 vertices, diameter log-uniform
//         [p0, p1]; a fraction p2 are MultiPolygons of 2-4 parts and a
//         fraction p3 carry 1-3 holes of half the diameter.
// kind 3: p0 random-walk LineStrings (2-200 vertices, step U(1, 20), uniform
//         heading), p1 Points and p2 GeometryCollections (a point, a short
//         walk and a small convex polygon each), shuffled; `count` ignored.
// kind 4: `count` star polygons with vertex counts log-uniform [p2, p3] and
//         radius log-uniform [p0, p1] (a batched-job draw).
extern "C" rz_status rz_synth_geometry(int kind, uint64_t seed, int64_t count, double W, double H,
                                       double p0, double p1, double p2, double p3,
                                       rz_geometry_buffers** out) {
  auto* b = new (std::nothrow) rz_geometry_buffers();
  if (!b) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  SplitMix64 rng{seed};
  std::vector<double> pts;
  if (kind == 1) {
    for (int64_t i = 0; i < count; ++i) {
      int n = (int)rng.irange(8, 64);
      double R = rng.uni(p0, p1);
      double cx = rng.uni(0.0, W), cy = rng.uni(0.0, H);
      star_ring(rng, cx, cy, R, n, 0.3, pts);
      b->begin_item(RZ_ITEM_POLYGONS);
      b->add_ring(pts.data(), (int64_t)pts.size() / 2);
      b->end_geometry();
    }
  } else if (kind == 2) {
    for (int64_t i = 0; i < count; ++i) {
      double d = rng.loguni(p0, p1);
      double cx = rng.uni(0.0, W), cy = rng.uni(0.0, H);
      int parts = rng.u01() < p2 ? (int)rng.irange(2, 4) : 1;
      int holes = rng.u01() < p3 ? (int)rng.irange(1, 3) : 0;
      b->begin_item(RZ_ITEM_POLYGONS);
      for (int k = 0; k < parts; ++k) {
        double px = cx + 1.1 * d * k, py = cy;
        int n = (int)rng.irange(4, 12);
        convex_ring(rng, px, py, d / 2.0, n, pts);
        b->add_ring(pts.data(), (int64_t)pts.size() / 2);
        for (int h = 0; h < holes; ++h) {
          double ang = kTwoPi * h / holes + rng.uni(0.0, 1.0);
          double hx = px + d / 8.0 * std::cos(ang), hy = py + d / 8.0 * std::sin(ang);
          int nh = (int)rng.irange(4, 12);
          convex_ring(rng, hx, hy, d / 4.0, nh, pts);
          b->add_ring(pts.data(), (int64_t)pts.size() / 2, true);
        }
      }
      b->end_geometry();
    }
  } else if (kind == 3) {
    int64_t nl = (int64_t)p0, np = (int64_t)p1, nc = (int64_t)p2;
    std::vector<uint8_t> type;
    type.insert(type.end(), (size_t)nl, 0);
    type.insert(type.end(), (size_t)np, 1);
    type.insert(type.end(), (size_t)nc, 2);
    for (size_t i = type.size(); i > 1; --i) {  // Fisher-Yates
      size_t j = (size_t)(rng.u01() * (double)i);
      std::swap(type[i - 1], type[j]);
    }
    for (uint8_t t : type) {
      if (t == 0) {
        random_walk(rng, W, H, (int)rng.irange(2, 200), 1.0, 20.0, pts);
        b->begin_item(RZ_ITEM_LINES);
        b->add_ring(pts.data(), (int64_t)pts.size() / 2);
      } else if (t == 1) {
        double pt[2] = {rng.uni(0.0, W), rng.uni(0.0, H)};
        b->begin_item(RZ_ITEM_POINTS);
        b->add_ring(pt, 1);
      } else {
        double pt[2] = {rng.uni(0.0, W), rng.uni(0.0, H)};
        b->begin_item(RZ_ITEM_POINTS);
        b->add_ring(pt, 1);
        random_walk(rng, W, H, (int)rng.irange(2, 20), 1.0, 20.0, pts);
        b->begin_item(RZ_ITEM_LINES);
        b->add_ring(pts.data(), (int64_t)pts.size() / 2);
        double cx = rng.uni(0.0, W), cy = rng.uni(0.0, H);
        convex_ring(rng, cx, cy, rng.uni(2.0, 20.0), (int)rng.irange(4, 12), pts);
        b->begin_item(RZ_ITEM_POLYGONS);
        b->add_ring(pts.data(), (int64_t)pts.size() / 2);
      }
      b->end_geometry();
    }
  } else if (kind == 4) {
    for (int64_t i = 0; i < count; ++i) {
      int n = (int)std::lround(rng.loguni(p2, p3));
      if (n < 3) n = 3;
      double R = rng.loguni(p0, p1);
      double cx = rng.uni(0.0, W), cy = rng.uni(0.0, H);
      star_ring(rng, cx, cy, R, n, 0.5, pts);
      b->begin_item(RZ_ITEM_POLYGONS);
      b->add_ring(pts.data(), (int64_t)pts.size() / 2);
      b->end_geometry();
    }
  } else {
    delete b;
    return rz::fail(RZ_VALUE_ERROR, "unknown synthetic geometry kind");
  }
  *out = b;
  return RZ_OK;
}

// dist 0: integer uniform on [lo, hi) cast to T (numpy astype: integers
//         wrap modulo 2^bits, floats convert exactly below 2^24 / 2^53);
// dist 1: standard normal (Box-Muller in f64) rounded to T;
// dist 2: real uniform on [lo, hi) rounded to T.
template <typename T>
static void fill_values(SplitMix64& rng, int dist, int64_t n, double lo, double hi, T* out) {
  for (int64_t i = 0; i < n; ++i) {
    if (dist == 1) {
      out[i] = (T)rng.normal();
    } else if (dist == 0) {
      int64_t span = (int64_t)(hi - lo);
      int64_t v = (int64_t)lo + (int64_t)(rng.u01() * (double)span);
      if (std::is_floating_point<T>::value) out[i] = (T)v;
      else out[i] = (T)(uint64_t)v;
    } else {
      out[i] = (T)rng.uni(lo, hi);
    }
  }
}

extern "C" rz_status rz_synth_field(rz_dtype dtype, int dist, uint64_t seed, int64_t n, double lo,
                                    double hi, void* out) {
  SplitMix64 rng{seed};
  switch (dtype) {
    case RZ_U8: fill_values(rng, dist, n, lo, hi, (uint8_t*)out); break;
    case RZ_U16: fill_values(rng, dist, n, lo, hi, (uint16_t*)out); break;
    case RZ_U32: fill_values(rng, dist, n, lo, hi, (uint32_t*)out); break;
    case RZ_U64: fill_values(rng, dist, n, lo, hi, (uint64_t*)out); break;
    case RZ_I8: fill_values(rng, dist, n, lo, hi, (int8_t*)out); break;
    case RZ_I16: fill_values(rng, dist, n, lo, hi, (int16_t*)out); break;
    case RZ_I32: fill_values(rng, dist, n, lo, hi, (int32_t*)out); break;
    case RZ_I64: fill_values(rng, dist, n, lo, hi, (int64_t*)out); break;
    case RZ_F32: fill_values(rng, dist, n, lo, hi, (float*)out); break;
    case RZ_F64: fill_values(rng, dist, n, lo, hi, (double*)out); break;
    default: return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
  }
  return RZ_OK;
}

// ------------------------------------------------- batched-job generator --

extern "C" rz_status rz_synth_job_spec(uint64_t seed, int64_t k, int32_t* side, int32_t* n_poly,
                                       int32_t* fun, int64_t* n_vertices) {
  if (k < 0) return rz::fail(RZ_VALUE_ERROR, "job index must be >= 0");
  const rzsynth::JobSpec j = rzsynth::job_spec(seed, (uint64_t)k);
  if (side) *side = j.side;
  if (n_poly) *n_poly = j.n_poly;
  if (fun) *fun = j.fun;
  if (n_vertices) {
    int64_t v = 0;
    for (int32_t p = 0; p < j.n_poly; ++p)
      v += rzsynth::poly_spec(seed, (uint64_t)k, (uint64_t)p, j.side).n_vert + 1;
    *n_vertices = v;
  }
  return RZ_OK;
}

extern "C" rz_status rz_synth_job_geometry(uint64_t seed, int64_t k, rz_geometry_buffers** out) {
  if (k < 0) return rz::fail(RZ_VALUE_ERROR, "job index must be >= 0");
  auto* b = new (std::nothrow) rz_geometry_buffers();
  if (!b) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  const rzsynth::JobSpec j = rzsynth::job_spec(seed, (uint64_t)k);
  std::vector<double> pts;
  for (int32_t p = 0; p < j.n_poly; ++p) {
    const rzsynth::PolySpec ps = rzsynth::poly_spec(seed, (uint64_t)k, (uint64_t)p, j.side);
    pts.resize(2 * (size_t)(ps.n_vert + 1));
    for (int32_t v = 0; v <= ps.n_vert; ++v)
      rzsynth::vertex(seed, (uint64_t)k, (uint64_t)p, ps, v, &pts[2 * (size_t)v], &pts[2 * (size_t)v + 1]);
    b->begin_item(RZ_ITEM_POLYGONS);
    b->add_ring(pts.data(), ps.n_vert + 1);
    b->end_geometry();
  }
  *out = b;
  return RZ_OK;
}

extern "C" rz_status rz_synth_job_field(uint64_t seed, int64_t k, double* out) {
  if (k < 0 || !out) return rz::fail(RZ_VALUE_ERROR, "bad job index or NULL output");
  const rzsynth::JobSpec j = rzsynth::job_spec(seed, (uint64_t)k);
  for (int32_t p = 0; p < j.n_poly; ++p) out[p] = rzsynth::field(seed, (uint64_t)k, (uint64_t)p);
  return RZ_OK;
}
