// Initial state of propagate() on the host: WKB ignition and fuel features
// burned into rasters the way the reference's GDALRasterizeGeometries calls
// do (crates/propag/src/propag.rs:837-890 rasterize_times, 929-978
// rasterize_fuels), and the initial references (propag.rs:433-449).
//
// GDAL itself is not linked.  The burn rules are GDAL 3.x's defaults for a
// call with no options and no transformer (ALL_TOUCHED off, MERGE_ALG
// REPLACE), restated here:
//   * pixel/line coordinates come from the inverse of the raster's
//     geotransform (the f32 GeoReference values widened to double), inverted
//     as GDALInvGeoTransform does;
//   * a geometry flattens into parts (a point, a line string, each polygon
//     ring; members of multi-geometries and collections in order) and burns
//     by its own type: points the pixel (floor x, floor y); line strings a
//     Bresenham line between the floored vertices of each segment; polygons,
//     multipolygons and collections a scan fill of the pixels whose centres
//     are inside: the scan line y + 0.5 crosses every non-horizontal edge
//     whose low end is at or below it and whose high end above it, crossings
//     round as floor(x + 0.5), sorted pairs [a, b) burn a..b-1, and a
//     horizontal edge on the scan line burns [floor(lo + 0.5), floor(hi + 0.5))
//     when it runs right to left;
//   * burns clip to the raster; later features replace earlier ones.
// A feature with a vertex that is not finite or lies 2^30 pixels or more
// from the origin burns nothing.
// A feature whose WKB does not parse is skipped, as the reference's
// filter_map over TryFrom does (propag.rs:208-215, 902-907).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "pg_internal.h"

namespace pg {
namespace {

enum GeomType { kPoint = 1, kLine = 2, kPolygon = 3, kMultiPoint = 4, kMultiLine = 5,
                kMultiPolygon = 6, kCollection = 7 };

// Flattened geometry: vertices of every part back to back.
struct Shape {
  std::vector<double> x, y;
  std::vector<size_t> part;  // vertex count of each part
};

class WkbReader {
 public:
  WkbReader(const uint8_t *p, size_t n) : p_(p), n_(n) {}

  // Parses one geometry at the cursor; appends its parts to `s`.  Returns
  // the geometry's base type, or 0 when the bytes are not valid WKB.
  int read(Shape &s, int depth = 0) {
    if (depth > 32 || off_ + 5 > n_) return 0;
    const uint8_t order = p_[off_];
    if (order > 1) return 0;
    le_ = order == 1;
    off_ += 1;
    uint32_t code = u32();
    // OGC 2.5D (0x80000000) and M (0x40000000) flags, or ISO +1000 (Z),
    // +2000 (M), +3000 (ZM); the EWKB SRID flag and mixed forms are invalid.
    const bool fz = code & 0x80000000u, fm = code & 0x40000000u;
    if (code & 0x30000000u) return 0;
    code &= 0x0FFFFFFFu;
    if ((fz || fm) && code >= 1000) return 0;
    if (code >= 4000) return 0;
    static const int extra[4] = {0, 1, 1, 2};
    dims_ = 2 + (fz ? 1 : 0) + (fm ? 1 : 0) + extra[code / 1000];
    const int base = static_cast<int>(code % 1000);
    const bool le = le_;
    const int dims = dims_;
    switch (base) {
      case kPoint: {
        double x, y;
        if (!point(x, y)) return 0;
        if (!(std::isnan(x) && std::isnan(y))) add_part(s, &x, &y, 1);  // else POINT EMPTY
        return kPoint;
      }
      case kLine: {
        if (!ring(s)) return 0;
        return kLine;
      }
      case kPolygon: {
        uint32_t nr;
        if (!count(nr)) return 0;
        for (uint32_t r = 0; r < nr; ++r)
          if (!ring(s)) return 0;
        return kPolygon;
      }
      case kMultiPoint:
      case kMultiLine:
      case kMultiPolygon:
      case kCollection: {
        uint32_t ng;
        if (!count(ng)) return 0;
        const int want = base == kCollection ? 0 : base - 3;
        for (uint32_t g = 0; g < ng; ++g) {
          const int t = read(s, depth + 1);
          if (t == 0 || (want != 0 && t != want)) return 0;
        }
        le_ = le;
        dims_ = dims;
        return base;
      }
      default:
        return 0;
    }
  }

 private:
  uint32_t u32() {
    uint32_t v;
    std::memcpy(&v, p_ + off_, 4);
    if (!le_) v = __builtin_bswap32(v);
    off_ += 4;
    return v;
  }
  double f64() {
    uint64_t v;
    std::memcpy(&v, p_ + off_, 8);
    if (!le_) v = __builtin_bswap64(v);
    off_ += 8;
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
  bool count(uint32_t &c) {
    if (off_ + 4 > n_) return false;
    c = u32();
    return true;
  }
  bool point(double &x, double &y) {
    if (off_ + 8 * static_cast<size_t>(dims_) > n_) return false;
    x = f64();
    y = f64();
    off_ += 8 * static_cast<size_t>(dims_ - 2);
    return true;
  }
  // A vertex count and its vertices, as one part (none when empty).
  bool ring(Shape &s) {
    uint32_t n;
    if (!count(n)) return false;
    if (static_cast<size_t>(n) * 8 * dims_ > n_ - off_) return false;
    const size_t at = s.x.size();
    s.x.resize(at + n);
    s.y.resize(at + n);
    for (uint32_t i = 0; i < n; ++i) point(s.x[at + i], s.y[at + i]);
    if (n) s.part.push_back(n);
    return true;
  }
  static void add_part(Shape &s, const double *x, const double *y, size_t n) {
    s.x.insert(s.x.end(), x, x + n);
    s.y.insert(s.y.end(), y, y + n);
    s.part.push_back(n);
  }

  const uint8_t *p_;
  size_t n_, off_ = 0;
  bool le_ = true;
  int dims_ = 2;
};

// GDALInvGeoTransform of the GDAL-ordered transform g.
bool invert(const double g[6], double inv[6]) {
  if (g[2] == 0.0 && g[4] == 0.0 && g[1] != 0.0 && g[5] != 0.0) {
    inv[0] = -g[0] / g[1];
    inv[1] = 1.0 / g[1];
    inv[2] = 0.0;
    inv[3] = -g[3] / g[5];
    inv[4] = 0.0;
    inv[5] = 1.0 / g[5];
    return true;
  }
  const double det = g[1] * g[5] - g[2] * g[4];
  const double mag = std::max(std::max(std::fabs(g[1]), std::fabs(g[2])),
                              std::max(std::fabs(g[4]), std::fabs(g[5])));
  if (std::fabs(det) <= 1e-10 * mag * mag) return false;
  const double r = 1.0 / det;
  inv[0] = (g[2] * g[3] - g[0] * g[5]) * r;
  inv[1] = g[5] * r;
  inv[2] = -g[2] * r;
  inv[3] = (-g[1] * g[3] + g[0] * g[4]) * r;
  inv[4] = -g[4] * r;
  inv[5] = g[1] * r;
  return true;
}

template <class V>
struct Raster {
  V *px;
  int w, h;
  void set(int x, int y, V v) const {
    if (x >= 0 && x < w && y >= 0 && y < h) px[static_cast<size_t>(y) * w + x] = v;
  }
  // Pixels a..b of row y, clipped to the raster.
  void span(int y, int a, int b, V v) const {
    a = std::max(a, 0);
    b = std::min(b, w - 1);
    for (int x = a; x <= b; ++x) px[static_cast<size_t>(y) * w + x] = v;
  }
};

inline int ifloor(double v) { return static_cast<int>(std::floor(v)); }

template <class V>
void burn_points(const Raster<V> &r, const Shape &s, V v) {
  for (size_t i = 0; i < s.x.size(); ++i) r.set(ifloor(s.x[i]), ifloor(s.y[i]), v);
}

template <class V>
void burn_lines(const Raster<V> &r, const Shape &s, V v) {
  size_t base = 0;
  for (size_t n : s.part) {
    for (size_t j = 1; j < n; ++j) {
      int ix = ifloor(s.x[base + j - 1]), iy = ifloor(s.y[base + j - 1]);
      const int ix1 = ifloor(s.x[base + j]), iy1 = ifloor(s.y[base + j]);
      const int ddx = std::abs(ix1 - ix), ddy = std::abs(iy1 - iy);
      const int sx = ix > ix1 ? -1 : 1, sy = iy > iy1 ? -1 : 1;
      const bool xmajor = ddx >= ddy;
      const int major = xmajor ? ddx : ddy, minor = xmajor ? ddy : ddx;
      const int e_minor = minor * 2, e_major = e_minor - major * 2;
      int err = e_minor - major;
      for (int k = 0; k <= major; ++k) {
        r.set(ix, iy, v);
        if (xmajor) {
          ix += sx;
          if (err > 0) { iy += sy; err += e_major; } else { err += e_minor; }
        } else {
          iy += sy;
          if (err > 0) { ix += sx; err += e_major; } else { err += e_minor; }
        }
      }
    }
    base += n;
  }
}

template <class V>
void burn_polygon(const Raster<V> &r, const Shape &s, V v) {
  const size_t nv = s.x.size();
  if (nv == 0) return;
  double ylo = s.y[0], yhi = s.y[0];
  for (size_t i = 1; i < nv; ++i) {
    ylo = std::min(ylo, s.y[i]);
    yhi = std::max(yhi, s.y[i]);
  }
  const int miny = std::max(static_cast<int>(ylo), 0);
  const int maxy = std::min(static_cast<int>(yhi), r.h - 1);
  const int maxx = r.w - 1;
  std::vector<int> xs;
  xs.reserve(nv);
  for (int y = miny; y <= maxy; ++y) {
    const double dy = y + 0.5;
    xs.clear();
    size_t base = 0;
    for (size_t n : s.part) {
      for (size_t i = 0; i < n; ++i) {
        const size_t i1 = base + (i == 0 ? n - 1 : i - 1), i2 = base + i;
        double y1 = s.y[i1], y2 = s.y[i2], x1, x2;
        if ((y1 < dy && y2 < dy) || (y1 > dy && y2 > dy)) continue;
        if (y1 < y2) {
          x1 = s.x[i1];
          x2 = s.x[i2];
        } else if (y1 > y2) {
          std::swap(y1, y2);
          x1 = s.x[i2];
          x2 = s.x[i1];
        } else {
          if (s.x[i1] > s.x[i2]) {
            const int a = ifloor(s.x[i2] + 0.5), b = ifloor(s.x[i1] + 0.5);
            if (!(a > maxx || b <= 0)) r.span(y, a, b - 1, v);
          }
          continue;
        }
        if (dy < y2 && dy >= y1) xs.push_back(ifloor((dy - y1) * (x2 - x1) / (y2 - y1) + x1 + 0.5));
      }
      base += n;
    }
    std::sort(xs.begin(), xs.end());
    for (size_t k = 0; k + 1 < xs.size(); k += 2)
      if (xs[k] <= maxx && xs[k + 1] > 0) r.span(y, xs[k], xs[k + 1] - 1, v);
  }
}

// Burns one WKB feature; false when it does not parse.
template <class V>
bool burn(const Raster<V> &r, const double inv[6], const uint8_t *wkb, size_t len, V v) {
  if (!wkb || len == 0) return false;
  Shape s;
  WkbReader rd(wkb, len);
  const int type = rd.read(s);
  if (type == 0) return false;
  for (size_t i = 0; i < s.x.size(); ++i) {
    const double x = s.x[i], y = s.y[i];
    s.x[i] = inv[0] + x * inv[1] + y * inv[2];
    s.y[i] = inv[3] + x * inv[4] + y * inv[5];
  }
  // A vertex that is not finite or lies 2^30 pixels or more away burns
  // nothing of its feature (GDAL's integer casts would be undefined there).
  for (size_t i = 0; i < s.x.size(); ++i)
    if (!(std::fabs(s.x[i]) < 1073741824.0) || !(std::fabs(s.y[i]) < 1073741824.0)) return true;
  if (type == kPoint || type == kMultiPoint)
    burn_points(r, s, v);
  else if (type == kLine || type == kMultiLine)
    burn_lines(r, s, v);
  else
    burn_polygon(r, s, v);
  return true;
}

pg_status geo_inverse(const GeoReference *geo, const char *who, double inv[6]) {
  if (!geo) return set_error(PG_ERR_INVALID, std::string(who) + ": geo is NULL");
  const GT &t = geo->transform.gt;
  const double g[6] = {t.x0, t.dx, t.rx, t.y0, t.ry, t.dy};
  if (!invert(g, inv)) return set_error(PG_ERR_INVALID, std::string(who) + ": geotransform is not invertible");
  if (geo->width > 0x7fffffffu || geo->height > 0x7fffffffu)
    return set_error(PG_ERR_INVALID, std::string(who) + ": raster too large");
  return PG_OK;
}

}  // namespace
}  // namespace pg

extern "C" {

pg_status pg_rasterize_times(const GeoReference *geo, const FFITimeFeature *features, size_t n,
                             float *time, uint32_t *burned) {
  double inv[6];
  if (pg_status s = pg::geo_inverse(geo, "pg_rasterize_times", inv)) return s;
  const size_t cells = static_cast<size_t>(geo->width) * geo->height;
  if (!time && cells) return pg::set_error(PG_ERR_INVALID, "pg_rasterize_times: time is NULL");
  if (n && !features) return pg::set_error(PG_ERR_INVALID, "pg_rasterize_times: features is NULL");
  std::fill(time, time + cells, FLT_MAX);
  const pg::Raster<float> r{time, static_cast<int>(geo->width), static_cast<int>(geo->height)};
  uint32_t ok = 0;
  for (size_t i = 0; i < n; ++i)
    ok += pg::burn(r, inv, features[i].geom_wkb, features[i].geom_wkb_len, features[i].time) ? 1 : 0;
  if (burned) *burned = ok;
  return PG_OK;
}

pg_status pg_rasterize_fuels(const GeoReference *geo, const FFIFuelFeature *features, size_t n,
                             uint8_t *fuel, uint32_t *burned) {
  double inv[6];
  if (pg_status s = pg::geo_inverse(geo, "pg_rasterize_fuels", inv)) return s;
  const size_t cells = static_cast<size_t>(geo->width) * geo->height;
  if (!fuel && cells) return pg::set_error(PG_ERR_INVALID, "pg_rasterize_fuels: fuel is NULL");
  if (n && !features) return pg::set_error(PG_ERR_INVALID, "pg_rasterize_fuels: features is NULL");
  std::fill(fuel, fuel + cells, static_cast<uint8_t>(255));
  const pg::Raster<uint8_t> r{fuel, static_cast<int>(geo->width), static_cast<int>(geo->height)};
  uint32_t ok = 0;
  for (size_t i = 0; i < n; ++i)
    ok += pg::burn(r, inv, features[i].geom_wkb, features[i].geom_wkb_len, features[i].code) ? 1 : 0;
  if (burned) *burned = ok;
  return PG_OK;
}

pg_status pg_init_refs(const GeoReference *geo, const float *time, uint16_t *refs_x, uint16_t *refs_y) {
  if (!geo || !time || !refs_x || !refs_y) return pg::set_error(PG_ERR_INVALID, "pg_init_refs: null pointer");
  if (geo->width > 65535u || geo->height > 65535u)
    return pg::set_error(PG_ERR_INVALID, "pg_init_refs: references are 16-bit; raster sides must be below 65536");
  const uint32_t w = geo->width, h = geo->height;
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      const size_t i = static_cast<size_t>(y) * w + x;
      const bool lit = time[i] != FLT_MAX;
      refs_x[i] = lit ? static_cast<uint16_t>(x) : 0xFFFFu;
      refs_y[i] = lit ? static_cast<uint16_t>(y) : 0xFFFFu;
    }
  return PG_OK;
}

}  // extern "C"
