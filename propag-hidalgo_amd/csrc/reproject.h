// Coordinate transformation between proj4 CRSs on the host, without PROJ:
// geographic (longitude/latitude in degrees) and transverse Mercator (UTM
// zones and general tmerc) on one of the common ellipsoids.  The reference
// transforms every ignition and fuel geometry from its CRS to the raster's
// with OGR (Geometry::transform_to_inplace, crates/propag/src/propag.rs:
// 854-857, 964-969) and drops a geometry whose transformation fails; this
// header does the same for the CRSs it knows and names the ones it does not.
//
// Transverse Mercator: Krueger's series in the third flattening n, as
// L. Krueger (1912) derived them and C. F. F. Karney ("Transverse Mercator
// with an accuracy of a few nanometers", J. Geodesy 85, 2011, eqs. 14-15,
// 35-36) gives them, here to n^4 (a few 0.1 mm within 4000 km of the
// central meridian); the conformal latitude is formed and inverted exactly
// (an iteration of the isometric-latitude relation, Snyder 1987, eqs. 3-4
// and 7-9), so only the Gauss-Krueger part is a series.
//
// Datums: no datum shift is applied.  Two CRSs on different datums without
// shift parameters are transformed as PROJ does by default for such a pair
// (a "ballpark" transformation: latitude and longitude carried over
// unchanged, each side on its own ellipsoid); a CRS with a real shift
// (+towgs84 other than zeros, +nadgrids, a datum that PROJ defines with a
// grid) is refused unless both sides carry the same one.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "crs.h"

namespace pg {

struct Ellipsoid {
  double a = 0.0, f = 0.0;  // semi-major axis (m), flattening
};

// Ellipsoids by proj4 name (a, 1/f; b for Clarke 1866 as PROJ defines it).
inline bool ellipsoid_by_name(const std::string &name, Ellipsoid &e) {
  struct Row {
    const char *name;
    double a, rf;
  };
  static const Row rows[] = {
      {"WGS84", 6378137.0, 298.257223563},   {"GRS80", 6378137.0, 298.257222101},
      {"intl", 6378388.0, 297.0},            {"clrk80", 6378249.145, 293.4663},
      {"bessel", 6377397.155, 299.1528128},  {"airy", 6377563.396, 299.3249646},
      {"krass", 6378245.0, 298.3},           {"WGS72", 6378135.0, 298.26},
      {"GRS67", 6378160.0, 298.247167427},   {"aust_SA", 6378160.0, 298.25},
  };
  for (const Row &r : rows)
    if (name == r.name) {
      e.a = r.a;
      e.f = 1.0 / r.rf;
      return true;
    }
  if (name == "clrk66") {  // a = 6378206.4, b = 6356583.8
    e.a = 6378206.4;
    e.f = (6378206.4 - 6356583.8) / 6378206.4;
    return true;
  }
  if (name == "sphere") {
    e.a = 6370997.0;
    e.f = 0.0;
    return true;
  }
  return false;
}

struct Crs {
  enum Kind { kGeographic, kTransverseMercator } kind = kGeographic;
  Ellipsoid e;
  double lon0 = 0.0, lat0 = 0.0;  // radians
  double k0 = 1.0, x0 = 0.0, y0 = 0.0;
  double to_meter = 1.0;
  std::string datum_key;  // equal keys: no datum shift between the two CRSs
  bool shift = false;     // the CRS carries a datum shift this header cannot apply
};

inline bool parse_number(const std::string &s, double &v) {
  if (s.empty()) return false;
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == 0 && std::isfinite(v);
}

// Parses a proj4 string.  Returns false with `err` for anything outside the
// supported set (so a caller can say which CRS it cannot transform).
inline bool parse_crs(const std::string &text, Crs &c, std::string &err) {
  const auto t = proj_tokens(text);
  if (t.empty()) {
    err = "empty proj4 string";
    return false;
  }
  const std::string proj = proj_param(t, "proj");
  if (proj != "longlat" && proj != "latlong" && proj != "lonlat" && proj != "latlon" && proj != "utm" &&
      proj != "tmerc") {
    err = "projection '" + proj + "' is not supported (geographic, utm and tmerc are)";
    return false;
  }
  static const char *known[] = {"proj",  "zone",  "south", "ellps", "datum", "a",     "b",        "rf",
                                "f",     "R",     "lat_0", "lon_0", "k",     "k_0",   "x_0",      "y_0",
                                "units", "towgs84", "nadgrids", "to_meter", "approx", "pm", "axis"};
  for (const auto &w : t) {
    const auto eq = w.find('=');
    const std::string key = w.substr(1, eq == std::string::npos ? std::string::npos : eq - 1);
    bool ok = false;
    for (const char *k : known) ok = ok || key == k;
    if (!ok || w[0] != '+') {
      err = "unsupported proj4 parameter '" + w + "'";
      return false;
    }
  }
  // ellipsoid: +datum, then +ellps, then explicit axes
  const std::string datum = proj_param(t, "datum");
  if (!datum.empty()) {
    if (datum == "WGS84") {
      ellipsoid_by_name("WGS84", c.e);
    } else if (datum == "NAD83") {
      ellipsoid_by_name("GRS80", c.e);
    } else {
      err = "datum '" + datum + "' needs a datum shift, which this library does not implement";
      return false;
    }
  }
  const std::string ellps = proj_param(t, "ellps");
  if (!ellps.empty() && !ellipsoid_by_name(ellps, c.e)) {
    err = "unknown ellipsoid '" + ellps + "'";
    return false;
  }
  double v;
  if (parse_number(proj_param(t, "R"), v)) {
    c.e.a = v;
    c.e.f = 0.0;
  }
  if (parse_number(proj_param(t, "a"), v)) {
    c.e.a = v;
    double b, rf, f;
    if (parse_number(proj_param(t, "b"), b)) c.e.f = (v - b) / v;
    else if (parse_number(proj_param(t, "rf"), rf) && rf != 0.0) c.e.f = 1.0 / rf;
    else if (parse_number(proj_param(t, "f"), f)) c.e.f = f;
  }
  if (c.e.a == 0.0) ellipsoid_by_name("GRS80", c.e);  // PROJ's default ellipsoid
  if (!(c.e.a > 0.0) || !(c.e.f >= 0.0 && c.e.f < 0.5)) {
    err = "invalid ellipsoid";
    return false;
  }
  // datum shift
  const std::string towgs = proj_param(t, "towgs84"), grids = proj_param(t, "nadgrids");
  bool zero_shift = true;
  if (!towgs.empty()) {
    size_t at = 0;
    while (at <= towgs.size()) {
      const size_t comma = towgs.find(',', at);
      double x;
      const std::string part = towgs.substr(at, comma == std::string::npos ? std::string::npos : comma - at);
      if (!parse_number(part, x)) {
        err = "invalid +towgs84";
        return false;
      }
      zero_shift = zero_shift && x == 0.0;
      if (comma == std::string::npos) break;
      at = comma + 1;
    }
  }
  c.shift = !zero_shift || (!grids.empty() && grids != "@null");
  c.datum_key = datum + "|" + towgs + "|" + grids;
  const std::string pm = proj_param(t, "pm");
  if (!pm.empty() && pm != "greenwich" && pm != "0") {
    err = "prime meridian '" + pm + "' is not supported";
    return false;
  }
  const std::string axis = proj_param(t, "axis");
  if (!axis.empty() && axis != "enu") {
    err = "axis order '" + axis + "' is not supported";
    return false;
  }
  const double deg = M_PI / 180.0;
  if (proj == "longlat" || proj == "latlong" || proj == "lonlat" || proj == "latlon") {
    c.kind = Crs::kGeographic;
    return true;
  }
  if (proj == "utm") {
    double zone;
    if (!parse_number(proj_param(t, "zone"), zone) || zone < 1 || zone > 60 || zone != std::floor(zone)) {
      err = "+proj=utm needs +zone=1..60";
      return false;
    }
    c.kind = Crs::kTransverseMercator;
    c.lon0 = (-183.0 + 6.0 * zone) * deg;
    c.lat0 = 0.0;
    c.k0 = 0.9996;
    c.x0 = 500000.0;
    c.y0 = proj_flag(t, "south") ? 10000000.0 : 0.0;
  } else if (proj == "tmerc") {
    c.kind = Crs::kTransverseMercator;
    double x;
    if (parse_number(proj_param(t, "lon_0"), x)) c.lon0 = x * deg;
    if (parse_number(proj_param(t, "lat_0"), x)) c.lat0 = x * deg;
    if (parse_number(proj_param(t, "k_0"), x) || parse_number(proj_param(t, "k"), x)) c.k0 = x;
    if (parse_number(proj_param(t, "x_0"), x)) c.x0 = x;
    if (parse_number(proj_param(t, "y_0"), x)) c.y0 = x;
  }
  const std::string units = proj_param(t, "units");
  if (units.empty() || units == "m") c.to_meter = 1.0;
  else if (units == "km") c.to_meter = 1000.0;
  else if (units == "ft") c.to_meter = 0.3048;
  else if (units == "us-ft") c.to_meter = 1200.0 / 3937.0;
  else {
    err = "units '" + units + "' are not supported";
    return false;
  }
  if (parse_number(proj_param(t, "to_meter"), v)) c.to_meter = v;
  return true;
}

// Transverse Mercator on one ellipsoid.
class TransverseMercator {
 public:
  explicit TransverseMercator(const Crs &c) : c_(c) {
    const double f = c.e.f;
    e_ = std::sqrt(f * (2.0 - f));
    const double n = f / (2.0 - f), n2 = n * n, n3 = n2 * n, n4 = n3 * n;
    A_ = c.e.a / (1.0 + n) * (1.0 + n2 / 4.0 + n4 / 64.0);
    alpha_[0] = n / 2.0 - 2.0 * n2 / 3.0 + 5.0 * n3 / 16.0 + 41.0 * n4 / 180.0;
    alpha_[1] = 13.0 * n2 / 48.0 - 3.0 * n3 / 5.0 + 557.0 * n4 / 1440.0;
    alpha_[2] = 61.0 * n3 / 240.0 - 103.0 * n4 / 140.0;
    alpha_[3] = 49561.0 * n4 / 161280.0;
    beta_[0] = n / 2.0 - 2.0 * n2 / 3.0 + 37.0 * n3 / 96.0 - n4 / 360.0;
    beta_[1] = n2 / 48.0 + n3 / 15.0 - 437.0 * n4 / 1440.0;
    beta_[2] = 17.0 * n3 / 480.0 - 37.0 * n4 / 840.0;
    beta_[3] = 4397.0 * n4 / 161280.0;
    double x, y;
    northing0_ = 0.0;
    forward_raw(c.lat0, c.lon0, x, y);  // the origin's northing on the central meridian
    northing0_ = y;
  }

  // (lat, lon) radians -> (x, y) in the CRS's units.  False outside the
  // hemisphere around the central meridian.
  bool forward(double lat, double lon, double &x, double &y) const {
    double l = std::remainder(lon - c_.lon0, 2.0 * M_PI);
    if (!(std::fabs(l) < M_PI / 2.0) || !(std::fabs(lat) <= M_PI / 2.0)) return false;
    double X, Y;
    forward_raw(lat, c_.lon0 + l, X, Y);
    x = (c_.x0 + X) / c_.to_meter;
    y = (c_.y0 + Y - northing0_) / c_.to_meter;
    return std::isfinite(x) && std::isfinite(y);
  }

  bool inverse(double x, double y, double &lat, double &lon) const {
    const double xi = (y * c_.to_meter - c_.y0 + northing0_) / (c_.k0 * A_);
    const double eta = (x * c_.to_meter - c_.x0) / (c_.k0 * A_);
    double xp = xi, ep = eta;
    for (int j = 1; j <= 4; ++j) {
      xp -= beta_[j - 1] * std::sin(2.0 * j * xi) * std::cosh(2.0 * j * eta);
      ep -= beta_[j - 1] * std::cos(2.0 * j * xi) * std::sinh(2.0 * j * eta);
    }
    // conformal latitude and longitude of the sphere (Gauss-Schreiber)
    const double chi = std::asin(std::sin(xp) / std::cosh(ep));
    lon = c_.lon0 + std::atan2(std::sinh(ep), std::cos(xp));
    lat = geodetic_of_conformal(chi);
    return std::isfinite(lat) && std::isfinite(lon);
  }

 private:
  // Conformal latitude chi of geodetic phi (Snyder 1987, eq. 3-1).
  double conformal(double phi) const {
    const double es = e_ * std::sin(phi);
    return 2.0 * std::atan(std::tan(M_PI / 4.0 + phi / 2.0) * std::pow((1.0 - es) / (1.0 + es), e_ / 2.0)) -
           M_PI / 2.0;
  }
  // Its inverse by fixed-point iteration (Snyder 1987, eq. 7-9).
  double geodetic_of_conformal(double chi) const {
    if (std::fabs(chi) >= M_PI / 2.0) return chi;
    const double t = std::tan(M_PI / 4.0 + chi / 2.0);
    double phi = chi;
    for (int it = 0; it < 30; ++it) {
      const double es = e_ * std::sin(phi);
      const double next = 2.0 * std::atan(t * std::pow((1.0 + es) / (1.0 - es), e_ / 2.0)) - M_PI / 2.0;
      if (std::fabs(next - phi) < 1e-15) return next;
      phi = next;
    }
    return phi;
  }
  void forward_raw(double lat, double lon, double &x, double &y) const {
    const double chi = conformal(lat);
    const double l = lon - c_.lon0;
    // Gauss-Schreiber coordinates on the conformal sphere
    const double xp = std::atan2(std::tan(chi), std::cos(l));
    const double ep = std::atanh(std::cos(chi) * std::sin(l));
    double xi = xp, eta = ep;
    for (int j = 1; j <= 4; ++j) {
      xi += alpha_[j - 1] * std::sin(2.0 * j * xp) * std::cosh(2.0 * j * ep);
      eta += alpha_[j - 1] * std::cos(2.0 * j * xp) * std::sinh(2.0 * j * ep);
    }
    x = c_.k0 * A_ * eta;
    y = c_.k0 * A_ * xi;
  }

  Crs c_;
  double e_, A_, alpha_[4], beta_[4], northing0_;
};

// src -> dst for points (x, y in each CRS's axis order: east, north).
class Transform {
 public:
  // Returns false with `err` when a CRS is not supported or the pair needs
  // a datum shift.
  bool init(const std::string &src, const std::string &dst, std::string &err) {
    identity_ = same_crs(src, dst);
    if (identity_) return true;
    std::string e;
    if (!parse_crs(src, s_, e)) {
      err = "source CRS '" + src + "': " + e;
      return false;
    }
    if (!parse_crs(dst, d_, e)) {
      err = "target CRS '" + dst + "': " + e;
      return false;
    }
    if (s_.datum_key != d_.datum_key && (s_.shift || d_.shift)) {
      err = "'" + src + "' -> '" + dst + "' needs a datum shift, which this library does not implement";
      return false;
    }
    if (s_.kind == Crs::kTransverseMercator) stm_.reset(new TransverseMercator(s_));
    if (d_.kind == Crs::kTransverseMercator) dtm_.reset(new TransverseMercator(d_));
    return true;
  }
  bool identity() const { return identity_; }

  // Transforms one point in place; false when it lies outside a
  // projection's domain (the caller drops the geometry).
  bool apply(double &x, double &y) const {
    if (identity_) return true;
    if (std::isnan(x) && std::isnan(y)) return true;  // POINT EMPTY
    const double deg = M_PI / 180.0;
    double lat, lon;
    if (stm_) {
      if (!stm_->inverse(x, y, lat, lon)) return false;
    } else {
      lon = x * deg;
      lat = y * deg;
      if (!(std::fabs(lat) <= M_PI / 2.0) || !std::isfinite(lon)) return false;
    }
    if (dtm_) return dtm_->forward(lat, lon, x, y);
    x = lon / deg;
    y = lat / deg;
    return std::isfinite(x) && std::isfinite(y);
  }

 private:
  struct Holder {
    TransverseMercator *p = nullptr;
    ~Holder() { delete p; }
    void reset(TransverseMercator *q) {
      delete p;
      p = q;
    }
    explicit operator bool() const { return p != nullptr; }
    TransverseMercator *operator->() const { return p; }
  };
  bool identity_ = false;
  Crs s_, d_;
  Holder stm_, dtm_;
};

// Rewrites the coordinates of one WKB geometry (OGC or ISO, 2-D/Z/M/ZM,
// either byte order: the encodings the rasteriser reads) through `tr` into
// `out`, keeping every other byte.  Returns 0 on success, 1 when the bytes
// are not WKB the rasteriser would read, 2 when a vertex lies outside a
// projection's domain.
class WkbTransformer {
 public:
  WkbTransformer(const Transform &tr, const uint8_t *p, size_t n, std::vector<uint8_t> &out)
      : tr_(tr), p_(p), n_(n), out_(out) {}

  int run() {
    out_.assign(p_, p_ + n_);
    return geometry(0);
  }

 private:
  int geometry(int depth) {
    if (depth > 32 || off_ + 5 > n_) return 1;
    const uint8_t order = p_[off_];
    if (order > 1) return 1;
    const bool le = order == 1;
    off_ += 1;
    uint32_t code = u32(le);
    const bool fz = code & 0x80000000u, fm = code & 0x40000000u;
    if (code & 0x30000000u) return 1;
    code &= 0x0FFFFFFFu;
    if ((fz || fm) && code >= 1000) return 1;
    if (code >= 4000) return 1;
    static const int extra[4] = {0, 1, 1, 2};
    const int dims = 2 + (fz ? 1 : 0) + (fm ? 1 : 0) + extra[code / 1000];
    switch (code % 1000) {
      case 1:
        return points(le, dims, 1);
      case 2: {
        uint32_t np;
        if (!count(le, np)) return 1;
        return points(le, dims, np);
      }
      case 3: {
        uint32_t nr;
        if (!count(le, nr)) return 1;
        for (uint32_t r = 0; r < nr; ++r) {
          uint32_t np;
          if (!count(le, np)) return 1;
          const int s = points(le, dims, np);
          if (s) return s;
        }
        return 0;
      }
      case 4:
      case 5:
      case 6:
      case 7: {
        uint32_t ng;
        if (!count(le, ng)) return 1;
        for (uint32_t g = 0; g < ng; ++g) {
          const int s = geometry(depth + 1);
          if (s) return s;
        }
        return 0;
      }
      default:
        return 1;
    }
  }
  uint32_t u32(bool le) {
    uint32_t v;
    std::memcpy(&v, p_ + off_, 4);
    if (!le) v = __builtin_bswap32(v);
    off_ += 4;
    return v;
  }
  bool count(bool le, uint32_t &c) {
    if (off_ + 4 > n_) return false;
    c = u32(le);
    return true;
  }
  double f64_at(size_t at, bool le) const {
    uint64_t v;
    std::memcpy(&v, p_ + at, 8);
    if (!le) v = __builtin_bswap64(v);
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
  void put_f64(size_t at, bool le, double d) {
    uint64_t v;
    std::memcpy(&v, &d, 8);
    if (!le) v = __builtin_bswap64(v);
    std::memcpy(out_.data() + at, &v, 8);
  }
  int points(bool le, int dims, uint32_t np) {
    const size_t stride = 8 * static_cast<size_t>(dims);
    if (np > (n_ - off_) / stride) return 1;
    for (uint32_t i = 0; i < np; ++i) {
      double x = f64_at(off_, le), y = f64_at(off_ + 8, le);
      if (!tr_.apply(x, y)) return 2;
      put_f64(off_, le, x);
      put_f64(off_ + 8, le, y);
      off_ += stride;
    }
    return 0;
  }

  const Transform &tr_;
  const uint8_t *p_;
  size_t n_, off_ = 0;
  std::vector<uint8_t> &out_;
};

}  // namespace pg
