// proj4 CRS strings on the host, without PROJ: comparison (same CRS or not)
// and the GeoTIFF keys a CRS maps to.  The reference hands these strings to
// GDAL (SpatialRef::from_proj4, crates/propag/src/loader.rs:57-65); this
// library covers the parts its outputs need and says so where it stops.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

namespace pg {

// The NUL-padded proj field of a GeoReference as a string (up to the first
// NUL, or all 1024 bytes).
inline std::string proj_of(const uint8_t (&proj)[1024]) {
  size_t n = 0;
  while (n < 1024 && proj[n]) ++n;
  return std::string(reinterpret_cast<const char *>(proj), n);
}

// proj4 parameters as sorted "+key=value" tokens, without the ones that do
// not change the CRS ("+no_defs", "+type=crs", "+wktext").
inline std::vector<std::string> proj_tokens(const std::string &s) {
  std::istringstream in(s);
  std::vector<std::string> t;
  for (std::string w; in >> w;) {
    if (w == "+no_defs" || w == "+type=crs" || w == "+wktext") continue;
    t.push_back(w);
  }
  std::sort(t.begin(), t.end());
  return t;
}

inline bool same_crs(const std::string &a, const std::string &b) { return proj_tokens(a) == proj_tokens(b); }

inline std::string proj_param(const std::vector<std::string> &t, const std::string &key) {
  const std::string k = "+" + key + "=";
  for (const auto &w : t)
    if (w.compare(0, k.size(), k) == 0) return w.substr(k.size());
  return {};
}

inline bool proj_flag(const std::vector<std::string> &t, const std::string &key) {
  return std::find(t.begin(), t.end(), "+" + key) != t.end();
}

// GeoTIFF 1.0 keys for a CRS (a subset of what GDAL's export writes):
//   EPSG:n / +init=epsg:n           -> geographic (4000 <= n < 5000) or projected code n
//   +proj=longlat|latlong, WGS84    -> geographic 4326
//   +proj=utm +zone=z [+south]:
//     datum or ellipsoid WGS84      -> projected 32600 + z / 32700 + z
//     any other ellipsoid           -> user-defined projected CS, projection
//                                      16000 + z / 16100 + z (UTM), metres,
//                                      ellipsoid code when known
//   anything else                   -> user-defined model, citation only
// The proj4 text is always recorded as the citation.
struct GeoKeys {
  int model = 32767;       // GTModelTypeGeoKey: 1 projected, 2 geographic
  int geographic = 0;      // GeographicTypeGeoKey
  int projected = 0;       // ProjectedCSTypeGeoKey
  int projection = 0;      // ProjectionGeoKey
  int ellipsoid = 0;       // GeogEllipsoidGeoKey
  int linear_units = 0;    // ProjLinearUnitsGeoKey
};

inline GeoKeys geokeys_of(const std::string &proj) {
  GeoKeys k;
  const auto t = proj_tokens(proj);
  std::string lower = proj;
  std::transform(lower.begin(), lower.end(), lower.begin(), [](unsigned char c) { return std::tolower(c); });
  int epsg = 0;
  for (const char *pre : {"+init=epsg:", "epsg:"}) {
    const auto at = lower.find(pre);
    if (at != std::string::npos) {
      epsg = std::atoi(lower.c_str() + at + std::string(pre).size());
      break;
    }
  }
  if (epsg > 0) {
    if (epsg >= 4000 && epsg < 5000) {
      k.model = 2;
      k.geographic = epsg;
    } else {
      k.model = 1;
      k.projected = epsg;
    }
    return k;
  }
  const std::string p = proj_param(t, "proj");
  const std::string datum = proj_param(t, "datum"), ellps = proj_param(t, "ellps");
  const bool wgs84 = datum == "WGS84" || (datum.empty() && ellps == "WGS84");
  if ((p == "longlat" || p == "latlong") && wgs84) {
    k.model = 2;
    k.geographic = 4326;
    return k;
  }
  if (p == "utm") {
    const int zone = std::atoi(proj_param(t, "zone").c_str());
    if (zone >= 1 && zone <= 60) {
      const bool south = proj_flag(t, "south");
      k.model = 1;
      if (wgs84) {
        k.projected = (south ? 32700 : 32600) + zone;
      } else {
        k.projected = 32767;
        k.projection = (south ? 16100 : 16000) + zone;
        k.linear_units = 9001;
        if (ellps == "GRS80") k.ellipsoid = 7019;
        else if (ellps == "intl") k.ellipsoid = 7022;
        else if (ellps == "clrk66") k.ellipsoid = 7008;
      }
    }
  }
  return k;
}

}  // namespace pg
