// libpropag.so: the reference's plugin ABI (include/propag_host.h) over the
// MI355X path of libpropag_hip.so.
//
// FFIPropagation_run follows FFIPropagation_run + propagate() of
// crates/propag/src/propag.rs:274-309, 323-574: validate the request
// (TryFrom<FFIPropagation>, propag.rs:196-235), rasterise the ignitions
// (rasterize_times, 837-890), load the terrain through the caller's
// callback (TerrainLoader for FFITerrainLoader, 813-835), set the initial
// references (433-449), run the burn pass and the propagation on the GPU
// (pg_propagate_host), and write the GeoTIFF and GeoPackage outputs
// (write_times / write_refs / write_boundaries, 576-750).  Errors come back
// as false plus a NUL-terminated message, never a crash (the reference
// catches panics the same way, 274-309).
//
// Where it differs on purpose: no MPI initialisation, so a process may run
// any number of propagations (the reference's second call fails); features
// are reprojected to the raster's CRS by reproject.h instead of PROJ
// (geographic and transverse Mercator CRSs, no datum shifts; other CRSs are
// refused with an error); the boundary layers describe this
// implementation's launch decomposition (8 x 8 tiles in one launch over the
// whole raster) instead of CUDA blocks and super-grid tiles.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/propag_host.h"
#include "crs.h"
#include "reproject.h"

namespace {

void copy_error(const std::string &msg, uintptr_t len, char *out) {
  if (!out || len == 0) return;
  const size_t n = std::min<size_t>(msg.size(), len - 1);
  std::memcpy(out, msg.data(), n);
  out[n] = 0;
}

std::string last(const char *what) { return std::string(what) + ": " + pg_last_error(); }

// The reference's some_path: NULL or "" is no path.
const char *some_path(const char *p) { return p && *p ? p : nullptr; }

// SpatialRef::from_proj4 of the features' CRS and of the raster's, then
// Geometry::transform_to_inplace of each feature (propag.rs:850-858,
// 960-970): features whose WKB parses are rewritten into `bufs` with their
// vertices in the raster's CRS; one with a vertex outside the target
// projection's domain is dropped, as the reference's filter_map drops a
// failed transformation; bytes that are not WKB are passed on for the
// rasteriser to skip.  With the raster's own CRS the features pass
// unchanged.
template <class Feature>
bool reproject(const char *crs, const GeoReference &geo, const Feature *in, size_t n, std::vector<Feature> &out,
               std::vector<std::vector<uint8_t>> &bufs, std::string &err) {
  if (!crs || !*crs) {
    err = "invalid CRS: empty proj4 string";
    return false;
  }
  const std::string raster = pg::proj_of(geo.proj);
  if (raster.empty()) {
    err = "invalid CRS: the GeoReference has no proj4 string";
    return false;
  }
  pg::Transform tr;
  if (!tr.init(crs, raster, err)) return false;
  out.assign(in, in + n);
  if (tr.identity()) return true;
  bufs.assign(n, {});
  size_t kept = 0;
  for (size_t i = 0; i < n; ++i) {
    Feature f = in[i];
    if (f.geom_wkb && f.geom_wkb_len) {
      const int st = pg::WkbTransformer(tr, f.geom_wkb, f.geom_wkb_len, bufs[i]).run();
      if (st == 2) continue;  // outside the target projection: dropped
      if (st == 0) f.geom_wkb = bufs[i].data();
    }
    out[kept++] = f;
  }
  out.resize(kept);
  return true;
}

bool run(const FFIPropagation &p, std::string &err) {
  const GeoReference &geo = p.settings.geo_ref;
  std::vector<FFITimeFeature> features;
  std::vector<std::vector<uint8_t>> wkbs;
  const size_t nin = p.initial_ignited_elements ? p.initial_ignited_elements_len : 0;
  if (!reproject(p.initial_ignited_elements_crs, geo, p.initial_ignited_elements, nin, features, wkbs, err))
    return false;
  const char *out_path = some_path(p.output_path);
  if (!out_path) {
    err = "invalid output_path";
    return false;
  }
  const size_t n = static_cast<size_t>(geo.width) * geo.height;
  std::vector<float> time(n);
  if (pg_rasterize_times(&geo, features.data(), features.size(), time.data(), nullptr) != PG_OK) {
    err = last("rasterize_times");
    return false;
  }
  // TerrainCuda::NULL columns, lent to the loader (propag.rs:813-835).
  std::vector<uint8_t> fuel(n, 0);
  std::vector<float> cols[9];
  for (auto &c : cols) c.assign(n, 0.0f);
  FFITerrain t{fuel.data(), cols[0].data(), cols[1].data(), cols[2].data(), cols[3].data(),
               cols[4].data(), cols[5].data(), cols[6].data(), cols[7].data(), cols[8].data()};
  if (!p.terrain_loader.load || !p.terrain_loader.load(p.terrain_loader.context, &geo, &t)) {
    err = "load_extent_error";
    return false;
  }
  std::vector<uint16_t> rx(n), ry(n);
  std::vector<uint32_t> bits((n + 31) / 32, 0u);
  if (pg_init_refs(&geo, time.data(), rx.data(), ry.data()) != PG_OK) {
    err = last("initial references");
    return false;
  }
  pg_stats stats{};
  if (pg_propagate_host(&p.settings, fuel.data(), cols[0].data(), cols[1].data(), cols[2].data(),
                        cols[3].data(), cols[4].data(), cols[5].data(), cols[6].data(), cols[7].data(),
                        cols[8].data(), time.data(), rx.data(), ry.data(), bits.data(), nullptr, nullptr,
                        &stats) != PG_OK) {
    err = last("propagate");
    return false;
  }
  if (pg_write_times_geotiff(out_path, &geo, time.data()) != PG_OK) {
    err = last("write_times");
    return false;
  }
  if (const char *path = some_path(p.refs_output_path)) {
    if (pg_write_refs_gpkg(path, &geo, time.data(), rx.data(), ry.data()) != PG_OK) {
      err = last("write_refs");
      return false;
    }
  }
  if (const char *path = some_path(p.grid_boundaries_out_path)) {
    if (pg_write_tiles_gpkg(path, "grids", &geo, geo.width, geo.height) != PG_OK) {
      err = last("write_boundaries");
      return false;
    }
  }
  if (const char *path = some_path(p.block_boundaries_out_path)) {
    if (pg_write_tiles_gpkg(path, "blocks", &geo, 8, 8) != PG_OK) {
      err = last("write_boundaries");
      return false;
    }
  }
  return true;
}

// GT::invert in f32 (crates/geometry/src/lib.rs:83-125).
bool invert(const GT &g, GT &out) {
  if (g.rx == 0.0f && g.ry == 0.0f && g.dx != 0.0f && g.dy != 0.0f) {
    out = GT{-g.x0 / g.dx, 1.0f / g.dx, 0.0f, -g.y0 / g.dy, 0.0f, 1.0f / g.dy};
    return true;
  }
  const float det = g.dx * g.dy - g.rx * g.ry;
  const float mag = std::max(std::max(std::fabs(g.dx), std::fabs(g.rx)), std::max(std::fabs(g.ry), std::fabs(g.dy)));
  if (std::fabs(det) <= 1e-10f * mag * mag) return false;
  const float r = 1.0f / det;
  out = GT{(g.rx * g.y0 - g.x0 * g.dy) * r, g.dy * r, -g.rx * r, (-g.dx * g.y0 + g.x0 * g.ry) * r, -g.ry * r, g.dx * r};
  return true;
}

// NumCast f32 -> u32 of a ceil()ed size: None below 0, at 2^32 or NaN.
bool to_u32(float v, uint32_t &out) {
  if (!(v > -1.0f && v < 4294967296.0f)) return false;
  out = static_cast<uint32_t>(v);
  return true;
}

bool make_geo(float ox, float oy, float dx, float dy, float w, float h, const char *proj, GeoReference *res) {
  if (!proj || !res) return false;
  // to_str(): the proj string must be UTF-8 (checked as well-formed sequences)
  for (const unsigned char *s = reinterpret_cast<const unsigned char *>(proj); *s;) {
    int extra = *s < 0x80 ? 0 : (*s >> 5) == 6 ? 1 : (*s >> 4) == 14 ? 2 : (*s >> 3) == 30 ? 3 : -1;
    if (extra < 0) return false;
    ++s;
    for (int i = 0; i < extra; ++i, ++s)
      if ((*s >> 6) != 2) return false;
  }
  GeoReference g{};
  g.transform.gt = GT{ox, dx, 0.0f, oy, 0.0f, dy};
  if (!invert(g.transform.gt, g.transform.inv)) return false;
  if (!to_u32(std::ceil(w), g.width) || !to_u32(std::ceil(h), g.height)) return false;
  const size_t n = std::min<size_t>(std::strlen(proj), sizeof g.proj);
  std::memcpy(g.proj, proj, n);
  *res = g;
  return true;
}

}  // namespace

extern "C" {

bool FFIPropagation_run(FFIPropagation propag, uintptr_t err_len, char *err_msg) {
  std::string err;
  try {
    if (run(propag, err)) return true;
  } catch (const std::exception &e) {
    err = e.what();
  } catch (...) {
    err = "unknown exception";
  }
  copy_error(err, err_len, err_msg);
  return false;
}

bool propag_rasterize_fuels(const FFIFuelFeature *fuels, uintptr_t fuels_len, const char *fuels_crs,
                            const GeoReference *geo_ref, uint8_t *result, char *err_msg, uintptr_t err_len) {
  std::string err;
  try {
    if (!geo_ref || !result) {
      err = "null GeoReference or result";
    } else {
      std::vector<FFIFuelFeature> feats;
      std::vector<std::vector<uint8_t>> wkbs;
      const size_t n = fuels && fuels_len ? fuels_len : 0;
      if (reproject(fuels_crs, *geo_ref, fuels, n, feats, wkbs, err)) {
        if (pg_rasterize_fuels(geo_ref, feats.data(), feats.size(), result, nullptr) == PG_OK) return true;
        err = last("rasterize_fuels");
      }
    }
  } catch (const std::exception &e) {
    err = e.what();
  }
  copy_error(err, err_len, err_msg);
  return false;
}

Settings create_settings(GeoReference geo_ref, float max_time) { return Settings{geo_ref, max_time}; }

FireSimpleCuda create_fire(T speed_max, T azimuth_max, T eccentricity) {
  return FireSimpleCuda{speed_max, azimuth_max, eccentricity};
}

PluginT create_plugin(void) { return PluginT{FFIPropagation_run, propag_rasterize_fuels}; }

uint32_t GeoReference_len(const GeoReference *self) { return self->width * self->height; }

float GeoReference_x0(const GeoReference *self) {
  const GT &g = self->transform.gt;
  return g.dx > 0.0f ? g.x0 : g.x0 + g.dx * static_cast<float>(self->width);
}
float GeoReference_x1(const GeoReference *self) {
  const GT &g = self->transform.gt;
  return g.dx > 0.0f ? g.x0 + g.dx * static_cast<float>(self->width) : g.x0;
}
float GeoReference_y0(const GeoReference *self) {
  const GT &g = self->transform.gt;
  return g.dy > 0.0f ? g.y0 : g.y0 + g.dy * static_cast<float>(self->height);
}
float GeoReference_y1(const GeoReference *self) {
  const GT &g = self->transform.gt;
  return g.dy > 0.0f ? g.y0 + g.dy * static_cast<float>(self->height) : g.y0;
}

bool GeoReference_south_up(float x0, float y0, float x1, float y1, float dx, float dy, const char *proj,
                           GeoReference *result) {
  return make_geo(x0, y0, dx, dy, (x1 - x0) / dx, (y1 - y0) / dy, proj, result);
}

bool GeoReference_north_up(float x0, float y0, float x1, float y1, float dx, float dy, const char *proj,
                           GeoReference *result) {
  return make_geo(x0, y1, dx, -dy, (x1 - x0) / dx, (y1 - y0) / dy, proj, result);
}

}  // extern "C"
