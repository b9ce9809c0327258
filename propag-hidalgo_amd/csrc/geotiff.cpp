// Arrival-time GeoTIFF writer: the file write_times produces with GDAL's
// GTiff driver (crates/propag/src/propag.rs:576-600): one f32 band, tiled
// 256 x 256 (TILED=YES, BLOCKXSIZE=BLOCKYSIZE=256), uncompressed, nodata
// FLT_MAX (GDAL_NODATA tag), the raster's geotransform and CRS.
//
// Written directly from the TIFF 6.0 and GeoTIFF 1.0 layouts: an 8-byte
// header, the tiles in row-major tile order (edge tiles padded with the
// nodata value), then one IFD with its out-of-line values.  The
// geotransform is stored as ModelPixelScale + ModelTiepoint when it is
// north-up without rotation, else as ModelTransformation (as GDAL does).
// Files that would pass 4 GiB are BigTIFF (GDAL's BIGTIFF=IF_NEEDED).
// CRS keys: see crs.h.

#include <cfloat>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "crs.h"
#include "pg_internal.h"

namespace pg {
namespace {

constexpr uint32_t kTile = 256;

enum TiffType : uint16_t { kAscii = 2, kShort = 3, kLong = 4, kDouble = 12, kLong8 = 16 };

struct Entry {
  uint16_t tag, type;
  uint64_t count;
  std::vector<uint8_t> data;  // little-endian values
};

template <class V>
Entry make(uint16_t tag, TiffType type, const std::vector<V> &vals) {
  Entry e{tag, type, vals.size(), {}};
  e.data.resize(vals.size() * sizeof(V));
  std::memcpy(e.data.data(), vals.data(), e.data.size());
  return e;
}

Entry ascii(uint16_t tag, const std::string &s) {
  Entry e{tag, kAscii, s.size() + 1, {}};
  e.data.assign(s.begin(), s.end());
  e.data.push_back(0);
  return e;
}

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put(FILE *f, const void *p, size_t n) { return std::fwrite(p, 1, n, f) == n; }

}  // namespace
}  // namespace pg

extern "C" pg_status pg_write_times_geotiff(const char *path, const GeoReference *geo, const float *time) {
  using namespace pg;
  if (!path || !*path || !geo || (!time && geo->width && geo->height))
    return set_error(PG_ERR_INVALID, "pg_write_times_geotiff: null argument");
  const uint32_t W = geo->width, H = geo->height;
  if (W == 0 || H == 0) return set_error(PG_ERR_INVALID, "pg_write_times_geotiff: empty raster");
  const uint32_t tx = (W + kTile - 1) / kTile, ty = (H + kTile - 1) / kTile;
  const uint64_t ntiles = static_cast<uint64_t>(tx) * ty;
  const uint64_t tile_bytes = static_cast<uint64_t>(kTile) * kTile * 4;
  const uint64_t data_end = 16 + ntiles * tile_bytes;
  const bool big = data_end + ntiles * 16 + (1u << 20) > 0xFFFFFFFFull;

  // -------- tags (sorted by tag number)
  const GT &g = geo->transform.gt;
  const std::string proj = proj_of(geo->proj);
  std::vector<Entry> tags;
  tags.push_back(make<uint32_t>(256, kLong, {W}));
  tags.push_back(make<uint32_t>(257, kLong, {H}));
  tags.push_back(make<uint16_t>(258, kShort, {32}));
  tags.push_back(make<uint16_t>(259, kShort, {1}));   // no compression
  tags.push_back(make<uint16_t>(262, kShort, {1}));   // BlackIsZero
  tags.push_back(make<uint16_t>(277, kShort, {1}));
  tags.push_back(make<uint16_t>(284, kShort, {1}));   // contiguous
  tags.push_back(make<uint32_t>(322, kLong, {kTile}));
  tags.push_back(make<uint32_t>(323, kLong, {kTile}));
  const size_t offsets_at = tags.size();
  if (big) {
    tags.push_back(make<uint64_t>(324, kLong8, std::vector<uint64_t>(ntiles)));
    tags.push_back(make<uint64_t>(325, kLong8, std::vector<uint64_t>(ntiles, tile_bytes)));
  } else {
    tags.push_back(make<uint32_t>(324, kLong, std::vector<uint32_t>(ntiles)));
    tags.push_back(make<uint32_t>(325, kLong, std::vector<uint32_t>(ntiles, static_cast<uint32_t>(tile_bytes))));
  }
  tags.push_back(make<uint16_t>(339, kShort, {3}));   // IEEE float
  if (g.rx == 0.0f && g.ry == 0.0f && g.dy < 0.0f) {
    tags.push_back(make<double>(33550, kDouble, {double(g.dx), -double(g.dy), 0.0}));
    tags.push_back(make<double>(33922, kDouble, {0.0, 0.0, 0.0, double(g.x0), double(g.y0), 0.0}));
  } else {
    tags.push_back(make<double>(34264, kDouble,
                                {double(g.dx), double(g.rx), 0.0, double(g.x0), double(g.ry), double(g.dy), 0.0,
                                 double(g.y0), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0}));
  }
  // GeoKeyDirectory: header, then (key, location, count, value) sorted by key.
  const GeoKeys k = geokeys_of(proj);
  std::string citation = proj + "|";
  std::vector<uint16_t> keys;
  auto key = [&](uint16_t id, uint16_t loc, uint16_t count, uint16_t value) {
    keys.insert(keys.end(), {id, loc, count, value});
  };
  key(1024, 0, 1, static_cast<uint16_t>(k.model));
  key(1025, 0, 1, 1);  // RasterPixelIsArea
  key(1026, 34737, static_cast<uint16_t>(citation.size()), 0);
  if (k.geographic) key(2048, 0, 1, static_cast<uint16_t>(k.geographic));
  if (k.ellipsoid) key(2056, 0, 1, static_cast<uint16_t>(k.ellipsoid));
  if (k.projected) key(3072, 0, 1, static_cast<uint16_t>(k.projected));
  if (k.projection) key(3074, 0, 1, static_cast<uint16_t>(k.projection));
  if (k.linear_units) key(3076, 0, 1, static_cast<uint16_t>(k.linear_units));
  std::vector<uint16_t> dir = {1, 1, 0, static_cast<uint16_t>(keys.size() / 4)};
  dir.insert(dir.end(), keys.begin(), keys.end());
  tags.push_back(make<uint16_t>(34735, kShort, dir));
  tags.push_back(ascii(34737, citation));  // GeoAsciiParams: '|'-terminated strings
  char nod[64];
  std::snprintf(nod, sizeof nod, "%.17g", static_cast<double>(FLT_MAX));
  tags.push_back(ascii(42113, nod));

  // -------- layout: header | tiles | IFD | out-of-line values
  const uint64_t entry_size = big ? 20 : 12, inline_max = big ? 8 : 4;
  const uint64_t ifd_at = data_end;
  const uint64_t ifd_size = (big ? 8 : 2) + tags.size() * entry_size + (big ? 8 : 4);
  uint64_t extra_at = ifd_at + ifd_size;
  std::vector<uint64_t> value_at(tags.size(), 0);
  for (size_t i = 0; i < tags.size(); ++i) {
    if (tags[i].data.size() > inline_max) {
      extra_at = (extra_at + 7) & ~uint64_t{7};
      value_at[i] = extra_at;
      extra_at += tags[i].data.size();
    }
  }
  if (!big && extra_at > 0xFFFFFFFFull) return set_error(PG_ERR_INVALID, "pg_write_times_geotiff: layout overflow");
  {
    Entry &off = tags[offsets_at];
    for (uint64_t t = 0; t < ntiles; ++t) {
      const uint64_t at = 16 + t * tile_bytes;
      if (big) std::memcpy(off.data.data() + 8 * t, &at, 8);
      else {
        const uint32_t a32 = static_cast<uint32_t>(at);
        std::memcpy(off.data.data() + 4 * t, &a32, 4);
      }
    }
  }

  File out;
  out.f = std::fopen(path, "wb");
  if (!out.f) return set_error(PG_ERR_IO, std::string("pg_write_times_geotiff: cannot create ") + path);
  bool ok = true;
  if (big) {
    const uint8_t h[16] = {'I', 'I', 43, 0, 8, 0, 0, 0};
    uint8_t hb[16];
    std::memcpy(hb, h, 8);
    std::memcpy(hb + 8, &ifd_at, 8);
    ok &= put(out.f, hb, 16);
  } else {
    uint8_t hb[16] = {'I', 'I', 42, 0};
    const uint32_t a = static_cast<uint32_t>(ifd_at);
    std::memcpy(hb + 4, &a, 4);
    ok &= put(out.f, hb, 16);  // 8 header bytes + 8 of padding before the tiles
  }
  std::vector<float> tile(static_cast<size_t>(kTile) * kTile);
  for (uint32_t j = 0; j < ty && ok; ++j)
    for (uint32_t i = 0; i < tx && ok; ++i) {
      std::fill(tile.begin(), tile.end(), FLT_MAX);
      const uint32_t x0 = i * kTile, y0 = j * kTile;
      const uint32_t w = std::min(kTile, W - x0), h = std::min(kTile, H - y0);
      for (uint32_t r = 0; r < h; ++r)
        std::memcpy(&tile[static_cast<size_t>(r) * kTile], time + static_cast<size_t>(y0 + r) * W + x0, w * 4);
      ok &= put(out.f, tile.data(), tile.size() * 4);
    }
  // IFD
  std::vector<uint8_t> ifd;
  auto app = [&](const void *p, size_t n) { ifd.insert(ifd.end(), static_cast<const uint8_t *>(p), static_cast<const uint8_t *>(p) + n); };
  if (big) {
    const uint64_t n = tags.size();
    app(&n, 8);
  } else {
    const uint16_t n = static_cast<uint16_t>(tags.size());
    app(&n, 2);
  }
  for (size_t i = 0; i < tags.size(); ++i) {
    const Entry &e = tags[i];
    app(&e.tag, 2);
    app(&e.type, 2);
    if (big) app(&e.count, 8);
    else {
      const uint32_t c = static_cast<uint32_t>(e.count);
      app(&c, 4);
    }
    uint8_t v[8] = {0};
    if (e.data.size() <= inline_max) std::memcpy(v, e.data.data(), e.data.size());
    else if (big) std::memcpy(v, &value_at[i], 8);
    else {
      const uint32_t a = static_cast<uint32_t>(value_at[i]);
      std::memcpy(v, &a, 4);
    }
    app(v, inline_max);
  }
  const uint64_t next = 0;
  app(&next, big ? 8 : 4);
  for (size_t i = 0; i < tags.size(); ++i) {
    if (!value_at[i]) continue;
    ifd.resize(value_at[i] - ifd_at, 0);
    app(tags[i].data.data(), tags[i].data.size());
  }
  ok &= put(out.f, ifd.data(), ifd.size());
  ok &= std::fflush(out.f) == 0;
  if (!ok) return set_error(PG_ERR_IO, std::string("pg_write_times_geotiff: write failed: ") + path);
  return PG_OK;
}
