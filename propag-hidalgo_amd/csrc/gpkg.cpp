// GeoPackage writers for the reference's vector outputs (write_refs,
// crates/propag/src/propag.rs:699-750; write_boundaries, 601-698): one
// feature table per file with an integer `fid` and a `geom` column, its
// gpkg_contents / gpkg_geometry_columns rows and the SRS row, following the
// OGC GeoPackage 1.2 encoding (application_id 'GPKG', user_version 10200,
// geometry blobs = "GP" header, little-endian, [minx maxx miny maxy]
// envelope, then 2-D WKB).
//
// Coordinates are formed as the reference forms them: f32 arithmetic of the
// geotransform (GeoTransform::backward, crates/geometry/src/lib.rs:170-177;
// cell centres add dx/2, dy/2 in f32), printed as the shortest decimal that
// round-trips the f32 (Rust's `{}` in the WKT the reference builds), and
// read back as a double (what GDAL parses from that WKT).
//
// SQLite comes from the system's libsqlite3.so.0, opened at run time; its C
// API is declared here (the library ships no header in this image).  The SRS
// row carries the EPSG code when crs.h recognises one, else the proj4 text
// as an undefined SRS (no WKT without PROJ).

#include <dlfcn.h>
#include <unistd.h>

#include <cctype>
#include <cfloat>
#include <charconv>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "crs.h"
#include "pg_internal.h"

namespace pg {
namespace {

struct sqlite3;
struct sqlite3_stmt;

struct Sqlite {
  int (*open_v2)(const char *, sqlite3 **, int, const char *);
  int (*close)(sqlite3 *);
  int (*exec)(sqlite3 *, const char *, int (*)(void *, int, char **, char **), void *, char **);
  int (*prepare_v2)(sqlite3 *, const char *, int, sqlite3_stmt **, const char **);
  int (*bind_blob)(sqlite3_stmt *, int, const void *, int, void (*)(void *));
  int (*bind_int64)(sqlite3_stmt *, int, long long);
  int (*bind_double)(sqlite3_stmt *, int, double);
  int (*bind_text)(sqlite3_stmt *, int, const char *, int, void (*)(void *));
  int (*step)(sqlite3_stmt *);
  int (*reset)(sqlite3_stmt *);
  int (*finalize)(sqlite3_stmt *);
  const char *(*errmsg)(sqlite3 *);
  void (*free)(void *);
};

constexpr int kOk = 0, kDone = 101;
constexpr int kOpenReadWrite = 0x2, kOpenCreate = 0x4;
void (*const kTransient)(void *) = reinterpret_cast<void (*)(void *)>(-1);

const Sqlite *sqlite(std::string &err) {
  static std::once_flag once;
  static Sqlite api;
  static std::string load_error;
  std::call_once(once, [] {
    void *h = dlopen("libsqlite3.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      load_error = std::string("cannot load libsqlite3.so.0: ") + dlerror();
      return;
    }
    auto get = [&](auto &fn, const char *name) {
      fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(h, name));
      if (!fn && load_error.empty()) load_error = std::string("libsqlite3 lacks ") + name;
    };
    get(api.open_v2, "sqlite3_open_v2");
    get(api.close, "sqlite3_close");
    get(api.exec, "sqlite3_exec");
    get(api.prepare_v2, "sqlite3_prepare_v2");
    get(api.bind_blob, "sqlite3_bind_blob");
    get(api.bind_int64, "sqlite3_bind_int64");
    get(api.bind_double, "sqlite3_bind_double");
    get(api.bind_text, "sqlite3_bind_text");
    get(api.step, "sqlite3_step");
    get(api.reset, "sqlite3_reset");
    get(api.finalize, "sqlite3_finalize");
    get(api.errmsg, "sqlite3_errmsg");
    get(api.free, "sqlite3_free");
  });
  err = load_error;
  return load_error.empty() ? &api : nullptr;
}

// f32 -> shortest round-trip decimal -> double.
double via_decimal(float v) {
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof buf - 1, v);
  *r.ptr = 0;
  return std::strtod(buf, nullptr);
}

struct Vec2 {
  float x, y;
};

// GeoTransform::backward of a cell corner, in f32.
Vec2 backward(const GT &g, uint32_t x, uint32_t y) {
  const float fx = static_cast<float>(x), fy = static_cast<float>(y);
  return {g.x0 + fx * g.dx + fy * g.rx, g.y0 + fx * g.ry + fy * g.dy};
}

struct Blob {
  std::vector<uint8_t> b;
  double minx = DBL_MAX, maxx = -DBL_MAX, miny = DBL_MAX, maxy = -DBL_MAX;
  template <class V>
  void put(V v) {
    const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
    b.insert(b.end(), p, p + sizeof(V));
  }
  // "GP" header with the envelope of pts, then WKB of `kind` (2 line
  // string, 3 polygon of one ring).
  void geometry(int32_t srs_id, uint32_t kind, const std::vector<double> &xy) {
    b.clear();
    double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
    for (size_t i = 0; i < xy.size(); i += 2) {
      x0 = std::min(x0, xy[i]);
      x1 = std::max(x1, xy[i]);
      y0 = std::min(y0, xy[i + 1]);
      y1 = std::max(y1, xy[i + 1]);
    }
    minx = std::min(minx, x0);
    maxx = std::max(maxx, x1);
    miny = std::min(miny, y0);
    maxy = std::max(maxy, y1);
    b.push_back('G');
    b.push_back('P');
    b.push_back(0);                 // version 1
    b.push_back((1u << 1) | 1u);    // envelope [minx, maxx, miny, maxy]; little-endian
    put(srs_id);
    put(x0), put(x1), put(y0), put(y1);
    b.push_back(1);                 // WKB little-endian
    put(kind);
    if (kind == 3) put(uint32_t{1});
    put(static_cast<uint32_t>(xy.size() / 2));
    for (double v : xy) put(v);
  }
};

class Gpkg {
 public:
  ~Gpkg() {
    if (ins_) api_->finalize(ins_);
    if (db_) api_->close(db_);
  }

  pg_status open(const char *path, const std::string &layer, const char *geom_type, const GeoReference *geo) {
    std::string err;
    api_ = sqlite(err);
    if (!api_) return set_error(PG_ERR_IO, "GeoPackage: " + err);
    ::unlink(path);  // the reference creates the dataset anew
    if (api_->open_v2(path, &db_, kOpenReadWrite | kOpenCreate, nullptr) != kOk)
      return fail(std::string("cannot create ") + path);
    layer_ = layer;
    const std::string proj = proj_of(geo->proj);
    const GeoKeys k = geokeys_of(proj);
    std::string org = "NONE";
    srs_ = 100000;
    if (k.projected && k.projected != 32767) {
      srs_ = k.projected;
      org = "EPSG";
    } else if (k.geographic) {
      srs_ = k.geographic;
      org = "EPSG";
    }
    std::string sql =
        "PRAGMA application_id = 1196444487; PRAGMA user_version = 10200; BEGIN;"
        "CREATE TABLE gpkg_spatial_ref_sys (srs_name TEXT NOT NULL, srs_id INTEGER NOT NULL PRIMARY KEY,"
        " organization TEXT NOT NULL, organization_coordsys_id INTEGER NOT NULL, definition TEXT NOT NULL,"
        " description TEXT);"
        "INSERT INTO gpkg_spatial_ref_sys VALUES ('Undefined cartesian SRS', -1, 'NONE', -1, 'undefined',"
        " 'undefined cartesian coordinate reference system');"
        "INSERT INTO gpkg_spatial_ref_sys VALUES ('Undefined geographic SRS', 0, 'NONE', 0, 'undefined',"
        " 'undefined geographic coordinate reference system');"
        "INSERT INTO gpkg_spatial_ref_sys VALUES ('WGS 84 geodetic', 4326, 'EPSG', 4326,"
        " 'GEOGCS[\"WGS 84\",DATUM[\"WGS_1984\",SPHEROID[\"WGS 84\",6378137,298.257223563]],"
        "PRIMEM[\"Greenwich\",0],UNIT[\"degree\",0.0174532925199433],AUTHORITY[\"EPSG\",\"4326\"]]',"
        " 'longitude/latitude coordinates in decimal degrees on the WGS 84 spheroid');"
        "CREATE TABLE gpkg_contents (table_name TEXT NOT NULL PRIMARY KEY, data_type TEXT NOT NULL,"
        " identifier TEXT UNIQUE, description TEXT DEFAULT '', last_change DATETIME NOT NULL DEFAULT"
        " (strftime('%Y-%m-%dT%H:%M:%fZ','now')), min_x DOUBLE, min_y DOUBLE, max_x DOUBLE, max_y DOUBLE,"
        " srs_id INTEGER, CONSTRAINT fk_gc_r_srs_id FOREIGN KEY (srs_id) REFERENCES"
        " gpkg_spatial_ref_sys(srs_id));"
        "CREATE TABLE gpkg_geometry_columns (table_name TEXT NOT NULL, column_name TEXT NOT NULL,"
        " geometry_type_name TEXT NOT NULL, srs_id INTEGER NOT NULL, z TINYINT NOT NULL, m TINYINT NOT NULL,"
        " CONSTRAINT pk_geom_cols PRIMARY KEY (table_name, column_name), CONSTRAINT fk_gc_tn FOREIGN KEY"
        " (table_name) REFERENCES gpkg_contents(table_name), CONSTRAINT fk_gc_srs FOREIGN KEY (srs_id)"
        " REFERENCES gpkg_spatial_ref_sys (srs_id));";
    if (!exec(sql)) return fail("schema");
    if (srs_ != 4326) {
      sqlite3_stmt *st = nullptr;
      if (api_->prepare_v2(db_, "INSERT INTO gpkg_spatial_ref_sys VALUES (?, ?, ?, ?, 'undefined', ?)", -1, &st,
                           nullptr) != kOk)
        return fail("srs");
      api_->bind_text(st, 1, proj.c_str(), -1, kTransient);
      api_->bind_int64(st, 2, srs_);
      api_->bind_text(st, 3, org.c_str(), -1, kTransient);
      api_->bind_int64(st, 4, srs_);
      api_->bind_text(st, 5, proj.c_str(), -1, kTransient);
      const int rc = api_->step(st);
      api_->finalize(st);
      if (rc != kDone) return fail("srs row");
    }
    if (!exec("CREATE TABLE \"" + layer + "\" (fid INTEGER PRIMARY KEY AUTOINCREMENT NOT NULL, geom " +
              geom_type + ");"))
      return fail("feature table");
    geom_type_ = geom_type;
    if (api_->prepare_v2(db_, ("INSERT INTO \"" + layer + "\" (geom) VALUES (?)").c_str(), -1, &ins_, nullptr) !=
        kOk)
      return fail("insert");
    return PG_OK;
  }

  pg_status add(uint32_t kind, const std::vector<double> &xy) {
    blob_.geometry(static_cast<int32_t>(srs_), kind, xy);
    api_->bind_blob(ins_, 1, blob_.b.data(), static_cast<int>(blob_.b.size()), kTransient);
    if (api_->step(ins_) != kDone) return fail("insert feature");
    api_->reset(ins_);
    ++count_;
    return PG_OK;
  }

  pg_status finish() {
    sqlite3_stmt *st = nullptr;
    if (api_->prepare_v2(db_,
                         "INSERT INTO gpkg_contents (table_name, data_type, identifier, description, min_x, min_y,"
                         " max_x, max_y, srs_id) VALUES (?, 'features', ?, '', ?, ?, ?, ?, ?)",
                         -1, &st, nullptr) != kOk)
      return fail("contents");
    api_->bind_text(st, 1, layer_.c_str(), -1, kTransient);
    api_->bind_text(st, 2, layer_.c_str(), -1, kTransient);
    if (count_) {
      api_->bind_double(st, 3, blob_.minx);
      api_->bind_double(st, 4, blob_.miny);
      api_->bind_double(st, 5, blob_.maxx);
      api_->bind_double(st, 6, blob_.maxy);
    }
    api_->bind_int64(st, 7, srs_);
    const int rc = api_->step(st);
    api_->finalize(st);
    if (rc != kDone) return fail("contents row");
    if (!exec("INSERT INTO gpkg_geometry_columns VALUES ('" + layer_ + "', 'geom', '" + geom_type_ + "', " +
              std::to_string(srs_) + ", 0, 0); COMMIT;"))
      return fail("commit");
    return PG_OK;
  }

 private:
  bool exec(const std::string &sql) {
    char *msg = nullptr;
    const int rc = api_->exec(db_, sql.c_str(), nullptr, nullptr, &msg);
    if (msg) api_->free(msg);
    return rc == kOk;
  }
  pg_status fail(const std::string &what) {
    std::string m = "GeoPackage " + what;
    if (db_) m += std::string(": ") + api_->errmsg(db_);
    return set_error(PG_ERR_IO, m);
  }

  const Sqlite *api_ = nullptr;
  sqlite3 *db_ = nullptr;
  sqlite3_stmt *ins_ = nullptr;
  std::string layer_, geom_type_;
  long long srs_ = 0;
  Blob blob_;
  uint64_t count_ = 0;
};

bool plain_name(const char *s) {
  if (!s || !*s) return false;
  for (const char *p = s; *p; ++p)
    if (!(std::isalnum(static_cast<unsigned char>(*p)) || *p == '_')) return false;
  return true;
}

}  // namespace
}  // namespace pg

extern "C" pg_status pg_write_refs_gpkg(const char *path, const GeoReference *geo, const float *time,
                                        const uint16_t *refs_x, const uint16_t *refs_y) {
  using namespace pg;
  if (!path || !*path || !geo || !time || !refs_x || !refs_y)
    return set_error(PG_ERR_INVALID, "pg_write_refs_gpkg: null argument");
  Gpkg g;
  if (pg_status s = g.open(path, "fire_references", "LINESTRING", geo)) return s;
  const GT &t = geo->transform.gt;
  const float hx = t.dx / 2.0f, hy = t.dy / 2.0f;
  std::vector<double> xy(4);
  const uint32_t W = geo->width, H = geo->height;
  for (uint32_t i = 0; i < W; ++i)
    for (uint32_t j = 0; j < H; ++j) {
      const size_t ix = static_cast<size_t>(j) * W + i;
      if (!(time[ix] < FLT_MAX)) continue;
      const Vec2 dst = backward(t, i, j), src = backward(t, refs_x[ix], refs_y[ix]);
      xy[0] = via_decimal(src.x + hx);
      xy[1] = via_decimal(src.y + hy);
      xy[2] = via_decimal(dst.x + hx);
      xy[3] = via_decimal(dst.y + hy);
      if (pg_status s = g.add(2, xy)) return s;
    }
  return g.finish();
}

extern "C" pg_status pg_write_tiles_gpkg(const char *path, const char *layer, const GeoReference *geo,
                                         uint32_t tile_w, uint32_t tile_h) {
  using namespace pg;
  if (!path || !*path || !geo) return set_error(PG_ERR_INVALID, "pg_write_tiles_gpkg: null argument");
  if (!plain_name(layer)) return set_error(PG_ERR_INVALID, "pg_write_tiles_gpkg: layer must be [A-Za-z0-9_]+");
  if (tile_w == 0 || tile_h == 0) return set_error(PG_ERR_INVALID, "pg_write_tiles_gpkg: empty tile");
  Gpkg g;
  if (pg_status s = g.open(path, layer, "POLYGON", geo)) return s;
  const GT &t = geo->transform.gt;
  std::vector<double> xy(10);
  for (uint32_t y = 0; y < geo->height; y += tile_h)
    for (uint32_t x = 0; x < geo->width; x += tile_w) {
      const uint32_t cx[4] = {x, x + tile_w, x + tile_w, x}, cy[4] = {y, y, y + tile_h, y + tile_h};
      for (int c = 0; c < 5; ++c) {
        const Vec2 p = backward(t, cx[c % 4], cy[c % 4]);
        xy[2 * c] = via_decimal(p.x);
        xy[2 * c + 1] = via_decimal(p.y);
      }
      if (pg_status s = g.add(3, xy)) return s;
    }
  return g.finish();
}
