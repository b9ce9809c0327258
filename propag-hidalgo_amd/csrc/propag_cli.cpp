// propag: the reference's command-line program (crates/propag/src/main.rs)
// on the MI355X path.  The same fixed scenario: a 1024 x 1024 south-up
// raster of 5 m cells in UTM 30 (GRS80), uniform terrain (fuel model 1, all
// moistures 0.1, wind 5 m/s blowing toward azimuth 0, flat), one point
// ignition at time 0 on the corner of cell (500, 500), 10 h; writes the
// arrival times to tiempos.tif and prints the reference's summary lines.
//
//   propag [output.tif]
//
// Steps as propagate() takes them (propag.rs:323-574): rasterise the
// ignition (pg_rasterize_times), build the terrain the loader returns
// (TerrainCuda as a constant TerrainLoader, propag.rs:756-760), initial
// references (pg_init_refs), burn + propagation on the GPU
// (pg_propagate_host), then write_times (pg_write_times_geotiff).

#include <cfloat>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/propag_host.h"

namespace {

int fail(const char *what) {
  std::fprintf(stderr, "propag: %s: %s\n", what, pg_last_error());
  return 1;
}

// WKB of POINT(x y), little-endian.
std::vector<uint8_t> point_wkb(double x, double y) {
  std::vector<uint8_t> b(21);
  b[0] = 1;
  const uint32_t type = 1;
  std::memcpy(&b[1], &type, 4);
  std::memcpy(&b[5], &x, 8);
  std::memcpy(&b[13], &y, 8);
  return b;
}

}  // namespace

int main(int argc, char **argv) {
  const char *out = argc > 1 ? argv[1] : "tiempos.tif";
  const float max_time = 60.0f * 60.0f * 10.0f;
  const float px = 5.0f;
  GeoReference geo;
  if (!GeoReference_south_up(0.0f, 0.0f, px * 1024.0f, px * 1024.0f, px, px,
                             "+proj=utm +zone=30 +ellps=GRS80 +towgs84=0,0,0,0,0,0,0 +units=m +no_defs +type=crs",
                             &geo)) {
    std::fprintf(stderr, "propag: invalid GeoReference\n");
    return 1;
  }
  const Settings settings = create_settings(geo, max_time);
  const uint32_t W = geo.width, H = geo.height;
  const size_t n = static_cast<size_t>(W) * H;

  // fire_pos = backward(width / 2 - 24 / 2, height / 2 - 24 / 2)
  const uint32_t fx = W / 2 - 24 / 2, fy = H / 2 - 24 / 2;
  const GT &g = geo.transform.gt;
  const float wx = g.x0 + static_cast<float>(fx) * g.dx + static_cast<float>(fy) * g.rx;
  const float wy = g.y0 + static_cast<float>(fx) * g.ry + static_cast<float>(fy) * g.dy;
  std::printf("fire_pos=(%u, %u)\n", fx, fy);
  const std::vector<uint8_t> wkb = point_wkb(wx, wy);
  const FFITimeFeature ign{0.0f, wkb.data(), wkb.size()};

  std::vector<float> time(n);
  uint32_t parsed = 0;
  if (pg_rasterize_times(&geo, &ign, 1, time.data(), &parsed) != PG_OK) return fail("rasterize_times");
  std::vector<uint16_t> rx(n), ry(n);
  if (pg_init_refs(&geo, time.data(), rx.data(), ry.data()) != PG_OK) return fail("initial references");

  std::vector<uint8_t> fuel(n, 1);
  std::vector<float> moist(n, 0.1f), wind(n, 5.0f), zero(n, 0.0f);
  std::vector<uint32_t> bits((n + 31) / 32, 0u);
  pg_stats st{};
  if (pg_propagate_host(&settings, fuel.data(), moist.data(), moist.data(), moist.data(), moist.data(),
                        moist.data(), wind.data(), zero.data(), zero.data(), zero.data(), time.data(), rx.data(),
                        ry.data(), bits.data(), nullptr, nullptr, &st) != PG_OK)
    return fail("propagate");
  std::printf("propagation: %llu sweeps, %.3f ms on the device (relaxation %.3f ms)\n",
              static_cast<unsigned long long>(st.iterations), st.elapsed_ms, st.relax_ms);

  float tmax = -FLT_MAX, tmin = FLT_MAX;
  size_t burned = 0;
  for (float t : time)
    if (t < FLT_MAX) {
      ++burned;
      tmax = t > tmax ? t : tmax;
      tmin = t < tmin ? t : tmin;
    }
  std::printf("config_max_time=%g\n", static_cast<double>(settings.max_time));
  if (burned) {
    std::printf("max_time=Some(%.1f)\n", static_cast<double>(tmax));
    std::printf("max_time=Some(%.1f)\n", static_cast<double>(tmin));
  } else {
    std::printf("max_time=None\nmax_time=None\n");
  }
  std::printf("num_times_after=%zu\n", burned);
  std::printf("Generating times geotiff\n");
  if (pg_write_times_geotiff(out, &geo, time.data()) != PG_OK) return fail("write_times");
  return 0;
}
