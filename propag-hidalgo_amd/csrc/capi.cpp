// C ABI of libpropag_hip.so (declared in include/propag_hip.h).
//
// Error convention: every entry point returns a pg_status; on failure the
// message is kept per thread for pg_last_error().  This replaces the
// reference's bool + strncpy into a caller buffer
// (crates/propag/src/propag.rs:274-309), which can leave the buffer
// unterminated.
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pg_internal.h"

namespace pg {
namespace {

thread_local std::string t_last_error;

std::mutex g_dev_mutex;
std::vector<DeviceInfo *> g_devices;

template <class T>
struct DevBuf {  // RAII device allocation for the host-buffer entry point
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T) + 16); }
};

}  // namespace

pg_status set_error(pg_status s, const std::string &msg) {
  t_last_error = msg;
  return s;
}

pg_status hip_error(hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return set_error(e == hipErrorOutOfMemory ? PG_ERR_OUT_OF_MEMORY : PG_ERR_HIP, msg);
}

void release_all() {
  release_launch_resources();
  std::lock_guard<std::mutex> lk(g_dev_mutex);
  int current = 0;
  const bool have = hipGetDevice(&current) == hipSuccess;
  for (size_t d = 0; d < g_devices.size(); ++d) {
    DeviceInfo *info = g_devices[d];
    if (!info) continue;
    if (hipSetDevice(static_cast<int>(d)) == hipSuccess && info->catalog) (void)hipFree(info->catalog);
    delete info;
    g_devices[d] = nullptr;
  }
  if (have) (void)hipSetDevice(current);
  (void)hipGetLastError();
}

// Registered once, when the library first sets up a device: by then the HIP
// runtime (and any tool wrapping it) is loaded and has registered its own
// exit handlers, so this one runs before them and the library's device
// memory, graphs and pinned blocks are released while the runtime is whole.
void release_at_exit() { release_all(); }

pg_status device_info(DeviceInfo **out) {
  int dev = 0;
  PG_HIP(hipGetDevice(&dev));
  static const bool registered = std::atexit(release_at_exit) == 0;
  (void)registered;
  std::lock_guard<std::mutex> lk(g_dev_mutex);
  if (g_devices.size() <= static_cast<size_t>(dev)) g_devices.resize(dev + 1, nullptr);
  if (!g_devices[dev]) {
    hipDeviceProp_t prop;
    PG_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return set_error(PG_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                             "; this library is built for gfx950 only");
    auto *info = new DeviceInfo{};
    info->device = dev;
    info->cu_count = prop.multiProcessorCount;
    int lds = 0;
    PG_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    info->lds_per_block = static_cast<size_t>(lds);
    DevCatalog host{};
    build_catalog(&host);
    PG_HIP(hipMalloc(reinterpret_cast<void **>(&info->catalog), sizeof(DevCatalog)));
    PG_HIP(hipMemcpy(info->catalog, &host, sizeof(DevCatalog), hipMemcpyHostToDevice));
    g_devices[dev] = info;
  }
  *out = g_devices[dev];
  return PG_OK;
}

}  // namespace pg

using namespace pg;

extern "C" {

const char *pg_last_error(void) { return t_last_error.c_str(); }

const char *pg_version(void) { return "propag-hip 0.2.0 (gfx950)"; }

pg_status pg_shutdown(void) {
  release_all();
  return PG_OK;
}

int pg_catalog_f32(int code, float out[16]) {
  if (code < 0 || code >= kNumFuels || !out) return -1;
  DevCatalog cat{};
  build_catalog(&cat);
  const DevFuel &f = cat.fuel[code];
  const float v[16] = {f.sigma, f.beta, f.flux_ratio, f.slope_k, f.wind_b, f.wind_e,
                       f.wind_k, f.live_rx, f.dead_rx, f.live_area_weight, f.dead_area_weight,
                       f.fine_dead_factor, f.live_ext_factor, f.bulk_density,
                       f.residence_time, f.mext};
  std::memcpy(out, v, sizeof(v));
  return f.np;
}

pg_status pg_burn(size_t n, const uint8_t *fuel, const float *d1hr, const float *d10hr,
                  const float *d100hr, const float *herb, const float *wood,
                  const float *wind_speed, const float *wind_azimuth, const float *slope,
                  const float *aspect, float *speed_max, float *azimuth_max,
                  float *eccentricity, void *stream) {
  if (n == 0) return PG_OK;
  if (!fuel || !d1hr || !d10hr || !d100hr || !herb || !wood || !wind_speed || !wind_azimuth ||
      !slope || !aspect || !speed_max || !azimuth_max || !eccentricity)
    return set_error(PG_ERR_INVALID, "pg_burn: null pointer");
  DeviceInfo *dev = nullptr;
  pg_status s = device_info(&dev);
  if (s != PG_OK) return s;
  return launch_burn(dev, n, fuel, d1hr, d10hr, d100hr, herb, wood, wind_speed, wind_azimuth, slope,
                     aspect, speed_max, azimuth_max, eccentricity, static_cast<hipStream_t>(stream));
}

pg_status pg_moisture(size_t n, const uint8_t *fuel, const float *slope, const float *aspect,
                      const pg_day_weather *weather, int hour, float *d1hr, float *d10hr, float *d100hr,
                      float *herb, float *wood, void *stream) {
  if (n == 0) return PG_OK;
  if (!fuel || !slope || !aspect || !weather || !d1hr || !d10hr || !d100hr || !herb || !wood)
    return set_error(PG_ERR_INVALID, "pg_moisture: null pointer");
  if (hour < 0 || hour > 23) return set_error(PG_ERR_INVALID, "pg_moisture: hour must be 0..23");
  DeviceInfo *dev = nullptr;
  pg_status s = device_info(&dev);
  if (s != PG_OK) return s;
  return launch_moisture(dev, n, fuel, slope, aspect, weather, hour, d1hr, d10hr, d100hr, herb, wood,
                         static_cast<hipStream_t>(stream));
}

pg_status pg_propagate(const Settings *settings, const float *speed_max, const float *azimuth_max,
                       const float *eccentricity, float *time, uint16_t *refs_x, uint16_t *refs_y,
                       uint32_t *boundaries, const pg_tuning *tuning, pg_stats *stats,
                       void *stream) {
  DeviceInfo *dev = nullptr;
  pg_status s = device_info(&dev);
  if (s != PG_OK) return s;
  return run_propagate(dev, settings, speed_max, azimuth_max, eccentricity, time, refs_x, refs_y,
                       boundaries, tuning, stats, static_cast<hipStream_t>(stream));
}

pg_status pg_propagate_host(const Settings *settings, const uint8_t *fuel, const float *d1hr,
                            const float *d10hr, const float *d100hr, const float *herb,
                            const float *wood, const float *wind_speed, const float *wind_azimuth,
                            const float *slope, const float *aspect, float *time,
                            uint16_t *refs_x, uint16_t *refs_y, uint32_t *boundaries,
                            float *fire_out, const pg_tuning *tuning, pg_stats *stats) {
  if (!settings) return set_error(PG_ERR_INVALID, "settings is NULL");
  if (!fuel || !d1hr || !d10hr || !d100hr || !herb || !wood || !wind_speed || !wind_azimuth ||
      !slope || !aspect || !time || !refs_x || !refs_y || !boundaries)
    return set_error(PG_ERR_INVALID, "pg_propagate_host: null pointer");
  DeviceInfo *dev = nullptr;
  pg_status s = device_info(&dev);
  if (s != PG_OK) return s;
  const size_t n = static_cast<size_t>(settings->geo_ref.width) * settings->geo_ref.height;
  if (n == 0) return set_error(PG_ERR_INVALID, "raster has no cells");
  const size_t nwords = (n + 31) / 32;
  DevBuf<uint8_t> d_fuel;
  DevBuf<float> d_cols[9], d_fire[3], d_time;
  DevBuf<uint16_t> d_rx, d_ry;
  DevBuf<uint32_t> d_bits;
  PG_HIP(d_fuel.alloc(n));
  for (auto &c : d_cols) PG_HIP(c.alloc(n));
  for (auto &c : d_fire) PG_HIP(c.alloc(n));
  PG_HIP(d_time.alloc(n));
  PG_HIP(d_rx.alloc(n));
  PG_HIP(d_ry.alloc(n));
  PG_HIP(d_bits.alloc(nwords));
  const float *cols[9] = {d1hr, d10hr, d100hr, herb, wood, wind_speed, wind_azimuth, slope, aspect};
  hipStream_t stream = nullptr;
  PG_HIP(hipMemcpyAsync(d_fuel.p, fuel, n, hipMemcpyHostToDevice, stream));
  for (int i = 0; i < 9; ++i)
    PG_HIP(hipMemcpyAsync(d_cols[i].p, cols[i], n * sizeof(float), hipMemcpyHostToDevice, stream));
  for (auto &c : d_fire) PG_HIP(hipMemsetAsync(c.p, 0, n * sizeof(float), stream));  // propag.rs:437-439
  PG_HIP(hipMemcpyAsync(d_time.p, time, n * sizeof(float), hipMemcpyHostToDevice, stream));
  PG_HIP(hipMemcpyAsync(d_rx.p, refs_x, n * sizeof(uint16_t), hipMemcpyHostToDevice, stream));
  PG_HIP(hipMemcpyAsync(d_ry.p, refs_y, n * sizeof(uint16_t), hipMemcpyHostToDevice, stream));
  s = launch_burn(dev, n, d_fuel.p, d_cols[0].p, d_cols[1].p, d_cols[2].p, d_cols[3].p, d_cols[4].p,
                  d_cols[5].p, d_cols[6].p, d_cols[7].p, d_cols[8].p, d_fire[0].p, d_fire[1].p,
                  d_fire[2].p, stream);
  if (s != PG_OK) return s;
  s = run_propagate(dev, settings, d_fire[0].p, d_fire[1].p, d_fire[2].p, d_time.p, d_rx.p, d_ry.p,
                    d_bits.p, tuning, stats, stream);
  if (s != PG_OK && s != PG_ERR_NOT_CONVERGED) return s;
  PG_HIP(hipMemcpyAsync(time, d_time.p, n * sizeof(float), hipMemcpyDeviceToHost, stream));
  PG_HIP(hipMemcpyAsync(refs_x, d_rx.p, n * sizeof(uint16_t), hipMemcpyDeviceToHost, stream));
  PG_HIP(hipMemcpyAsync(refs_y, d_ry.p, n * sizeof(uint16_t), hipMemcpyDeviceToHost, stream));
  PG_HIP(hipMemcpyAsync(boundaries, d_bits.p, nwords * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  if (fire_out)
    for (int i = 0; i < 3; ++i)
      PG_HIP(hipMemcpyAsync(fire_out + i * n, d_fire[i].p, n * sizeof(float), hipMemcpyDeviceToHost, stream));
  PG_HIP(hipStreamSynchronize(stream));
  return s;
}

}  // extern "C"
