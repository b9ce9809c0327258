// Internal declarations shared by the HIP sources of libpropag_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/propag_hip.h"

namespace pg {

constexpr int kMaxParticles = 5;
constexpr int kNumFuels = 14;  // codes 0..13; 0 (NoFuel) has no particles

// One fuel particle as the burn kernel reads it (firelib.rs:61-77 reduced to
// what Fuel::rx_int_rbqig touches).  mclass: 0 d1hr, 1 d10hr, 2 d100hr,
// 3 herb, 4 wood (Particle::moisture, firelib.rs:699-713).
struct DevParticle {
  int32_t mclass;
  int32_t dead;
  float area_weight;
  float sigma_factor;
  float load;
};

// One fuel model (firelib.rs:305-334 reduced to the runtime fields).
struct DevFuel {
  int32_t np;         // particles, dead first
  int32_t has_live;
  int32_t pad0, pad1;
  float live_area_weight, dead_area_weight, live_rx, dead_rx;
  float fine_dead_factor, live_ext_factor, bulk_density, flux_ratio;
  float slope_k, wind_b, wind_b_inv, wind_e;
  float wind_k, mext, inv_mext, pad2;
  DevParticle p[kMaxParticles];
  // Derivation intermediates kept for inspection (not read on the device).
  float sigma, beta, residence_time, pad3;
};

struct DevCatalog {
  DevFuel fuel[kNumFuels];
};

// Host-side derivation of the f32 catalog (catalog.cpp).
void build_catalog(DevCatalog *out);

// Thread-local error plumbing (capi.cpp).
pg_status set_error(pg_status s, const std::string &msg);
pg_status hip_error(hipError_t e, const char *what);

#define PG_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return ::pg::hip_error(e_, #call); \
  } while (0)

// Per-device resources created once (capi.cpp).
struct DeviceInfo {
  int device;
  int cu_count;
  size_t lds_per_block;
  DevCatalog *catalog;  // device copy
};
pg_status device_info(DeviceInfo **out);

// Frees the propagation's per-device workspaces and cached launch graphs
// (propag.hip).
void release_launch_resources();

// Stage F launcher (burn.hip).
pg_status launch_burn(const DeviceInfo *dev, size_t n, const uint8_t *fuel,
                      const float *d1, const float *d10, const float *d100,
                      const float *herb, const float *wood, const float *ws,
                      const float *waz, const float *slope, const float *aspect,
                      float *speed, float *az, float *ecc, hipStream_t stream);

// Stage F for a scenario on a resident terrain (burn.hip): constant
// moisture (d1, d10, d100, herb, wood), constant wind unless the per-cell
// wind rasters are given.  Writes all n cells.
pg_status launch_burn_weather(const DeviceInfo *dev, size_t n, const uint8_t *fuel, const float *slope,
                              const float *aspect, const float moisture[5], float wind_speed, float wind_azimuth,
                              const float *wind_speed_field, const float *wind_azimuth_field, float *speed,
                              float *az, float *ecc, hipStream_t stream);

// Moisture pre-pass launcher (moisture.hip).
pg_status launch_moisture(const DeviceInfo *dev, size_t n, const uint8_t *fuel, const float *slope,
                          const float *aspect, const pg_day_weather *w, int hour, float *d1hr, float *d10hr,
                          float *d100hr, float *herb, float *wood, hipStream_t stream);

// Stage P driver (propag.hip).
pg_status run_propagate(const DeviceInfo *dev, const Settings *settings,
                        const float *speed, const float *az, const float *ecc,
                        float *time, uint16_t *refs_x, uint16_t *refs_y,
                        uint32_t *boundaries, const pg_tuning *tuning,
                        pg_stats *stats, hipStream_t stream);

}  // namespace pg
