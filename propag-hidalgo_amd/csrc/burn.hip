// Stage F: per-cell Rothermel/BEHAVE burn (firelib) on gfx950.
//
// Replaces the Rust-CUDA kernel cuda_standard_simple_burn
// (crates/firelib/src/cuda.rs:57-95) which evaluates Catalog::burn_simple
// (crates/firelib/src/firelib.rs:394-397,1016-1188) per cell.
//
// Layout: SoA rasters (1 B fuel + 9 f32 inputs, 3 f32 outputs) streamed once,
// 49 B/cell; one thread per cell.  The 14-fuel catalog (2.5 KB) is staged in
// LDS once per workgroup because lanes index it by their own fuel code.
// Arithmetic is f32 in the reference's operation order, compiled without FMA
// contraction; the per-cell arithmetic, not HBM, bounds the pass.
#include <cfloat>

#include "firelib_dev.h"

namespace pg {
namespace {

using firelib::Fire;
using firelib::burn_cell;
using firelib::stage_catalog;

struct BurnArgs {
  const uint8_t *fuel;
  const float *d1, *d10, *d100, *herb, *wood, *ws, *waz, *slope, *aspect;
  float *speed, *az, *ecc;
};

// One cell per thread (coalesced 4-byte accesses): at 72 VGPRs seven waves
// per SIMD hide the per-cell arithmetic's latency better than four cells per
// thread with 16-byte accesses did (18.3 against 20.0 us per 1024^2 pass).
__global__ __launch_bounds__(256) void burn_scalar(size_t n, BurnArgs a, const DevCatalog *cat) {
  __shared__ DevFuel s_cat[kNumFuels];
  stage_catalog(s_cat, cat);
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    Fire f;
    if (burn_cell(s_cat, a.fuel[i], a.d1[i], a.d10[i], a.d100[i], a.herb[i], a.wood[i], a.ws[i],
                  a.waz[i], a.slope[i], a.aspect[i], f)) {
      a.speed[i] = f.speed;
      a.az[i] = f.az;
      a.ecc[i] = f.ecc;
    }
  }
}

// A scenario on a resident terrain: fuel, slope and aspect per cell,
// spatially constant moisture and either a constant wind or per-cell wind
// rasters (null pointers select the constants).  Writes every cell: cells
// without a catalog fuel get a null fire (speed 0), as the caller's
// pre-zeroed outputs would hold after cuda_standard_simple_burn.
struct WeatherArgs {
  const uint8_t *fuel;
  const float *slope, *aspect;
  const float *ws_f, *waz_f;
  float d1, d10, d100, herb, wood, ws, waz;
  float *speed, *az, *ecc;
};

__global__ __launch_bounds__(256) void burn_weather(size_t n, WeatherArgs a, const DevCatalog *cat) {
  __shared__ DevFuel s_cat[kNumFuels];
  stage_catalog(s_cat, cat);
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float ws = a.ws_f ? a.ws_f[i] : a.ws;
    const float waz = a.waz_f ? a.waz_f[i] : a.waz;
    Fire f{0.0f, 0.0f, 0.0f};
    (void)burn_cell(s_cat, a.fuel[i], a.d1, a.d10, a.d100, a.herb, a.wood, ws, waz, a.slope[i], a.aspect[i], f);
    a.speed[i] = f.speed;
    a.az[i] = f.az;
    a.ecc[i] = f.ecc;
  }
}

}  // namespace

pg_status launch_burn(const DeviceInfo *dev, size_t n, const uint8_t *fuel, const float *d1,
                      const float *d10, const float *d100, const float *herb, const float *wood,
                      const float *ws, const float *waz, const float *slope, const float *aspect,
                      float *speed, float *az, float *ecc, hipStream_t stream) {
  if (n == 0) return PG_OK;
  BurnArgs a{fuel, d1, d10, d100, herb, wood, ws, waz, slope, aspect, speed, az, ecc};
  // One thread per cell up to 16 workgroups per CU; the grid-stride loop
  // takes the rest.
  size_t blocks = (n + 255) / 256;
  const size_t cap = static_cast<size_t>(dev->cu_count) * 16;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(burn_scalar, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, n, a, dev->catalog);
  PG_HIP(hipGetLastError());
  return PG_OK;
}

pg_status launch_burn_weather(const DeviceInfo *dev, size_t n, const uint8_t *fuel, const float *slope,
                              const float *aspect, const float moisture[5], float wind_speed, float wind_azimuth,
                              const float *wind_speed_field, const float *wind_azimuth_field, float *speed,
                              float *az, float *ecc, hipStream_t stream) {
  if (n == 0) return PG_OK;
  WeatherArgs a{fuel, slope, aspect, wind_speed_field, wind_azimuth_field,
                moisture[0], moisture[1], moisture[2], moisture[3], moisture[4], wind_speed, wind_azimuth,
                speed, az, ecc};
  size_t blocks = (n + 255) / 256;
  const size_t cap = static_cast<size_t>(dev->cu_count) * 16;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(burn_weather, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, n, a, dev->catalog);
  PG_HIP(hipGetLastError());
  return PG_OK;
}

}  // namespace pg
