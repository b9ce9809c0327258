// Stage F: per-cell Rothermel/BEHAVE burn (firelib) on gfx950.
//
// Replaces the Rust-CUDA kernel cuda_standard_simple_burn
// (crates/firelib/src/cuda.rs:57-95) which evaluates Catalog::burn_simple
// (crates/firelib/src/firelib.rs:394-397,1016-1188) per cell.
//
// Layout: SoA rasters (1 B fuel + 9 f32 inputs, 3 f32 outputs) streamed once,
// so the kernel is HBM-bound at 49 B/cell.  Each thread owns 4 consecutive
// cells and moves them with 16-byte loads/stores when the rasters are 16-byte
// aligned; the 14-fuel catalog (2.5 KB) is staged in LDS once per workgroup
// because lanes index it by their own fuel code.  Arithmetic is f32 in the
// reference's operation order, compiled without FMA contraction.
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

// 4 cells per thread with 16-byte accesses; requires n % 4 == 0 and 16-byte
// aligned float rasters / 4-byte aligned fuel raster (checked by the host).
__global__ __launch_bounds__(256) void burn_vec4(size_t n4, BurnArgs a, const DevCatalog *cat) {
  __shared__ DevFuel s_cat[kNumFuels];
  stage_catalog(s_cat, cat);
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < n4; q += stride) {
    const uint32_t codes = reinterpret_cast<const uint32_t *>(a.fuel)[q];
    const float4 d1 = reinterpret_cast<const float4 *>(a.d1)[q];
    const float4 d10 = reinterpret_cast<const float4 *>(a.d10)[q];
    const float4 d100 = reinterpret_cast<const float4 *>(a.d100)[q];
    const float4 herb = reinterpret_cast<const float4 *>(a.herb)[q];
    const float4 wood = reinterpret_cast<const float4 *>(a.wood)[q];
    const float4 ws = reinterpret_cast<const float4 *>(a.ws)[q];
    const float4 waz = reinterpret_cast<const float4 *>(a.waz)[q];
    const float4 sl = reinterpret_cast<const float4 *>(a.slope)[q];
    const float4 as = reinterpret_cast<const float4 *>(a.aspect)[q];
    Fire f[4];
    bool ok[4];
    ok[0] = burn_cell(s_cat, codes & 0xffu, d1.x, d10.x, d100.x, herb.x, wood.x, ws.x, waz.x, sl.x, as.x, f[0]);
    ok[1] = burn_cell(s_cat, (codes >> 8) & 0xffu, d1.y, d10.y, d100.y, herb.y, wood.y, ws.y, waz.y, sl.y, as.y, f[1]);
    ok[2] = burn_cell(s_cat, (codes >> 16) & 0xffu, d1.z, d10.z, d100.z, herb.z, wood.z, ws.z, waz.z, sl.z, as.z, f[2]);
    ok[3] = burn_cell(s_cat, codes >> 24, d1.w, d10.w, d100.w, herb.w, wood.w, ws.w, waz.w, sl.w, as.w, f[3]);
    if (ok[0] && ok[1] && ok[2] && ok[3]) {
      reinterpret_cast<float4 *>(a.speed)[q] = make_float4(f[0].speed, f[1].speed, f[2].speed, f[3].speed);
      reinterpret_cast<float4 *>(a.az)[q] = make_float4(f[0].az, f[1].az, f[2].az, f[3].az);
      reinterpret_cast<float4 *>(a.ecc)[q] = make_float4(f[0].ecc, f[1].ecc, f[2].ecc, f[3].ecc);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!ok[k]) continue;  // leave the caller's value (cuda.rs:89-93)
        a.speed[4 * q + k] = f[k].speed;
        a.az[4 * q + k] = f[k].az;
        a.ecc[4 * q + k] = f[k].ecc;
      }
    }
  }
}

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

inline bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

}  // namespace

pg_status launch_burn(const DeviceInfo *dev, size_t n, const uint8_t *fuel, const float *d1,
                      const float *d10, const float *d100, const float *herb, const float *wood,
                      const float *ws, const float *waz, const float *slope, const float *aspect,
                      float *speed, float *az, float *ecc, hipStream_t stream) {
  if (n == 0) return PG_OK;
  BurnArgs a{fuel, d1, d10, d100, herb, wood, ws, waz, slope, aspect, speed, az, ecc};
  const float *fl[] = {d1, d10, d100, herb, wood, ws, waz, slope, aspect, speed, az, ecc};
  bool vec = (n % 4 == 0) && aligned(fuel, 4);
  for (const float *p : fl) vec = vec && aligned(p, 16);
  const unsigned block = 256;
  // Enough workgroups to cover every CU several times; the grid-stride loop
  // takes the rest.
  const size_t units = vec ? n / 4 : n;
  size_t blocks = (units + block - 1) / block;
  const size_t cap = static_cast<size_t>(dev->cu_count) * 16;
  if (blocks > cap) blocks = cap;
  if (vec)
    hipLaunchKernelGGL(burn_vec4, dim3(static_cast<unsigned>(blocks)), dim3(block), 0, stream, n / 4, a,
                       dev->catalog);
  else
    hipLaunchKernelGGL(burn_scalar, dim3(static_cast<unsigned>(blocks)), dim3(block), 0, stream, n, a,
                       dev->catalog);
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
