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

#include "pg_internal.h"

namespace pg {
namespace {

constexpr float kSmidgen = 1e-6f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kFtMinPerMs = 0.00508f;  // 1 ft/min in m/s

struct Fire {
  float speed, az, ecc;
};

// Fuel::eta_m (firelib.rs:1060-1076)
__device__ inline float eta_m(const DevFuel &f, bool alive, float life_moist, float wfmd) {
  float life_mext;
  if (!alive) {
    life_mext = f.mext;
  } else if (f.has_live) {
    const float fdmois = wfmd * f.fine_dead_factor;
    const float live_mext = f.live_ext_factor * (1.0f - fdmois * f.inv_mext) - 0.226f;
    life_mext = fmaxf(live_mext, f.mext);
  } else {
    life_mext = 0.0f;
  }
  if (life_moist >= life_mext) return 0.0f;
  const float rt = life_moist / life_mext;
  return 1.0f - 2.59f * rt + 5.11f * rt * rt - 3.52f * rt * rt * rt;
}

// Fuel::burn_simple for one cell.  Returns false for a code that names no
// fuel with particles (Catalog::get, firelib.rs:382-387).
__device__ inline bool burn_cell(const DevFuel *cat, uint32_t code, float d1, float d10,
                                 float d100, float herb, float wood, float ws_ms,
                                 float waz, float slope, float aspect, Fire &out) {
  if (code >= static_cast<uint32_t>(kNumFuels)) return false;
  const DevFuel &f = cat[code];
  if (f.np == 0) return false;
  const float moist[5] = {d1, d10, d100, herb, wood};

  // rx_int_rbqig (firelib.rs:1031-1058)
  float wfmd = 0.0f, alive_moist = 0.0f, dead_moist = 0.0f, rbqig = 0.0f;
#pragma unroll
  for (int i = 0; i < kMaxParticles; ++i) {
    if (i >= f.np) break;
    const DevParticle &p = f.p[i];
    const float m = moist[p.mclass];
    const float law = p.dead ? f.dead_area_weight : f.live_area_weight;
    rbqig += (250.0f + 1116.0f * m) * p.area_weight * law * p.sigma_factor;
    if (p.dead) {
      dead_moist += p.area_weight * m;
      wfmd += m * p.sigma_factor * p.load;
    } else {
      alive_moist += p.area_weight * m;
    }
  }
  const float rx_int = f.live_rx * eta_m(f, true, alive_moist, wfmd) +
                       f.dead_rx * eta_m(f, false, dead_moist, wfmd);
  const float speed0 = rx_int * f.flux_ratio / (f.bulk_density * rbqig);

  // calculate_wind_dependent_vars / wind_slope_situation (firelib.rs:1078-1167)
  const float ws = ws_ms / kFtMinPerMs;
  const float phi_slope = f.slope_k * slope * slope;
  const float phi_wind = f.wind_k * powf(ws, f.wind_b);
  const float phi_ew = phi_slope + phi_wind;
  const float upslope = aspect >= kPi ? aspect - kPi : aspect + kPi;
  float ew, s, a;
  bool limit = true;
  if (speed0 < kSmidgen) {  // NoSpread
    ew = 0.0f; s = 0.0f; a = 0.0f; limit = false;
  } else if (phi_ew < kSmidgen) {  // NoSlopeNoWind
    ew = 0.0f; s = speed0; a = 0.0f; limit = false;
  } else if (slope < kSmidgen) {  // WindNoSlope
    ew = ws; s = speed0 * (1.0f + phi_ew); a = waz;
  } else if (ws_ms < kSmidgen || fabsf(upslope - waz) < kSmidgen) {  // SlopeNoWind | UpSlope
    ew = powf(phi_ew * f.wind_e, f.wind_b_inv); s = speed0 * (1.0f + phi_ew); a = upslope;
  } else {  // CrossSlope
    const float wind_rate = speed0 * phi_wind;
    const float slope_rate = speed0 * phi_slope;
    const float split = upslope <= waz ? waz - upslope : 2.0f * kPi - upslope + waz;
    const float x = slope_rate + wind_rate * cosf(split);
    const float y = wind_rate * sinf(split);
    const float rv = sqrtf(x * x + y * y);
    const float smax2 = speed0 + rv;
    const float phi2 = smax2 / speed0 - 1.0f;
    const float al = asinf(fabsf(y) / rv);
    float split2;
    if (x >= 0.0f && y >= 0.0f) split2 = al;
    else if (x >= 0.0f) split2 = 2.0f * kPi - al;
    else if (y >= 0.0f) split2 = kPi - al;
    else split2 = kPi + al;
    float az2 = upslope + split2;
    if (az2 > 2.0f * kPi) az2 = az2 - 2.0f * kPi;
    ew = powf(phi2 * f.wind_e, f.wind_b_inv); s = smax2; a = az2;
  }
  const float max_wind = 0.9f * rx_int;
  if (limit && ew > max_wind) {  // check_wind_limit
    s = speed0 * (1.0f + f.wind_k * powf(max_wind, f.wind_b));
    ew = max_wind;
  }
  const float lw = 1.0f + 0.002840909f * ew;  // Fuel::eccentricity
  out.speed = s * kFtMinPerMs;
  out.az = a;
  out.ecc = sqrtf(lw * lw - 1.0f) / lw;
  return true;
}

__device__ inline void stage_catalog(DevFuel *s_cat, const DevCatalog *cat) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(cat);
  uint32_t *dst = reinterpret_cast<uint32_t *>(s_cat);
  for (unsigned i = threadIdx.x; i < sizeof(DevCatalog) / 4; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

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
