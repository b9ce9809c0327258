// Per-cell firelib burn (Fuel::burn_simple, crates/firelib/src/firelib.rs:
// 394-397,1016-1188) as device functions: used by the Stage F kernels
// (burn.hip) and by the fused scenario admission (propag.hip).  Arithmetic is
// f32 in the reference's operation order; the sources that include this file
// are compiled without FMA contraction.
#pragma once

#include "pg_internal.h"

namespace pg {
namespace firelib {

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

}  // namespace firelib
}  // namespace pg
