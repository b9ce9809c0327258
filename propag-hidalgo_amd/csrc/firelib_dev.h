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

// x^y for the burn's positive finite bases (wind speed, effective-wind
// terms): exp2(y * log2 x) with x = m * 2^e split exactly (m in [0.5, 1)),
// log2 m from the hardware log2 (V_LOG_F32, |log2 m| <= 1, so its error is
// absolute and at most an ulp of 1), e + log2 m and the product with y
// formed in double, 2^(fraction) from the hardware exp2 (V_EXP_F32) on
// [-0.5, 0.5] and the integer part applied exactly by ldexp.  For the
// exponents used here (|y| <= 3) the result is within about 3e-7 relative
// of the exact power; 0^y (y > 0) is 0 and negative, infinite and NaN bases
// take powf.
__device__ inline float pow_pos(float x, float y) {
  if (!(x > 0.0f && x <= 3.40282347e38f)) return x == 0.0f && y > 0.0f ? 0.0f : powf(x, y);
  const int e = __builtin_amdgcn_frexp_expf(x);
  const float m = __builtin_amdgcn_frexp_mantf(x);
  const double t = static_cast<double>(y) * (static_cast<double>(e) + static_cast<double>(__builtin_amdgcn_logf(m)));
  const double n = rint(t);
  return ldexpf(__builtin_amdgcn_exp2f(static_cast<float>(t - n)), static_cast<int>(n));
}

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
  const float phi_wind = f.wind_k * pow_pos(ws, f.wind_b);
  const float phi_ew = phi_slope + phi_wind;
  const float upslope = aspect >= kPi ? aspect - kPi : aspect + kPi;
  // Each situation sets s and a, and ew either directly or, for the two
  // that need (phi * wind_e)^(1 / wind_b), through pe (one power call site,
  // so a wave runs it once whichever of the two its lanes take).
  float ew = 0.0f, s, a, pe = 0.0f;
  bool limit = true, ew_pow = false;
  if (speed0 < kSmidgen) {  // NoSpread
    s = 0.0f; a = 0.0f; limit = false;
  } else if (phi_ew < kSmidgen) {  // NoSlopeNoWind
    s = speed0; a = 0.0f; limit = false;
  } else if (slope < kSmidgen) {  // WindNoSlope
    ew = ws; s = speed0 * (1.0f + phi_ew); a = waz;
  } else if (ws_ms < kSmidgen || fabsf(upslope - waz) < kSmidgen) {  // SlopeNoWind | UpSlope
    pe = phi_ew; ew_pow = true; s = speed0 * (1.0f + phi_ew); a = upslope;
  } else {  // CrossSlope
    const float wind_rate = speed0 * phi_wind;
    const float slope_rate = speed0 * phi_slope;
    const float split = upslope <= waz ? waz - upslope : 2.0f * kPi - upslope + waz;
    float sn, cs;
    sincosf(split, &sn, &cs);
    const float x = slope_rate + wind_rate * cs;
    const float y = wind_rate * sn;
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
    pe = phi2; ew_pow = true; s = smax2; a = az2;
  }
  if (ew_pow) ew = pow_pos(pe * f.wind_e, f.wind_b_inv);
  const float max_wind = 0.9f * rx_int;
  if (limit && ew > max_wind) {  // check_wind_limit
    s = speed0 * (1.0f + f.wind_k * pow_pos(max_wind, f.wind_b));
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
