// Moisture pre-pass: per-cell dead and live fuel moisture at one hour of a
// day of weather, the rasters the burn pass reads.
//
// Replaces the host-side chain crates/firelib/src/fuel_moisture.rs:26-63
// (create_terrain_with_fuel_moisture) -> crates/dmoist/src/lib.rs:55-145
// (calculate_hourly_fuel_moisture), evaluated per cell on the device.  The
// model: a base fine-fuel moisture from a (temperature, humidity) table
// (lib.rs:197-262), a shading correction by aspect, slope, cloud cover,
// month and hour (lib.rs:472-631), an ignition probability table
// (lib.rs:272-454) lowered by the rain of the last six days (lib.rs:147-195,
// 633-650), the 1-h moisture as moisture x probability ratio (lib.rs:652-
// 663), 10-h and 100-h moistures interpolated between the next finer
// class's values at 06:00 and 15:00 (lib.rs:665-703), live moisture by month
// (lib.rs:705-716); ratios = percentages / 100.  The tables are the
// reference's data.  Arithmetic is f32 in the reference's order, compiled
// without contraction.
//
// Per cell it reads fuel (1 B), slope and aspect (8 B) and writes five f32
// rasters (20 B): HBM-bound at 29 B/cell.  The weather is per scenario (the
// kernel argument); the per-hour base moisture and the rain effect are
// folded on the host.  The three tables (about 1 KB) are staged in LDS per
// workgroup since lanes index them by their own cell's classes.
#include <cmath>

#include "pg_internal.h"

namespace pg {
namespace {

constexpr float kPiF = 3.14159274101257324f;  // f32 pi, as the reference's f32::PI

// Ignition probability by [shading class][temperature class][moisture - 1].
__constant__ unsigned char kIgnition[4][9][18] = {
    {{100, 100, 100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 30, 20, 20, 20, 10, 10},
     {100, 100, 90, 80, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 20, 10, 10, 10},
     {100, 100, 80, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 20, 10, 10, 10, 10}},
    {{100, 100, 100, 80, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 20, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 100, 80, 70, 60, 50, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 20, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 80, 70, 60, 50, 50, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10, 10}},
    {{100, 100, 90, 80, 70, 60, 50, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 100, 80, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 90, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 80, 70, 60, 50, 50, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10, 10}},
    {{100, 100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10},
     {100, 100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 20, 10, 10, 10},
     {100, 100, 80, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 50, 40, 30, 30, 30, 20, 20, 20, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 90, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 90, 70, 60, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10},
     {100, 80, 70, 60, 50, 50, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10, 10},
     {100, 80, 70, 60, 50, 40, 40, 30, 30, 20, 20, 20, 10, 10, 10, 10, 10, 10}},
};

// Shading correction by [month class][shaded][side * 2 + steep][hour class].
__constant__ unsigned char kShadeCorr[3][2][8][6] = {
    {{{3, 1, 0, 0, 1, 3}, {4, 2, 1, 1, 2, 4}, {2, 1, 0, 0, 1, 4}, {2, 0, 0, 1, 3, 5},
      {3, 1, 0, 0, 1, 3}, {3, 1, 1, 1, 1, 3}, {3, 1, 0, 0, 1, 3}, {5, 3, 1, 0, 0, 2}},
     {{5, 4, 3, 3, 4, 5}, {5, 4, 3, 3, 4, 5}, {4, 4, 3, 4, 4, 5}, {4, 4, 3, 4, 4, 5},
      {4, 4, 3, 3, 4, 5}, {4, 4, 3, 3, 4, 5}, {5, 4, 3, 3, 4, 4}, {5, 4, 3, 3, 4, 4}}},
    {{{4, 2, 1, 1, 2, 4}, {4, 3, 3, 3, 3, 4}, {4, 2, 1, 1, 2, 4}, {3, 1, 1, 2, 4, 5},
      {4, 2, 1, 1, 2, 4}, {4, 2, 1, 1, 2, 4}, {4, 2, 1, 1, 2, 4}, {5, 4, 2, 1, 1, 3}},
     {{5, 5, 4, 4, 5, 5}, {5, 5, 4, 4, 5, 5}, {5, 4, 4, 4, 5, 5}, {5, 4, 4, 4, 5, 5},
      {5, 4, 4, 4, 4, 5}, {5, 4, 4, 4, 4, 5}, {5, 5, 4, 4, 4, 5}, {5, 5, 4, 4, 4, 5}}},
    {{{5, 4, 3, 3, 4, 5}, {5, 5, 5, 5, 5, 5}, {5, 4, 3, 3, 4, 5}, {5, 4, 3, 2, 5, 5},
      {5, 4, 3, 2, 4, 5}, {5, 3, 1, 1, 3, 5}, {5, 4, 3, 3, 4, 5}, {5, 5, 4, 2, 3, 5}},
     {{5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5},
      {5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5}, {5, 5, 5, 5, 5, 5}}},
};

// Per-scenario, per-hour values folded on the host: for the three hours the
// cell needs (the requested one, 06:00, 15:00), the base moisture (table by
// temperature and humidity only), temperature and cloud cover; plus the
// day's rain effect, month and live moisture ratio.
struct HourArgs {
  float base, temperature, cloud;
  int hour;
};
struct MoistArgs {
  HourArgs h[3];  // requested hour, 06:00, 15:00
  float rain_max;
  float live;
  int month;
  int hour;
};

struct MoistLds {
  unsigned char ign[4 * 9 * 18];
  unsigned char corr[3 * 2 * 8 * 6];
};

__device__ inline float shading_of(float cloud, int fuel) {
  const float s = cloud * 0.75f;
  return (fuel >= 7 && fuel <= 12) ? (s < 50.0f ? 75.0f : 100.0f) : s;
}

// Shading correction (lib.rs:472-614); NaN for a month or aspect outside
// the tables.
__device__ inline float correction(const MoistLds &t, float aspect_deg, float slope_deg, float shading, int month,
                                   int hour_in) {
  int hour = hour_in;
  if (!(hour_in >= 6 && hour_in <= 19)) {
    hour = 19;
    shading = 100.0f;
  } else if (hour_in < 8) {
    hour = 8;
    shading = 100.0f;
  }
  const int sh = shading < 50.0f ? 0 : 1;
  const int hi = min(max((hour - 8) / 2, 0), 5);
  int mo;
  if (month >= 5 && month <= 8)
    mo = 0;
  else if (month == 2 || month == 3 || month == 4 || month == 9 || month == 10)
    mo = 1;
  else if (month == 11 || month == 12 || month == 1)
    mo = 2;
  else
    return __builtin_nanf("");
  const float a = slope_deg == 0.0f ? 180.0f : aspect_deg;  // flat counts as south
  int side;
  if ((a >= 315.0f && a <= 360.0f) || (a >= 0.0f && a < 45.0f))
    side = 0;
  else if (a >= 45.0f && a < 135.0f)
    side = 1;
  else if (a >= 135.0f && a < 225.0f)
    side = 2;
  else if (a >= 225.0f && a < 315.0f)
    side = 3;
  else
    return __builtin_nanf("");
  const int row = side * 2 + (slope_deg <= 30.0f ? 0 : 1);
  return static_cast<float>(t.corr[((mo * 2 + sh) * 8 + row) * 6 + hi]);
}

__device__ inline float ignition(const MoistLds &t, float temperature, float cloud, float moisture, int fuel) {
  if (fuel == 0 || !(moisture > 0.0f && moisture < 18.0f)) return 0.0f;
  const float s = shading_of(cloud, fuel);
  const int si = (s >= 0.0f && s < 10.0f) ? 0 : (s >= 10.0f && s < 50.0f) ? 1 : (s >= 50.0f && s < 90.0f) ? 2 : 3;
  float tf = 9.0f - ceilf(temperature / 5.0f);
  tf = tf >= 0.0f ? (tf > 8.0f ? 8.0f : tf) : 0.0f;  // NaN -> 0
  const int mi = max(static_cast<int>(floorf(moisture)) - 1, 0);
  return static_cast<float>(t.ign[(si * 9 + static_cast<int>(tf)) * 18 + mi]);
}

__device__ inline float rain_corrected(float p, float effect) {
  float adj = 0.0f;
  if (effect > 0.0f && effect < 5.0f)
    adj = -5.0f;
  else if (effect >= 5.0f && effect < 15.0f)
    adj = -10.0f;
  else if (effect >= 15.0f && effect < 35.0f)
    adj = -20.0f;
  else if (effect >= 35.0f)
    adj = -30.0f;
  const float r = p + adj;
  return r < 0.0f ? 0.0f : (r > 100.0f ? 100.0f : r);
}

// 1-h moisture (%) of a cell at one of the three hours.
__device__ inline float fine(const MoistLds &t, const MoistArgs &a, const HourArgs &h, float aspect_deg,
                             float slope_deg, int fuel) {
  const float m =
      fuel == 0 ? 0.0f : h.base + correction(t, aspect_deg, slope_deg, shading_of(h.cloud, fuel), a.month, h.hour);
  const float p = ignition(t, h.temperature, h.cloud, m, fuel);
  const float pw = rain_corrected(p, a.rain_max);
  if (pw <= 0.0f) return m > 0.0f ? 30.0f : 0.0f;
  return p / pw * m;
}

// Next-thicker class (lib.rs:665-703).
__device__ inline float thicker(int hour, float m, float m6, float m15) {
  if (m == 0.0f) return 0.0f;
  const float at6 = m15 + (0.4142f * (m6 - m15));
  const float at15 = m6 - (0.5571f * (m6 - m15));
  if (fabsf(at6 - at15) < 0.001f) return at6 < 0.001f ? m : at6;
  if (hour >= 0 && hour < 6) return at15 + (((at6 - at15) / 15.0f) * static_cast<float>(hour + 15));
  if (hour >= 6 && hour < 15) return at6 + (((at15 - at6) / 9.0f) * static_cast<float>(hour - 6));
  if (hour >= 15 && hour < 24) return at15 + (((at6 - at15) / 15.0f) * static_cast<float>(hour - 15));
  return m;
}

__global__ __launch_bounds__(256) void moisture_kernel(size_t n, const uint8_t *__restrict__ fuel,
                                                       const float *__restrict__ slope,
                                                       const float *__restrict__ aspect, MoistArgs a,
                                                       float *__restrict__ d1, float *__restrict__ d10,
                                                       float *__restrict__ d100, float *__restrict__ herb,
                                                       float *__restrict__ wood) {
  __shared__ MoistLds t;
  for (unsigned i = threadIdx.x; i < sizeof(t.ign); i += blockDim.x) t.ign[i] = (&kIgnition[0][0][0])[i];
  for (unsigned i = threadIdx.x; i < sizeof(t.corr); i += blockDim.x) t.corr[i] = (&kShadeCorr[0][0][0][0])[i];
  __syncthreads();
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int f = fuel[i];
    const float aspect_deg = aspect[i] * 180.0f / kPiF;
    const float slope_deg = atanf(slope[i]) * 180.0f / kPiF;
    const float m = fine(t, a, a.h[0], aspect_deg, slope_deg, f);
    const float m6 = fine(t, a, a.h[1], aspect_deg, slope_deg, f);
    const float m15 = fine(t, a, a.h[2], aspect_deg, slope_deg, f);
    const float t10 = thicker(a.hour, m, m6, m15);
    const float t6 = thicker(6, m6, m6, m15);
    const float t15 = thicker(15, m15, m6, m15);
    d1[i] = m / 100.0f;
    d10[i] = t10 / 100.0f;
    d100[i] = thicker(a.hour, t10, t6, t15) / 100.0f;
    herb[i] = a.live;
    wood[i] = a.live;
  }
}

// Base fine-fuel moisture (lib.rs:197-262): the day table, by temperature
// class and humidity in 5 % steps (the reference uses it at every hour).
int base_moisture(float temperature, float humidity) {
  static const unsigned char day[6][21] = {
      {1, 2, 2, 3, 4, 5, 5, 6, 7, 8, 8, 8, 9, 9, 10, 11, 12, 12, 13, 13, 14},
      {1, 2, 2, 3, 4, 5, 5, 6, 7, 7, 7, 8, 9, 9, 10, 10, 11, 12, 13, 13, 13},
      {1, 2, 2, 3, 4, 5, 5, 6, 6, 7, 7, 8, 9, 9, 9, 10, 11, 12, 12, 12, 13},
      {1, 1, 2, 2, 3, 4, 5, 5, 6, 7, 7, 8, 8, 8, 9, 10, 10, 11, 12, 12, 13},
      {1, 1, 2, 2, 3, 4, 4, 5, 6, 7, 7, 8, 8, 8, 9, 10, 10, 11, 12, 12, 13},
      {1, 1, 2, 2, 3, 4, 4, 5, 6, 7, 7, 8, 8, 8, 9, 10, 10, 11, 12, 12, 12},
  };
  float hf = std::floor(humidity / 5.0f);
  hf = hf >= 0.0f ? (hf > 20.0f ? 20.0f : hf) : 0.0f;  // NaN -> 0
  int ti;
  if (temperature < 0.0f)
    ti = 0;
  else if (temperature <= 9.0f)
    ti = 1;
  else if (temperature <= 20.0f)
    ti = 2;
  else if (temperature <= 31.0f)
    ti = 3;
  else if (temperature <= 42.0f)
    ti = 4;
  else
    ti = 5;
  return day[ti][static_cast<int>(hf)];
}

// Rain effect of the six previous days, the largest (lib.rs:147-195).
float rain_effect_max(int month, const float rain[6]) {
  static const float table[2][4][6] = {
      {{-1, -1, 0, 0, 0, 0}, {-1, 5, 1, 0, 0, 0}, {-1, 15, 5, 1, 0, 0}, {-1, 35, 15, 5, 1, 0}},
      {{-1, -1, -1, -1, 1, 0}, {-1, -1, -1, -1, 5, 1}, {-1, -1, -1, 15, 5, 1}, {-1, -1, 35, 15, 5, 1}},
  };
  float best = -INFINITY;
  for (int i = 0; i < 6; ++i) {
    float e = rain[i];
    if (rain[i] > 0.0f) {
      const int season = (month >= 4 && month <= 10) ? 0 : 1;
      const float r = rain[i];
      const int cls = (r >= 1.0f && r < 5.0f) ? 0 : (r >= 5.0f && r < 15.0f) ? 1 : (r >= 15.0f && r < 35.0f) ? 2 : 3;
      const float t = table[season][cls][i];
      if (t > -1.0f) e = t;
    }
    best = std::fmax(best, e);
  }
  return best;
}

float live_ratio(int month) {
  float v;
  switch (month) {
    case 1: case 2: case 6: case 11: case 12: v = 100.0f; break;
    case 3: case 4: case 5: v = 200.0f; break;
    case 7: case 8: v = 80.0f; break;
    case 9: case 10: v = 90.0f; break;
    default: return std::nanf("");
  }
  return v / 100.0f;
}

}  // namespace

pg_status launch_moisture(const DeviceInfo *dev, size_t n, const uint8_t *fuel, const float *slope,
                          const float *aspect, const pg_day_weather *w, int hour, float *d1hr, float *d10hr,
                          float *d100hr, float *herb, float *wood, hipStream_t stream) {
  MoistArgs a{};
  const int hours[3] = {hour, 6, 15};
  for (int k = 0; k < 3; ++k) {
    const int h = hours[k];
    a.h[k].base = static_cast<float>(base_moisture(w->temperature[h], w->humidity[h]));
    a.h[k].temperature = w->temperature[h];
    a.h[k].cloud = w->cloud_cover[h];
    a.h[k].hour = h;
  }
  a.rain_max = rain_effect_max(w->month, w->rain_6_days);
  a.live = live_ratio(w->month);
  a.month = w->month;
  a.hour = hour;
  size_t blocks = (n + 255) / 256;
  const size_t cap = static_cast<size_t>(dev->cu_count) * 16;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(moisture_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, n, fuel, slope,
                     aspect, a, d1hr, d10hr, d100hr, herb, wood);
  PG_HIP(hipGetLastError());
  return PG_OK;
}

}  // namespace pg
