/*
 * Oracle (test infrastructure only): the dead/live fuel moisture pre-pass
 * that feeds the burn pass its moisture rasters, restated from
 * crates/dmoist/src/lib.rs (hourly model of the reference, driven by
 * crates/firelib/src/fuel_moisture.rs:26-63).  See oracle.h for who may use
 * this.
 *
 * Per hour h of one day the model reads a base fine-fuel moisture from a
 * (temperature, relative humidity) table, corrects it for shading by
 * aspect/slope/cloud/month/hour, reads an ignition probability, lowers it by
 * the rain of the last six days, and scales the moisture by the ratio of the
 * two probabilities (1-h moisture).  10-h and 100-h moistures interpolate
 * between the values at 06:00 and 15:00 of the next finer class; live
 * moisture depends on the month only.  All arithmetic is f32 as in the
 * reference; the tables are the reference's data.
 *
 * Pinned by tests/test_oracle_moisture.py against the reference's own
 * fixture files (crates/dmoist/fixtures, copied as data to
 * tests/golden/dmoist/).
 */
#include "oracle.h"

#include <math.h>

/* Rain effect of one day (crates/dmoist/src/lib.rs:147-183): by season
 * (April-October or not), rain class and days back; -1 entries, no rain or
 * rain below every class keep the rain amount itself. */
static float rain_effect(int month, float rain, int days_back) {
  static const float table[2][4][6] = {
      {{-1, -1, 0, 0, 0, 0}, {-1, 5, 1, 0, 0, 0}, {-1, 15, 5, 1, 0, 0}, {-1, 35, 15, 5, 1, 0}},
      {{-1, -1, -1, -1, 1, 0}, {-1, -1, -1, -1, 5, 1}, {-1, -1, -1, 15, 5, 1}, {-1, -1, 35, 15, 5, 1}},
  };
  if (rain > 0.0f) {
    const int season = (month >= 4 && month <= 10) ? 0 : 1;
    int cls;
    if (rain >= 1.0f && rain < 5.0f)
      cls = 0;
    else if (rain >= 5.0f && rain < 15.0f)
      cls = 1;
    else if (rain >= 15.0f && rain < 35.0f)
      cls = 2;
    else
      cls = 3; /* 35 and above, and (0, 1) */
    const float e = table[season][cls][days_back];
    if (e > -1.0f) return e;
  }
  return rain;
}

/* lib.rs:185-195 */
float oracle_rain_effect_max(int month, const float rain[6]) {
  float m = -INFINITY;
  for (int i = 0; i < 6; ++i) {
    const float e = rain_effect(month, rain[i], i);
    m = fmaxf(m, e);
  }
  return m;
}

/* Base fine-fuel moisture (lib.rs:197-262): the day table only, whatever the
 * hour (the reference's hour index is fixed at 0). */
int oracle_base_moisture(float temperature, float humidity) {
  static const unsigned char day[6][21] = {
      {1, 2, 2, 3, 4, 5, 5, 6, 7, 8, 8, 8, 9, 9, 10, 11, 12, 12, 13, 13, 14},
      {1, 2, 2, 3, 4, 5, 5, 6, 7, 7, 7, 8, 9, 9, 10, 10, 11, 12, 13, 13, 13},
      {1, 2, 2, 3, 4, 5, 5, 6, 6, 7, 7, 8, 9, 9, 9, 10, 11, 12, 12, 12, 13},
      {1, 1, 2, 2, 3, 4, 5, 5, 6, 7, 7, 8, 8, 8, 9, 10, 10, 11, 12, 12, 13},
      {1, 1, 2, 2, 3, 4, 4, 5, 6, 7, 7, 8, 8, 8, 9, 10, 10, 11, 12, 12, 13},
      {1, 1, 2, 2, 3, 4, 4, 5, 6, 7, 7, 8, 8, 8, 9, 10, 10, 11, 12, 12, 12},
  };
  float hf = floorf(humidity / 5.0f);
  hf = hf >= 0.0f ? (hf > 20.0f ? 20.0f : hf) : 0.0f; /* NaN -> 0, as a saturating cast */
  const int hi = (int)hf;
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
    ti = 5; /* above 42, and NaN */
  return day[ti][hi];
}

/* lib.rs:264-270 */
float oracle_shading(float cloud, int fuel) {
  float s = cloud * 0.75f;
  if (fuel >= 7 && fuel <= 12) s = s < 50.0f ? 75.0f : 100.0f;
  return s;
}

/* Ignition probability (lib.rs:272-454): by shading class, temperature
 * class and the whole part of the corrected moisture (1..17). */
float oracle_ignition_probability(float temperature, float cloud, float moisture, int fuel) {
  static const unsigned char t[4][9][18] = {
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
  if (fuel == 0) return 0.0f;
  if (!(moisture > 0.0f && moisture < 18.0f)) return 0.0f;
  const float s = oracle_shading(cloud, fuel);
  int si;
  if (s >= 0.0f && s < 10.0f)
    si = 0;
  else if (s >= 10.0f && s < 50.0f)
    si = 1;
  else if (s >= 50.0f && s < 90.0f)
    si = 2;
  else
    si = 3;
  float tf = 9.0f - ceilf(temperature / 5.0f);
  tf = tf >= 0.0f ? (tf > 8.0f ? 8.0f : tf) : 0.0f; /* NaN -> 0 */
  /* whole part minus one; the reference's index underflows (a panic) for a
   * moisture below 1, which its tables never produce: the first column is
   * used there */
  int mi = (int)floorf(moisture) - 1;
  if (mi < 0) mi = 0;
  return (float)t[si][(int)tf][mi];
}

/* Shading correction of the base moisture (lib.rs:472-614).  Aspect and
 * slope in degrees; NaN for a month outside 1..12 or an aspect outside
 * [0, 360]. */
static float shading_correction(float aspect_deg, float slope_deg, float shading, int month, int hour_in) {
  static const unsigned char t[3][2][8][6] = {
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
  /* night hours use 19:00 fully shaded, early morning 08:00 fully shaded */
  int hour = hour_in;
  if (!(hour_in >= 6 && hour_in <= 19)) {
    hour = 19;
    shading = 100.0f;
  } else if (hour_in < 8) {
    hour = 8;
    shading = 100.0f;
  }
  const int sh = shading < 50.0f ? 0 : 1;
  int hi = (hour - 8) / 2;
  hi = hi < 0 ? 0 : (hi > 5 ? 5 : hi);
  int mo;
  if (month >= 5 && month <= 8)
    mo = 0;
  else if (month == 2 || month == 3 || month == 4 || month == 9 || month == 10)
    mo = 1;
  else if (month == 11 || month == 12 || month == 1)
    mo = 2;
  else
    return NAN;
  const float a = slope_deg == 0.0f ? 180.0f : aspect_deg; /* flat counts as south */
  int side;
  if ((a >= 315.0f && a <= 360.0f) || (a >= 0.0f && a < 45.0f))
    side = 0; /* north */
  else if (a >= 45.0f && a < 135.0f)
    side = 1; /* east */
  else if (a >= 135.0f && a < 225.0f)
    side = 2; /* south */
  else if (a >= 225.0f && a < 315.0f)
    side = 3; /* west */
  else
    return NAN;
  const int row = side * 2 + (slope_deg <= 30.0f ? 0 : 1);
  return (float)t[mo][sh][row][hi];
}

/* lib.rs:616-631 */
float oracle_corrected_moisture(float base, float cloud, int month, float aspect_deg, float slope_deg, int fuel,
                                int hour) {
  if (fuel == 0) return 0.0f;
  return base + shading_correction(aspect_deg, slope_deg, oracle_shading(cloud, fuel), month, hour);
}

/* lib.rs:633-650 */
float oracle_rain_corrected_probability(float probability, float effect) {
  float adj;
  if (effect > 0.0f && effect < 5.0f)
    adj = -5.0f;
  else if (effect >= 5.0f && effect < 15.0f)
    adj = -10.0f;
  else if (effect >= 15.0f && effect < 35.0f)
    adj = -20.0f;
  else if (effect >= 35.0f)
    adj = -30.0f;
  else
    adj = 0.0f;
  const float p = probability + adj;
  return p < 0.0f ? 0.0f : (p > 100.0f ? 100.0f : p);
}

/* lib.rs:652-663 */
float oracle_fine_moisture(float moisture, float prob_dry, float prob_wet) {
  if (prob_wet <= 0.0f) return moisture > 0.0f ? 30.0f : 0.0f;
  return prob_dry / prob_wet * moisture;
}

/* Next-thicker class (lib.rs:665-703): values at 06:00 and 15:00 from the
 * finer class's, then linear in the hour between them. */
float oracle_thicker_moisture(int hour, float m, float m6, float m15) {
  if (m == 0.0f) return 0.0f;
  const float at6 = m15 + (0.4142f * (m6 - m15));
  const float at15 = m6 - (0.5571f * (m6 - m15));
  if (fabsf(at6 - at15) < 0.001f) return at6 < 0.001f ? m : at6;
  if (hour >= 0 && hour < 6) return at15 + (((at6 - at15) / 15.0f) * (float)(hour + 15));
  if (hour >= 6 && hour < 15) return at6 + (((at15 - at6) / 9.0f) * (float)(hour - 6));
  if (hour >= 15 && hour < 24) return at15 + (((at6 - at15) / 15.0f) * (float)(hour - 15));
  return m;
}

/* lib.rs:705-716 */
float oracle_live_moisture(int month) {
  switch (month) {
    case 1: case 2: case 6: case 11: case 12: return 100.0f;
    case 3: case 4: case 5: return 200.0f;
    case 7: case 8: return 80.0f;
    case 9: case 10: return 90.0f;
    default: return NAN;
  }
}

/* 1-h moisture of one hour (lib.rs:71-110). */
static float hour_fine(const oracle_day_weather *w, float aspect_deg, float slope_deg, int fuel, float rain_max,
                       int hour) {
  const float base = (float)oracle_base_moisture(w->temperature[hour], w->humidity[hour]);
  const float corr = oracle_corrected_moisture(base, w->cloud[hour], w->month, aspect_deg, slope_deg, fuel, hour);
  const float p = oracle_ignition_probability(w->temperature[hour], w->cloud[hour], corr, fuel);
  const float pr = oracle_rain_corrected_probability(p, rain_max);
  return oracle_fine_moisture(corr, p, pr);
}

/* Moisture ratios of every cell at one hour, as
 * crates/firelib/src/fuel_moisture.rs:38-62 fills TerrainCuda from
 * calculate_hourly_fuel_moisture (lib.rs:55-145): percentages / 100; aspect
 * (rad) and slope (rise/run) converted to degrees in f32 as lib.rs:78-79. */
void oracle_moisture(size_t n, const uint8_t *fuel, const float *slope, const float *aspect,
                     const oracle_day_weather *w, int hour, float *d1hr, float *d10hr, float *d100hr, float *herb,
                     float *wood) {
  const float pi = 3.14159274101257324f; /* f32 pi */
  const float rain_max = oracle_rain_effect_max(w->month, w->rain);
  const float live = oracle_live_moisture(w->month) / 100.0f;
  for (size_t i = 0; i < n; ++i) {
    const float aspect_deg = aspect[i] * 180.0f / pi;
    const float slope_deg = atanf(slope[i]) * 180.0f / pi;
    const int f = fuel[i];
    const float m = hour_fine(w, aspect_deg, slope_deg, f, rain_max, hour);
    const float m6 = hour_fine(w, aspect_deg, slope_deg, f, rain_max, 6);
    const float m15 = hour_fine(w, aspect_deg, slope_deg, f, rain_max, 15);
    const float t = oracle_thicker_moisture(hour, m, m6, m15);
    const float t6 = oracle_thicker_moisture(6, m6, m6, m15);
    const float t15 = oracle_thicker_moisture(15, m15, m6, m15);
    d1hr[i] = m / 100.0f;
    d10hr[i] = t / 100.0f;
    d100hr[i] = oracle_thicker_moisture(hour, t, t6, t15) / 100.0f;
    herb[i] = live;
    wood[i] = live;
  }
}
