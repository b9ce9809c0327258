/*
 * CPU oracle for the two GPU stages of propag25 (albertov/propag-hidalgo).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the timed CPU baseline.  The product library never links or calls it.
 *
 * Stage F: a restatement of firelib's Rothermel/BEHAVE per-cell pass
 *   (crates/firelib/src/firelib.rs:996-1189) with the fuel catalog derived the
 *   way crates/firelib/src/lib.rs:125-268 + firelib.rs:554-979 derive it, in
 *   f32 (the reference's float::T) and as an FP64 twin of the same equations.
 * Stage P: a serial transcription of the update rule of
 *   crates/propag/src/propag.cu:124-207,269-302,345-375 and the DDA of
 *   crates/propag/src/propag.h:89-135, iterated as exact Jacobi sweeps (every
 *   cell reads the state of the previous sweep, as inside one cooperative
 *   launch of propag.cu:62-109) until a sweep changes nothing, then the
 *   post-pass of propag.cu:110-117.
 */
#ifndef PROPAG_ORACLE_H
#define PROPAG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Stage F -------------------------------------------- */

/* Per-cell SoA burn, f32.  Writes outputs only where the fuel code names a
 * catalog fuel with particles (cuda.rs:89-93); other cells keep what the
 * caller put there.  Returns the number of cells written. */
size_t oracle_burn_f32(size_t n, const uint8_t *fuel, const float *d1hr,
                       const float *d10hr, const float *d100hr,
                       const float *herb, const float *wood,
                       const float *wind_speed, const float *wind_azimuth,
                       const float *slope, const float *aspect,
                       float *speed_max, float *azimuth_max,
                       float *eccentricity, int nthreads);

/* Same equations in double; the catalog is derived in double straight from
 * the imperial table (no f32 unit round trips).  Inputs are the f32 inputs
 * widened.  Outputs in m/s, rad, ratio. */
size_t oracle_burn_f64(size_t n, const uint8_t *fuel, const float *d1hr,
                       const float *d10hr, const float *d100hr,
                       const float *herb, const float *wood,
                       const float *wind_speed, const float *wind_azimuth,
                       const float *slope, const float *aspect,
                       double *speed_max, double *azimuth_max,
                       double *eccentricity);

/* Catalog introspection for tests: writes the f32 scalar constants of fuel
 * `code` (0..13) into out[0..16] in this order:
 * sigma, beta, flux_ratio, slope_k, wind_b, wind_e, wind_k,
 * live_rx_factor, dead_rx_factor, live_area_weight, dead_area_weight,
 * fine_dead_factor, live_ext_factor, bulk_density, residence_time, mext.
 * Returns the number of particles (0 = no fuel). */
int oracle_catalog_f32(int code, float *out);
int oracle_catalog_f64(int code, double *out);

/* ---------------- Stage P -------------------------------------------- */

typedef struct {
  uint32_t width, height;
  float dx, dy;     /* gt.dx, gt.dy of the geotransform (dy < 0: north-up) */
  float max_time;
} oracle_grid;

typedef struct {
  int64_t sweeps;        /* Jacobi sweeps run, including the final quiet one */
  int64_t relaxations;   /* in-bounds cells evaluated                         */
  int64_t improvements;  /* cell updates committed                            */
  int64_t dda_steps;     /* DDA cells visited                                 */
} oracle_stats;

/* Boundary bitset of propag.cu:345-375 (bits must be zeroed by the caller). */
void oracle_boundaries(const oracle_grid *g, const float *speed_max,
                       const float *azimuth_max, const float *eccentricity,
                       uint32_t *boundaries);

/* Arrival-time propagation.  time/refs_x/refs_y are in-out and follow the
 * reference's initial state (propag.rs:433-449): ignited cells hold their
 * time and refs == their own position, the rest FLT_MAX and 0xFFFF.
 * formula 0: the closed algebraic form of time_from shared with the HIP
 *            kernel (cos(az - az_max) from cell offsets, no trig per step);
 * formula 1: the literal reference arithmetic with libm atan2f/cosf and
 *            IEEE division/sqrt (used to quantify rounding sensitivity).
 * max_sweeps <= 0 means unbounded.  nthreads > 1 parallelises each sweep
 * (results are identical: a sweep only reads the previous state).
 * Returns 0 on convergence, 1 if max_sweeps was hit. */
int oracle_propagate(const oracle_grid *g, const float *speed_max,
                     const float *azimuth_max, const float *eccentricity,
                     float *time, uint16_t *refs_x, uint16_t *refs_y,
                     uint32_t *boundaries, int formula, int64_t max_sweeps,
                     int nthreads, oracle_stats *stats);

/* One evaluation of time_from (propag.cu:269-294) for tests. */
float oracle_time_from(const oracle_grid *g, float from_time, int fx, int fy,
                       int tx, int ty, float speed_max, float azimuth_max,
                       float eccentricity, int formula);

/* The f32 DDA walk of propag.h:91-135 from `from` to `to` (inclusive).
 * Writes up to max_out cells as (x, y) pairs and returns the number of
 * cells in the walk (may exceed max_out). */
int oracle_dda(int fx, int fy, int tx, int ty, int32_t *out_xy, int max_out);

/* Deterministic sin/cos in double (Cody-Waite reduction by pi/2 and Taylor
 * polynomials); only IEEE +,-,*,/ so results are identical on any IEEE-754
 * machine compiled without contraction. */
void oracle_sincos(double x, double *s, double *c);

/* ---------------- moisture pre-pass (crates/dmoist/src/lib.rs) ------- */

/* One day of weather of a scenario: hourly temperature (C), relative
 * humidity (%), cloud cover (%), the rain (mm) of the six previous days
 * (index 0 = the day before) and the month (1..12). */
typedef struct oracle_day_weather {
  float temperature[24], humidity[24], cloud[24];
  float rain[6];
  int month;
} oracle_day_weather;

float oracle_rain_effect_max(int month, const float rain[6]);
int oracle_base_moisture(float temperature, float humidity);
float oracle_shading(float cloud, int fuel);
float oracle_ignition_probability(float temperature, float cloud, float moisture, int fuel);
float oracle_corrected_moisture(float base, float cloud, int month, float aspect_deg, float slope_deg, int fuel,
                                int hour);
float oracle_rain_corrected_probability(float probability, float effect);
float oracle_fine_moisture(float moisture, float prob_dry, float prob_wet);
float oracle_thicker_moisture(int hour, float m, float m6, float m15);
float oracle_live_moisture(int month);
/* Moisture ratios (d1hr, d10hr, d100hr, herb, wood) of n cells at one hour,
 * from per-cell fuel code, slope (rise/run) and aspect (rad). */
void oracle_moisture(size_t n, const uint8_t *fuel, const float *slope, const float *aspect,
                     const oracle_day_weather *w, int hour, float *d1hr, float *d10hr, float *d100hr, float *herb,
                     float *wood);

#ifdef __cplusplus
}
#endif
#endif
