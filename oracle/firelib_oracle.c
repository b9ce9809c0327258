/*
 * Oracle (test infrastructure only): Rothermel/BEHAVE fire behaviour per
 * cell, restated from crates/firelib/src/firelib.rs and the NFFL catalog of
 * crates/firelib/src/lib.rs:125-268.  See oracle.h for who may use this.
 *
 * f32 path.  The reference builds its catalog at compile time
 * (firelib.rs:554-979) with f32 +,-,*,/ and with exp/powf evaluated in f64
 * soft-float then rounded to f32 (crates/firelib/src/f32.rs:23-34).  Every
 * quantity also takes the SI round trip of crates/firelib/src/units.rs:26-115
 * (imperial -> SI in ParticleDef::standard, SI -> imperial in Particle::make),
 * which this file repeats so the f32 constants carry the same rounding.
 * Runtime per-cell arithmetic follows firelib.rs:1016-1188 in f32 with libm
 * powf/cosf/sinf/asinf/sqrtf.
 *
 * f64 path.  The same equations in double with the catalog derived directly
 * from the imperial table (what fireLib.c does), used to state the f32
 * error and to compare against fireLib.c.
 *
 * Conventions (TerrainCuda, firelib.rs:95-124): wind speed in m/s converted
 * to ft/min by dividing by 0.00508; wind azimuth, aspect in radians
 * (compass, the direction the wind blows toward / the downslope direction);
 * output speed_max in m/s (ft/min * 0.00508), azimuth_max in rad.
 *
 * Pinned by: tests/test_oracle_firelib.py (fireLib.c FP64 golden vectors,
 * closed-form situations, f32 vs f64 error bound).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXP 5
#define NFUELS 14
#define SMIDGEN_F 1e-6f
#define SMIDGEN_D 1e-6

enum { DEAD = 0, HERB = 1, WOOD = 2, NOPART = 3 };

/* The NFFL table as lib.rs:126-266 states it (depth ft, mext, particles of
 * (type, load lb/ft2, savr 1/ft)); NFFL13's 10-h load is 0.0580 there.
 * Each decimal literal is kept twice, rounded once to float and once to
 * double, so neither path inherits the other's rounding. */
typedef struct {
  float depth, mext;
  double depth_d, mext_d;
  int np;
  struct { int type; float load, savr; double load_d; } p[MAXP];
} fuel_def;

#define FUEL(depth, mext, np) depth##f, mext##f, depth, mext, np
#define PART(t, load, savr) {t, load##f, savr##f, load}

static const fuel_def TABLE[NFUELS] = {
    {FUEL(0.1, 0.01, 0), {{0}}},
    {FUEL(1.0, 0.12, 1), {PART(DEAD, 0.0340, 3500.0)}},
    {FUEL(1.0, 0.15, 4),
     {PART(DEAD, 0.0920, 3000.0), PART(DEAD, 0.0460, 109.0),
      PART(DEAD, 0.0230, 30.0), PART(HERB, 0.0230, 1500.0)}},
    {FUEL(2.5, 0.25, 1), {PART(DEAD, 0.1380, 1500.0)}},
    {FUEL(6.0, 0.2, 4),
     {PART(DEAD, 0.2300, 2000.0), PART(DEAD, 0.1840, 109.0),
      PART(DEAD, 0.0920, 30.0), PART(WOOD, 0.2300, 1500.0)}},
    {FUEL(2.0, 0.2, 3),
     {PART(DEAD, 0.0460, 2000.0), PART(DEAD, 0.0230, 109.0),
      PART(WOOD, 0.0920, 1500.0)}},
    {FUEL(2.5, 0.25, 3),
     {PART(DEAD, 0.0690, 1750.0), PART(DEAD, 0.1150, 109.0),
      PART(DEAD, 0.0920, 30.0)}},
    {FUEL(2.5, 0.40, 4),
     {PART(DEAD, 0.0520, 1750.0), PART(DEAD, 0.0860, 109.0),
      PART(DEAD, 0.0690, 30.0), PART(WOOD, 0.0170, 1550.0)}},
    {FUEL(0.2, 0.30, 3),
     {PART(DEAD, 0.0690, 2000.0), PART(DEAD, 0.0460, 109.0),
      PART(DEAD, 0.1150, 30.0)}},
    {FUEL(0.2, 0.25, 3),
     {PART(DEAD, 0.1340, 2500.0), PART(DEAD, 0.0190, 109.0),
      PART(DEAD, 0.0070, 30.0)}},
    {FUEL(1.0, 0.25, 4),
     {PART(DEAD, 0.1380, 2000.0), PART(DEAD, 0.0920, 109.0),
      PART(DEAD, 0.2300, 30.0), PART(WOOD, 0.0920, 1500.0)}},
    {FUEL(1.0, 0.15, 3),
     {PART(DEAD, 0.0690, 1500.0), PART(DEAD, 0.2070, 109.0),
      PART(DEAD, 0.2530, 30.0)}},
    {FUEL(2.3, 0.20, 3),
     {PART(DEAD, 0.1840, 1500.0), PART(DEAD, 0.6440, 109.0),
      PART(DEAD, 0.7590, 30.0)}},
    {FUEL(3.0, 0.25, 3),
     {PART(DEAD, 0.3220, 1500.0), PART(DEAD, 0.0580, 109.0),
      PART(DEAD, 1.2880, 30.0)}},
};

/* ------------------------------------------------------------------ f32 */

typedef struct {
  int type, size_class, mclass; /* mclass 0..4 = d1, d10, d100, herb, wood */
  float load, savr, density, heat, si_total, si_eff;
  float area_weight, size_class_weight, surface_area, sigma_factor;
} part32;

typedef struct {
  int np, n_dead, has_live, has_particles;
  part32 p[MAXP];
  float mext, inv_mext, live_area_weight, dead_area_weight, fine_dead_factor,
      live_ext_factor, bulk_density, residence_time, flux_ratio, slope_k,
      wind_b, wind_b_inv, wind_e, wind_k, sigma, beta, live_rx, dead_rx;
} fuel32;

static fuel32 CAT32[NFUELS];
static int CAT32_READY = 0;

static float expd_f(float x) { return (float)exp((double)x); }
static float powd_f(float x, float y) { return (float)pow((double)x, (double)y); }
static float safe_div_f(float a, float b) { return b > SMIDGEN_F ? a / b : 0.0f; }

/* units.rs: F = 0.3048, P = 0.4535924, BTU = 1054.35; powi(2) = F*F,
 * powi(3) = F*(F*F) (square-and-multiply). */
static const float FOOT_TO_M = 0.3048f, POUND_TO_KG = 0.4535924f,
                   BTU_ = 1054.35f;
static float load_factor(void) { return POUND_TO_KG / (FOOT_TO_M * FOOT_TO_M); }
static float density_factor(void) {
  float a2 = FOOT_TO_M * FOOT_TO_M;
  return POUND_TO_KG / (FOOT_TO_M * a2);
}
static float savr_factor(void) { return 1.0f / FOOT_TO_M; }
static float heat_factor(void) { return BTU_ / POUND_TO_KG; }

/* ParticleDef::standard (firelib.rs:557-573) then back to imperial
 * (Particle::make, firelib.rs:658-689). */
typedef struct { int type; float load, savr, density, heat; } pimp32;
static pimp32 particle_roundtrip(int type, float load, float savr) {
  pimp32 r;
  float load_si = load * load_factor();
  float savr_si = savr * savr_factor();
  float dens_si = 32.0f * density_factor();
  float heat_si = 8000.0f * heat_factor();
  r.type = type;
  r.load = load_si / load_factor();
  r.savr = savr_si / savr_factor();
  r.density = dens_si / density_factor();
  r.heat = heat_si / heat_factor();
  return r;
}

static int size_class32(float savr) { /* firelib.rs:577-592 */
  if (savr > 1200.0f) return 0;
  if (savr > 192.0f) return 1;
  if (savr > 96.0f) return 2;
  if (savr > 48.0f) return 3;
  if (savr > 16.0f) return 4;
  return 5;
}
static int same_life(int t, int life_dead) { return (t == DEAD) == life_dead; }
static float surface_area32(const pimp32 *p) { /* firelib.rs:643-648 */
  return safe_div_f(p->load * p->savr, p->density);
}

static void build_fuel32(int code) {
  const fuel_def *d = &TABLE[code];
  fuel32 *f = &CAT32[code];
  pimp32 defs[MAXP];
  memset(f, 0, sizeof(*f));
  for (int i = 0; i < d->np; i++)
    defs[i] = particle_roundtrip(d->p[i].type, d->p[i].load, d->p[i].savr);

  /* Particle::make in definition order; area weights are summed over the
   * definition order (firelib.rs:605-642). */
  part32 made[MAXP];
  for (int i = 0; i < d->np; i++) {
    const pimp32 *q = &defs[i];
    part32 *p = &made[i];
    int dead = q->type == DEAD;
    float total = 0.0f;
    for (int k = 0; k < d->np; k++)
      total += same_life(defs[k].type, dead) ? surface_area32(&defs[k]) : 0.0f;
    p->type = q->type;
    p->load = q->load;
    p->savr = q->savr;
    p->density = q->density;
    p->heat = q->heat;
    p->si_total = 0.0555f;
    p->si_eff = 0.01f;
    p->surface_area = surface_area32(q);
    p->area_weight = safe_div_f(p->surface_area, total);
    p->size_class = size_class32(q->savr);
    p->sigma_factor = expd_f(safe_div_f(-138.0f, q->savr));
  }
  for (int i = 0; i < d->np; i++) {
    int dead = made[i].type == DEAD;
    float s = 0.0f;
    for (int k = 0; k < d->np; k++) {
      float term = 0.0f;
      if (same_life(made[k].type, dead) && made[k].size_class == made[i].size_class)
        term = made[k].area_weight;
      s += term;
    }
    made[i].size_class_weight = s;
    if (made[i].type == DEAD)
      made[i].mclass = made[i].size_class <= 1 ? 0 : (made[i].size_class <= 3 ? 1 : 2);
    else
      made[i].mclass = made[i].type == HERB ? 3 : 4;
  }
  /* Sort: dead first, then alive, each in definition order. */
  int n = 0;
  for (int i = 0; i < d->np; i++)
    if (made[i].type == DEAD) f->p[n++] = made[i];
  f->n_dead = n;
  for (int i = 0; i < d->np; i++)
    if (made[i].type != DEAD) f->p[n++] = made[i];
  f->np = n;
  f->has_live = f->n_dead < n;
  f->has_particles = n > 0;

  /* FuelDef derivations (firelib.rs:716-868) over the sorted particles. */
  float depth_si = d->depth * FOOT_TO_M;
  float depth = depth_si / FOOT_TO_M;
  float total_area = 0.0f;
  for (int i = 0; i < n; i++) total_area += f->p[i].surface_area;
  float law[2] = {0.0f, 0.0f}, lsavr[2] = {0.0f, 0.0f}, lheat[2] = {0.0f, 0.0f},
        lseff[2] = {0.0f, 0.0f}, lload[2] = {0.0f, 0.0f}, lfine[2] = {0.0f, 0.0f};
  /* index 0 = alive, 1 = dead */
  for (int life = 0; life < 2; life++) {
    int dead = life == 1;
    for (int i = 0; i < n; i++) {
      const part32 *p = &f->p[i];
      int s = same_life(p->type, dead);
      law[life] += s ? p->surface_area / total_area : 0.0f;
      lsavr[life] += s ? p->area_weight * p->savr : 0.0f;
      lheat[life] += s ? p->area_weight * p->heat : 0.0f;
      lseff[life] += s ? p->area_weight * p->si_eff : 0.0f;
      lload[life] += s ? (p->size_class_weight * p->load) * (1.0f - p->si_total) : 0.0f;
      if (s) {
        if (dead)
          lfine[life] += p->load * p->sigma_factor;
        else
          lfine[life] += p->load * expd_f(-500.0f / p->savr);
      } else {
        lfine[life] += 0.0f;
      }
    }
  }
  float sigma = law[0] * lsavr[0] + law[1] * lsavr[1];
  float beta_acc = 0.0f;
  for (int i = 0; i < n; i++) beta_acc += f->p[i].load / f->p[i].density;
  float beta = beta_acc / depth;
  float ratio = beta / (3.348f / powd_f(sigma, 0.8189f));
  /* gamma (firelib.rs:815-824) */
  float sigma15 = powd_f(sigma, 1.5f);
  float gamma_max = sigma15 / (495.0f + 0.0594f * sigma15);
  float aa = 133.0f / powd_f(sigma, 0.7913f);
  float gamma = (gamma_max * powd_f(ratio, aa)) * expd_f(aa * (1.0f - ratio));
  float eta_s[2];
  for (int life = 0; life < 2; life++) {
    float seff = lseff[life];
    if (seff > SMIDGEN_F) {
      float eta = 0.174f / powd_f(seff, 0.19f);
      eta_s[life] = eta < 1.0f ? eta : 1.0f;
    } else {
      eta_s[life] = 1.0f;
    }
  }
  float rx[2];
  for (int life = 0; life < 2; life++)
    rx[life] = ((lload[life] * lheat[life]) * eta_s[life]) * gamma;
  /* wind_bke (firelib.rs:856-867) */
  float wind_b = 0.02526f * powd_f(sigma, 0.54f);
  float c = 7.47f * expd_f(-0.133f * powd_f(sigma, 0.55f));
  float e = 0.715f * expd_f(-0.000359f * sigma);
  float wind_k = c * powd_f(ratio, -e);
  float wind_e = powd_f(ratio, e) / c;
  float bulk = 0.0f;
  for (int i = 0; i < n; i++) bulk += f->p[i].load / depth;
  float flux = expd_f((0.792f + 0.681f * sqrtf(sigma)) * (beta + 0.1f)) /
               (192.0f + 0.2595f * sigma);

  f->mext = d->mext;
  f->inv_mext = 1.0f / d->mext;
  f->live_area_weight = law[0];
  f->dead_area_weight = law[1];
  f->live_rx = rx[0];
  f->dead_rx = rx[1];
  f->fine_dead_factor = 1.0f / lfine[1];
  f->live_ext_factor = 2.9f * safe_div_f(lfine[1], lfine[0]);
  f->bulk_density = bulk;
  f->residence_time = 384.0f / sigma;
  f->flux_ratio = flux;
  f->slope_k = 5.275f * powd_f(beta, -0.3f);
  f->wind_b = wind_b;
  f->wind_b_inv = 1.0f / wind_b;
  f->wind_e = wind_e;
  f->wind_k = wind_k;
  f->sigma = sigma;
  f->beta = beta;
}

static void ensure_cat32(void) {
  if (CAT32_READY) return;
  for (int i = 0; i < NFUELS; i++) build_fuel32(i);
  CAT32_READY = 1;
}

typedef struct { float speed_max, azimuth_max, eccentricity; } fire32;

static const float PI_F = 3.14159265358979323846f;
static const double PI_D = 3.14159265358979323846;

/* Fuel::eta_m, firelib.rs:1060-1076 */
static float eta_m32(const fuel32 *f, int alive, float life_moist, float wfmd) {
  float life_mext;
  if (!alive) {
    life_mext = f->mext;
  } else if (f->has_live) {
    float fdmois = wfmd * f->fine_dead_factor;
    float live_mext = f->live_ext_factor * (1.0f - fdmois * f->inv_mext) - 0.226f;
    life_mext = fmaxf(live_mext, f->mext);
  } else {
    life_mext = 0.0f;
  }
  if (life_moist >= life_mext) return 0.0f;
  float rt = life_moist / life_mext;
  return 1.0f - 2.59f * rt + 5.11f * rt * rt - 3.52f * rt * rt * rt;
}

static int burn32(int code, float d1, float d10, float d100, float herb,
                  float wood, float ws_ms, float waz, float slope,
                  float aspect, fire32 *out) {
  if (code < 0 || code >= NFUELS) return 0;
  const fuel32 *f = &CAT32[code];
  if (!f->has_particles) return 0;
  float moist[5] = {d1, d10, d100, herb, wood};
  /* rx_int_rbqig, firelib.rs:1031-1058 */
  float wfmd = 0.0f, alive_moist = 0.0f, dead_moist = 0.0f, rbqig = 0.0f;
  for (int i = 0; i < f->np; i++) {
    const part32 *p = &f->p[i];
    float m = moist[p->mclass];
    float law = p->type == DEAD ? f->dead_area_weight : f->live_area_weight;
    rbqig += (250.0f + 1116.0f * m) * p->area_weight * law * p->sigma_factor;
    if (p->type == DEAD) {
      dead_moist += p->area_weight * m;
      wfmd += m * p->sigma_factor * p->load;
    } else {
      alive_moist += p->area_weight * m;
    }
  }
  float rx_int = f->live_rx * eta_m32(f, 1, alive_moist, wfmd) +
                 f->dead_rx * eta_m32(f, 0, dead_moist, wfmd);
  float rbq = f->bulk_density * rbqig;
  float speed0 = rx_int * f->flux_ratio / rbq;

  /* calculate_wind_dependent_vars, firelib.rs:1078-1145 */
  float ws = ws_ms / 0.00508f; /* ft/min */
  float phi_slope = f->slope_k * slope * slope;
  float phi_wind = f->wind_k * powf(ws, f->wind_b);
  float phi_ew = phi_slope + phi_wind;
  float speed_max1 = speed0 * (1.0f + phi_ew);
  float upslope = aspect >= PI_F ? aspect - PI_F : aspect + PI_F;
  float max_wind = 0.9f * rx_int;
  float pew, ew, s, a;
  int limit = 1;
  if (speed0 < SMIDGEN_F) {
    pew = phi_ew; ew = 0.0f; s = 0.0f; a = 0.0f; limit = 0;
  } else if (phi_ew < SMIDGEN_F) {
    pew = phi_ew; ew = 0.0f; s = speed0; a = 0.0f; limit = 0;
  } else if (slope < SMIDGEN_F) {
    pew = phi_ew; ew = ws; s = speed_max1; a = waz;
  } else if (ws_ms < SMIDGEN_F || fabsf(upslope - waz) < SMIDGEN_F) {
    pew = phi_ew; ew = powf(phi_ew * f->wind_e, f->wind_b_inv); s = speed_max1; a = upslope;
  } else {
    float wind_rate = speed0 * phi_wind;
    float slope_rate = speed0 * phi_slope;
    float split = upslope <= waz ? waz - upslope : 2.0f * PI_F - upslope + waz;
    float x = slope_rate + wind_rate * cosf(split);
    float y = wind_rate * sinf(split);
    float rv = sqrtf(x * x + y * y);
    float speed_max2 = speed0 + rv;
    float phi_ew2 = speed_max2 / speed0 - 1.0f;
    float eff = powf(phi_ew2 * f->wind_e, f->wind_b_inv);
    float al = asinf(fabsf(y) / rv);
    float split2;
    if (x >= 0.0f && y >= 0.0f) split2 = al;
    else if (x >= 0.0f) split2 = 2.0f * PI_F - al;
    else if (y >= 0.0f) split2 = PI_F - al;
    else split2 = PI_F + al;
    float az2 = upslope + split2;
    if (az2 > 2.0f * PI_F) az2 = az2 - 2.0f * PI_F;
    pew = phi_ew2; ew = eff; s = speed_max2; a = az2;
  }
  (void)pew;
  if (limit && ew > max_wind) {
    float phi_max = f->wind_k * powf(max_wind, f->wind_b);
    s = speed0 * (1.0f + phi_max);
    ew = max_wind;
  }
  /* Fuel::eccentricity, firelib.rs:1171-1174 */
  float lw = 1.0f + 0.002840909f * ew;
  float ecc = sqrtf(lw * lw - 1.0f) / lw;
  out->speed_max = s * 0.00508f;
  out->azimuth_max = a;
  out->eccentricity = ecc;
  return 1;
}

size_t oracle_burn_f32(size_t n, const uint8_t *fuel, const float *d1hr,
                       const float *d10hr, const float *d100hr,
                       const float *herb, const float *wood,
                       const float *wind_speed, const float *wind_azimuth,
                       const float *slope, const float *aspect,
                       float *speed_max, float *azimuth_max,
                       float *eccentricity, int nthreads) {
  ensure_cat32();
  size_t written = 0;
  (void)nthreads;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) reduction(+ : written) num_threads(nthreads)
#endif
  for (long long i = 0; i < (long long)n; i++) {
    fire32 fr;
    if (burn32(fuel[i], d1hr[i], d10hr[i], d100hr[i], herb[i], wood[i],
               wind_speed[i], wind_azimuth[i], slope[i], aspect[i], &fr)) {
      speed_max[i] = fr.speed_max;
      azimuth_max[i] = fr.azimuth_max;
      eccentricity[i] = fr.eccentricity;
      written++;
    }
  }
  return written;
}

int oracle_catalog_f32(int code, float *out) {
  ensure_cat32();
  if (code < 0 || code >= NFUELS) return 0;
  const fuel32 *f = &CAT32[code];
  float v[16] = {f->sigma, f->beta, f->flux_ratio, f->slope_k, f->wind_b,
                 f->wind_e, f->wind_k, f->live_rx, f->dead_rx,
                 f->live_area_weight, f->dead_area_weight, f->fine_dead_factor,
                 f->live_ext_factor, f->bulk_density, f->residence_time,
                 f->mext};
  memcpy(out, v, sizeof(v));
  return f->np;
}

/* ------------------------------------------------------------------ f64 */

typedef struct {
  int type, size_class, mclass;
  double load, savr, density, heat, si_total, si_eff;
  double area_weight, size_class_weight, surface_area, sigma_factor;
} part64;

typedef struct {
  int np, n_dead, has_live, has_particles;
  part64 p[MAXP];
  double mext, live_area_weight, dead_area_weight, fine_dead_load,
      live_ext_factor, bulk_density, residence_time, flux_ratio, slope_k,
      wind_b, wind_e, wind_k, sigma, beta, live_rx, dead_rx;
} fuel64;

static fuel64 CAT64[NFUELS];
static int CAT64_READY = 0;

static double safe_div_d(double a, double b) { return b > SMIDGEN_D ? a / b : 0.0; }
static int size_class64(double savr) {
  if (savr > 1200.0) return 0;
  if (savr > 192.0) return 1;
  if (savr > 96.0) return 2;
  if (savr > 48.0) return 3;
  if (savr > 16.0) return 4;
  return 5;
}

static void build_fuel64(int code) {
  const fuel_def *d = &TABLE[code];
  fuel64 *f = &CAT64[code];
  memset(f, 0, sizeof(*f));
  part64 made[MAXP];
  for (int i = 0; i < d->np; i++) {
    part64 *p = &made[i];
    p->type = d->p[i].type;
    p->load = d->p[i].load_d;
    p->savr = (double)d->p[i].savr;
    p->density = 32.0;
    p->heat = 8000.0;
    p->si_total = 0.0555;
    p->si_eff = 0.01;
    p->surface_area = safe_div_d(p->load * p->savr, p->density);
    p->sigma_factor = exp(safe_div_d(-138.0, p->savr));
    p->size_class = size_class64(p->savr);
  }
  for (int i = 0; i < d->np; i++) {
    int dead = made[i].type == DEAD;
    double total = 0.0;
    for (int k = 0; k < d->np; k++)
      if (same_life(made[k].type, dead)) total += made[k].surface_area;
    made[i].area_weight = safe_div_d(made[i].surface_area, total);
  }
  for (int i = 0; i < d->np; i++) {
    int dead = made[i].type == DEAD;
    double s = 0.0;
    for (int k = 0; k < d->np; k++)
      if (same_life(made[k].type, dead) && made[k].size_class == made[i].size_class)
        s += made[k].area_weight;
    made[i].size_class_weight = s;
    if (made[i].type == DEAD)
      made[i].mclass = made[i].size_class <= 1 ? 0 : (made[i].size_class <= 3 ? 1 : 2);
    else
      made[i].mclass = made[i].type == HERB ? 3 : 4;
  }
  int n = 0;
  for (int i = 0; i < d->np; i++)
    if (made[i].type == DEAD) f->p[n++] = made[i];
  f->n_dead = n;
  for (int i = 0; i < d->np; i++)
    if (made[i].type != DEAD) f->p[n++] = made[i];
  f->np = n;
  f->has_live = f->n_dead < n;
  f->has_particles = n > 0;
  double depth = d->depth_d, mext = d->mext_d;

  double total_area = 0.0;
  for (int i = 0; i < n; i++) total_area += f->p[i].surface_area;
  double law[2] = {0}, lsavr[2] = {0}, lheat[2] = {0}, lseff[2] = {0},
         lload[2] = {0}, lfine[2] = {0};
  for (int life = 0; life < 2; life++) {
    int dead = life == 1;
    for (int i = 0; i < n; i++) {
      const part64 *p = &f->p[i];
      if (!same_life(p->type, dead)) continue;
      law[life] += p->surface_area / total_area;
      lsavr[life] += p->area_weight * p->savr;
      lheat[life] += p->area_weight * p->heat;
      lseff[life] += p->area_weight * p->si_eff;
      lload[life] += p->size_class_weight * p->load * (1.0 - p->si_total);
      lfine[life] += dead ? p->load * p->sigma_factor
                          : p->load * exp(-500.0 / p->savr);
    }
  }
  double sigma = law[0] * lsavr[0] + law[1] * lsavr[1];
  double beta = 0.0;
  for (int i = 0; i < n; i++) beta += f->p[i].load / f->p[i].density;
  beta /= depth;
  double ratio = beta / (3.348 / pow(sigma, 0.8189));
  double sigma15 = pow(sigma, 1.5);
  double gamma_max = sigma15 / (495.0 + 0.0594 * sigma15);
  double aa = 133.0 / pow(sigma, 0.7913);
  double gamma = gamma_max * pow(ratio, aa) * exp(aa * (1.0 - ratio));
  double eta_s[2];
  for (int life = 0; life < 2; life++) {
    double seff = lseff[life];
    eta_s[life] = 1.0;
    if (seff > SMIDGEN_D) {
      double eta = 0.174 / pow(seff, 0.19);
      eta_s[life] = eta < 1.0 ? eta : 1.0;
    }
  }
  double c = 7.47 * exp(-0.133 * pow(sigma, 0.55));
  double e = 0.715 * exp(-0.000359 * sigma);
  double bulk = 0.0;
  for (int i = 0; i < n; i++) bulk += f->p[i].load;
  f->mext = mext;
  f->live_area_weight = law[0];
  f->dead_area_weight = law[1];
  f->live_rx = lload[0] * lheat[0] * eta_s[0] * gamma;
  f->dead_rx = lload[1] * lheat[1] * eta_s[1] * gamma;
  f->fine_dead_load = lfine[1];
  f->live_ext_factor = 2.9 * safe_div_d(lfine[1], lfine[0]);
  f->bulk_density = bulk / depth;
  f->residence_time = 384.0 / sigma;
  f->flux_ratio = exp((0.792 + 0.681 * sqrt(sigma)) * (beta + 0.1)) /
                  (192.0 + 0.2595 * sigma);
  f->slope_k = 5.275 * pow(beta, -0.3);
  f->wind_b = 0.02526 * pow(sigma, 0.54);
  f->wind_k = c * pow(ratio, -e);
  f->wind_e = pow(ratio, e) / c;
  f->sigma = sigma;
  f->beta = beta;
}

static void ensure_cat64(void) {
  if (CAT64_READY) return;
  for (int i = 0; i < NFUELS; i++) build_fuel64(i);
  CAT64_READY = 1;
}

static double eta_m64(const fuel64 *f, int alive, double lm, double wfmd) {
  double life_mext;
  if (!alive) {
    life_mext = f->mext;
  } else if (f->has_live) {
    double fdmois = wfmd / f->fine_dead_load;
    double live_mext = f->live_ext_factor * (1.0 - fdmois / f->mext) - 0.226;
    life_mext = fmax(live_mext, f->mext);
  } else {
    life_mext = 0.0;
  }
  if (lm >= life_mext) return 0.0;
  double rt = lm / life_mext;
  return 1.0 - 2.59 * rt + 5.11 * rt * rt - 3.52 * rt * rt * rt;
}

static int burn64(int code, double d1, double d10, double d100, double herb,
                  double wood, double ws_ms, double waz, double slope,
                  double aspect, double out[3]) {
  if (code < 0 || code >= NFUELS) return 0;
  const fuel64 *f = &CAT64[code];
  if (!f->has_particles) return 0;
  double moist[5] = {d1, d10, d100, herb, wood};
  double wfmd = 0.0, alive_moist = 0.0, dead_moist = 0.0, rbqig = 0.0;
  for (int i = 0; i < f->np; i++) {
    const part64 *p = &f->p[i];
    double m = moist[p->mclass];
    double law = p->type == DEAD ? f->dead_area_weight : f->live_area_weight;
    rbqig += (250.0 + 1116.0 * m) * p->area_weight * law * p->sigma_factor;
    if (p->type == DEAD) {
      dead_moist += p->area_weight * m;
      wfmd += m * p->sigma_factor * p->load;
    } else {
      alive_moist += p->area_weight * m;
    }
  }
  double rx_int = f->live_rx * eta_m64(f, 1, alive_moist, wfmd) +
                  f->dead_rx * eta_m64(f, 0, dead_moist, wfmd);
  double speed0 = rx_int * f->flux_ratio / (f->bulk_density * rbqig);
  double ws = ws_ms / 0.00508;
  double phi_slope = f->slope_k * slope * slope;
  double phi_wind = f->wind_k * pow(ws, f->wind_b);
  double phi_ew = phi_slope + phi_wind;
  double speed_max1 = speed0 * (1.0 + phi_ew);
  double upslope = aspect >= PI_D ? aspect - PI_D : aspect + PI_D;
  double max_wind = 0.9 * rx_int;
  double ew, s, a;
  int limit = 1;
  if (speed0 < SMIDGEN_D) {
    ew = 0.0; s = 0.0; a = 0.0; limit = 0;
  } else if (phi_ew < SMIDGEN_D) {
    ew = 0.0; s = speed0; a = 0.0; limit = 0;
  } else if (slope < SMIDGEN_D) {
    ew = ws; s = speed_max1; a = waz;
  } else if (ws_ms < SMIDGEN_D || fabs(upslope - waz) < SMIDGEN_D) {
    ew = pow(phi_ew * f->wind_e, 1.0 / f->wind_b); s = speed_max1; a = upslope;
  } else {
    double wind_rate = speed0 * phi_wind;
    double slope_rate = speed0 * phi_slope;
    double split = upslope <= waz ? waz - upslope : 2.0 * PI_D - upslope + waz;
    double x = slope_rate + wind_rate * cos(split);
    double y = wind_rate * sin(split);
    double rv = sqrt(x * x + y * y);
    double speed_max2 = speed0 + rv;
    double phi_ew2 = speed_max2 / speed0 - 1.0;
    double al = asin(fabs(y) / rv);
    double split2;
    if (x >= 0.0 && y >= 0.0) split2 = al;
    else if (x >= 0.0) split2 = 2.0 * PI_D - al;
    else if (y >= 0.0) split2 = PI_D - al;
    else split2 = PI_D + al;
    double az2 = upslope + split2;
    if (az2 > 2.0 * PI_D) az2 -= 2.0 * PI_D;
    ew = pow(phi_ew2 * f->wind_e, 1.0 / f->wind_b); s = speed_max2; a = az2;
  }
  if (limit && ew > max_wind) {
    s = speed0 * (1.0 + f->wind_k * pow(max_wind, f->wind_b));
    ew = max_wind;
  }
  double lw = 1.0 + 0.002840909 * ew;
  out[0] = s * 0.00508;
  out[1] = a;
  out[2] = sqrt(lw * lw - 1.0) / lw;
  return 1;
}

size_t oracle_burn_f64(size_t n, const uint8_t *fuel, const float *d1hr,
                       const float *d10hr, const float *d100hr,
                       const float *herb, const float *wood,
                       const float *wind_speed, const float *wind_azimuth,
                       const float *slope, const float *aspect,
                       double *speed_max, double *azimuth_max,
                       double *eccentricity) {
  ensure_cat64();
  size_t written = 0;
  for (size_t i = 0; i < n; i++) {
    double o[3];
    if (burn64(fuel[i], d1hr[i], d10hr[i], d100hr[i], herb[i], wood[i],
               wind_speed[i], wind_azimuth[i], slope[i], aspect[i], o)) {
      speed_max[i] = o[0];
      azimuth_max[i] = o[1];
      eccentricity[i] = o[2];
      written++;
    }
  }
  return written;
}

int oracle_catalog_f64(int code, double *out) {
  ensure_cat64();
  if (code < 0 || code >= NFUELS) return 0;
  const fuel64 *f = &CAT64[code];
  double v[16] = {f->sigma, f->beta, f->flux_ratio, f->slope_k, f->wind_b,
                  f->wind_e, f->wind_k, f->live_rx, f->dead_rx,
                  f->live_area_weight, f->dead_area_weight,
                  1.0 / f->fine_dead_load, f->live_ext_factor,
                  f->bulk_density, f->residence_time, f->mext};
  memcpy(out, v, sizeof(v));
  return f->np;
}
