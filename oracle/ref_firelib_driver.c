/*
 * Driver for the reference's own C fireLib (crates/firelib-sys/fireLib.c,
 * FP64 BEHAVE), compiled from /root/reference into oracle/_ref/ by
 * oracle/Makefile.  TEST INFRASTRUCTURE ONLY: it produces the FP64 golden
 * vectors of tests/golden/ and is never part of the product.
 *
 * It builds the catalog through fireLib's public API
 * (Fire_FuelCatalogCreate / Fire_FuelModelCreate / Fire_FuelParticleAdd) with
 * the particle table of crates/firelib/src/lib.rs:125-268, the one the
 * reference's GPU path uses (NFFL13 10-h load 0.0580, not fireLib's 1.0580),
 * and evaluates cells the way the reference's quickcheck test does
 * (crates/firelib/src/lib.rs:312-367): Fire_SpreadNoWindNoSlope then
 * Fire_SpreadWindSlopeMax, with wind in ft/min and angles in degrees.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "fireLib.h"

typedef struct {
  double depth, mext;
  size_t np;
  struct { size_t type; double load, savr; } p[4];
} model_def;

static const model_def MODELS[14] = {
    {0.1, 0.01, 0, {{0, 0, 0}}},
    {1.0, 0.12, 1, {{FIRE_TYPE_DEAD, 0.0340, 3500.0}}},
    {1.0, 0.15, 4,
     {{FIRE_TYPE_DEAD, 0.0920, 3000.0}, {FIRE_TYPE_DEAD, 0.0460, 109.0},
      {FIRE_TYPE_DEAD, 0.0230, 30.0}, {FIRE_TYPE_HERB, 0.0230, 1500.0}}},
    {2.5, 0.25, 1, {{FIRE_TYPE_DEAD, 0.1380, 1500.0}}},
    {6.0, 0.2, 4,
     {{FIRE_TYPE_DEAD, 0.2300, 2000.0}, {FIRE_TYPE_DEAD, 0.1840, 109.0},
      {FIRE_TYPE_DEAD, 0.0920, 30.0}, {FIRE_TYPE_WOOD, 0.2300, 1500.0}}},
    {2.0, 0.2, 3,
     {{FIRE_TYPE_DEAD, 0.0460, 2000.0}, {FIRE_TYPE_DEAD, 0.0230, 109.0},
      {FIRE_TYPE_WOOD, 0.0920, 1500.0}}},
    {2.5, 0.25, 3,
     {{FIRE_TYPE_DEAD, 0.0690, 1750.0}, {FIRE_TYPE_DEAD, 0.1150, 109.0},
      {FIRE_TYPE_DEAD, 0.0920, 30.0}}},
    {2.5, 0.40, 4,
     {{FIRE_TYPE_DEAD, 0.0520, 1750.0}, {FIRE_TYPE_DEAD, 0.0860, 109.0},
      {FIRE_TYPE_DEAD, 0.0690, 30.0}, {FIRE_TYPE_WOOD, 0.0170, 1550.0}}},
    {0.2, 0.30, 3,
     {{FIRE_TYPE_DEAD, 0.0690, 2000.0}, {FIRE_TYPE_DEAD, 0.0460, 109.0},
      {FIRE_TYPE_DEAD, 0.1150, 30.0}}},
    {0.2, 0.25, 3,
     {{FIRE_TYPE_DEAD, 0.1340, 2500.0}, {FIRE_TYPE_DEAD, 0.0190, 109.0},
      {FIRE_TYPE_DEAD, 0.0070, 30.0}}},
    {1.0, 0.25, 4,
     {{FIRE_TYPE_DEAD, 0.1380, 2000.0}, {FIRE_TYPE_DEAD, 0.0920, 109.0},
      {FIRE_TYPE_DEAD, 0.2300, 30.0}, {FIRE_TYPE_WOOD, 0.0920, 1500.0}}},
    {1.0, 0.15, 3,
     {{FIRE_TYPE_DEAD, 0.0690, 1500.0}, {FIRE_TYPE_DEAD, 0.2070, 109.0},
      {FIRE_TYPE_DEAD, 0.2530, 30.0}}},
    {2.3, 0.20, 3,
     {{FIRE_TYPE_DEAD, 0.1840, 1500.0}, {FIRE_TYPE_DEAD, 0.6440, 109.0},
      {FIRE_TYPE_DEAD, 0.7590, 30.0}}},
    {3.0, 0.25, 3,
     {{FIRE_TYPE_DEAD, 0.3220, 1500.0}, {FIRE_TYPE_DEAD, 0.0580, 109.0},
      {FIRE_TYPE_DEAD, 1.2880, 30.0}}},
};

static FuelCatalogPtr make_catalog(void) {
  static char name[] = "rust-table";
  FuelCatalogPtr cat = Fire_FuelCatalogCreate(name, 13);
  if (!cat) return NULL;
  for (size_t m = 0; m < 14; m++) {
    char nm[16], desc[16];
    snprintf(nm, sizeof nm, "M%02zu", m);
    snprintf(desc, sizeof desc, "model %zu", m);
    if (Fire_FuelModelCreate(cat, m, nm, desc, MODELS[m].depth, MODELS[m].mext,
                             1.0, MODELS[m].np) != FIRE_STATUS_OK)
      return NULL;
    for (size_t p = 0; p < MODELS[m].np; p++)
      if (Fire_FuelParticleAdd(cat, m, MODELS[m].p[p].type, MODELS[m].p[p].load,
                               MODELS[m].p[p].savr, 32.0, 8000.0, 0.0555,
                               0.0100) != FIRE_STATUS_OK)
        return NULL;
  }
  return cat;
}

/* Evaluates n cells.  Inputs follow TerrainCuda units (m/s, radians);
 * outputs are spreadMax in m/s, azimuthMax in radians, eccentricity, and
 * spread0 / rxInt in the fireLib units (ft/min, BTU/ft2/min).  Returns the
 * number of cells evaluated or -1 on a catalog error. */
long ref_firelib_eval(size_t n, const uint8_t *fuel, const double *d1,
                      const double *d10, const double *d100,
                      const double *herb, const double *wood,
                      const double *wind_ms, const double *wind_az_rad,
                      const double *slope, const double *aspect_rad,
                      double *speed_ms, double *az_rad, double *ecc,
                      double *spread0, double *rx_int) {
  FuelCatalogPtr cat = make_catalog();
  if (!cat) return -1;
  for (size_t i = 0; i < n; i++) {
    size_t m = fuel[i];
    if (m > 13) {
      speed_ms[i] = az_rad[i] = ecc[i] = spread0[i] = rx_int[i] = 0.0;
      continue;
    }
    double moist[FIRE_MCLASSES] = {d1[i], d10[i], d100[i], 0.0, herb[i], wood[i]};
    /* force a full recomputation for every cell (no input caching) */
    Fuel_CombustionFlag(cat, m) = 0;
    Fire_SpreadNoWindNoSlope(cat, m, moist);
    Fire_SpreadWindSlopeMax(cat, m, wind_ms[i] / 0.00508,
                            wind_az_rad[i] * 57.29577951308232, slope[i],
                            aspect_rad[i] * 57.29577951308232);
    speed_ms[i] = Fuel_SpreadMax(cat, m) * 0.00508;
    az_rad[i] = Fuel_AzimuthMax(cat, m) * 1.7453292519943295e-2;
    ecc[i] = Fuel_Eccentricity(cat, m);
    spread0[i] = Fuel_Spread0(cat, m);
    rx_int[i] = Fuel_RxIntensity(cat, m);
  }
  Fire_FuelCatalogDestroy(cat);
  return (long)n;
}
