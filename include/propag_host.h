/*
 * propag_host.h -- the plugin ABI of libpropag.so.
 *
 * Drop-in for the header the reference generates with cbindgen from
 * crates/propag/src/propag.rs (crates/propag/build.rs:20-33) and that the
 * QGIS plugin compiles against (qgis-plugin/propag-algo/src/plugincontainer.h):
 * the same type names, field order and #[repr(C)] layout, and the same
 * exported functions.  The plugin dlopen()s libpropag.so and looks up
 * FFIPropagation_run and propag_rasterize_fuels
 * (qgis-plugin/propag-algo/src/plugincontainer.cpp:9-26).
 *
 * Behind it, libpropag.so runs the MI355X path of libpropag_hip.so
 * (propag_hip.h): WKB rasterisation, the terrain callback, the burn pass and
 * the propagation on the GPU, then writes the outputs.
 */
#ifndef PROPAG_HOST_H
#define PROPAG_HOST_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include "geometry.h"
#include "propag_hip.h" /* Settings, FFITimeFeature, FFIFuelFeature */

#ifdef __cplusplus
extern "C" {
#endif

typedef float T; /* the reference's float::T, named in its generated header */

/* Stage F output of one cell (crates/propag/src/propag.rs:55-60). */
typedef struct FireSimpleCuda {
  T speed_max;   /* m/s  */
  T azimuth_max; /* rad  */
  T eccentricity;
} FireSimpleCuda;

/* The terrain columns a loader fills: W*H values each, row-major in the
 * layout of the GeoReference it is given (propag.rs:763-775). */
typedef struct FFITerrain {
  uint8_t *fuel_code;
  T *d1hr;
  T *d10hr;
  T *d100hr;
  T *herb;
  T *wood;
  T *wind_speed;
  T *wind_azimuth;
  T *slope;
  T *aspect;
} FFITerrain;

/* Terrain callback (propag.rs:777-783): called once per run with the run's
 * GeoReference and columns the library allocated and set to zero (fuel
 * code 0 = no fuel); returns false to abort the run.  The columns are lent
 * for the duration of the call. */
typedef bool (*LoadFn)(void *context, const GeoReference *geo_ref, FFITerrain *terrain);

typedef struct FFITerrainLoader {
  LoadFn load;
  void *context;
} FFITerrainLoader;

/* One run (propag.rs:177-188).  output_path is required (a GeoTIFF of
 * arrival times is written there); the other paths are optional (NULL or
 * "" = not written): refs_output_path a GPKG layer "fire_references" of
 * lines from each burned cell's reference to the cell, the boundary paths
 * GPKG layers "blocks" / "grids" of the launch decomposition.  The
 * ignitions are WKB in the CRS named by initial_ignited_elements_crs
 * (proj4), which must be the raster's own CRS (this library links no
 * reprojection); features whose WKB does not parse are skipped. */
typedef struct FFIPropagation {
  Settings settings;
  const char *output_path;
  const char *refs_output_path;
  const char *block_boundaries_out_path;
  const char *grid_boundaries_out_path;
  const FFITimeFeature *initial_ignited_elements;
  uintptr_t initial_ignited_elements_len;
  const char *initial_ignited_elements_crs;
  FFITerrainLoader terrain_loader;
} FFIPropagation;

typedef bool (*RunFn)(FFIPropagation propag, uintptr_t err_len, char *err_msg);
typedef bool (*RasterizeFuelsFn)(const FFIFuelFeature *fuels, uintptr_t fuels_len,
                                 const char *fuels_crs, const GeoReference *geo_ref,
                                 uint8_t *result, char *err_msg, uintptr_t err_len);

typedef struct PluginT {
  RunFn run;
  RasterizeFuelsFn rasterize_fuels;
} PluginT;

/* Runs one propagation and writes its outputs (propag.rs:274-309).  Returns
 * false on any error, with the message copied into err_msg (at most
 * err_len - 1 bytes, always NUL-terminated when err_len > 0).  Synchronous;
 * unlike the reference it may be called any number of times per process
 * (no MPI initialisation). */
bool FFIPropagation_run(FFIPropagation propag, uintptr_t err_len, char *err_msg);

/* Burns fuel features into result (W*H bytes, caller-owned; 255 where no
 * feature burned) (propag.rs:893-927).  fuels_crs must name the raster's
 * CRS.  Same error convention as FFIPropagation_run. */
bool propag_rasterize_fuels(const FFIFuelFeature *fuels, uintptr_t fuels_len,
                            const char *fuels_crs, const GeoReference *geo_ref,
                            uint8_t *result, char *err_msg, uintptr_t err_len);

/* Constructors exported for the header's users (propag.rs:45-51, 68-81,
 * 803-811). */
Settings create_settings(GeoReference geo_ref, float max_time);
FireSimpleCuda create_fire(T speed_max, T azimuth_max, T eccentricity);
PluginT create_plugin(void);

/* The output writers FFIPropagation_run uses (pg_write_times_geotiff,
 * pg_write_refs_gpkg, pg_write_tiles_gpkg) are declared in propag_hip.h. */

#ifdef __cplusplus
}
#endif
#endif
