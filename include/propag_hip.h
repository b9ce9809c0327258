/*
 * propag_hip.h -- C ABI of the MI355X (gfx950) implementation of propag25's
 * two GPU stages: the firelib Rothermel/BEHAVE burn pass and the arrival-time
 * propagation.  Plain pointers and sizes only; every entry point returns a
 * pg_status and leaves a message for pg_last_error() on failure.
 *
 * The layouts of GT / GeoTransform / GeoReference (geometry.h), Settings
 * and the feature records are identical to the reference's #[repr(C)] types
 * so a caller can pass the same bytes:
 *   GT, GeoTransform, GeoReference   crates/geometry/src/lib.rs:20-29,128-133,180-187
 *   Settings                         crates/propag/src/propag.rs:38-43
 *   FFITimeFeature, FFIFuelFeature   crates/propag/src/propag.rs:120-125,150-155
 *
 * Device pointers are HIP device allocations on the current device; streams
 * are hipStream_t passed as void* (NULL = the default stream).
 */
#ifndef PROPAG_HIP_H
#define PROPAG_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "geometry.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

typedef struct Settings {
  GeoReference geo_ref;
  float max_time;
} Settings;

typedef enum pg_status {
  PG_OK = 0,
  PG_ERR_INVALID = 1,     /* bad argument (sizes, null pointers, layout)   */
  PG_ERR_HIP = 2,         /* a HIP runtime call failed                      */
  PG_ERR_NO_DEVICE = 3,   /* no usable gfx950 device                        */
  PG_ERR_NOT_CONVERGED = 4, /* the iteration cap was reached                */
  PG_ERR_OUT_OF_MEMORY = 5, /* device allocation or admission refused       */
  PG_ERR_TIMEOUT = 6,     /* a device-side wait exceeded its spin limit     */
  PG_ERR_BUSY = 7,        /* serving queue full / slot unavailable          */
  PG_ERR_IO = 8           /* an output file could not be written            */
} pg_status;

/* Propagation schedule knobs.  Zero fields take the library defaults. */
typedef struct pg_tuning {
  uint32_t tile_w, tile_h;      /* cells per tile: 8 x 8, 16 x 4 or 32 x 2, one
                                   cell per lane of a wave (0 x 0 = 8 x 8).  The
                                   handle API takes the context's shape
                                   (pg_ctx_set_tile); 0 x 0 or that shape.  */
  uint32_t workgroups;          /* persistent workgroups (0 = fill the device) */
  uint32_t max_iterations;      /* Jacobi sweep cap (0 = 64 * (W + H) + 1024)  */
  uint32_t mode;                /* 0 = persistent kernel, 1 = one launch per sweep,
                                   2 = persistent kernel through a cached hipGraph */
  uint32_t spin_limit_log2;     /* device wait limit, 2^n polls (0 = 22; 32 or
                                   more = no wait: a grid barrier that finds a
                                   workgroup not yet arrived fails the launch
                                   with PG_ERR_TIMEOUT, for tests of that path) */
  uint32_t waves_per_group;     /* persistent kernel waves per workgroup (0 = auto) */
  uint32_t profile;             /* 1: time the sweep phases (fills pg_stats.phase_ms) */
} pg_tuning;

/* Counters of one propagation (device-side counts copied back). */
typedef struct pg_stats {
  uint64_t iterations;      /* Jacobi sweeps, including the final quiet one */
  uint64_t relaxations;     /* cells whose update rule was evaluated        */
  uint64_t improvements;    /* cell updates committed                       */
  uint64_t dda_steps;       /* DDA cells visited by boundary walks          */
  uint64_t tiles_visited;   /* tile visits (live tiles summed over sweeps)  */
  double   elapsed_ms;      /* device time of the whole call (HIP events)   */
  double   relax_ms;        /* device time of the relaxation kernel(s) only */
  double   work_ms;         /* mean per workgroup: time in tile work          */
  double   wait_ms;         /* mean per workgroup: time waiting at barriers   */
  double   crit_ms;         /* sum over sweeps of the slowest workgroup's work */
  uint32_t workgroups;      /* workgroups of the relaxation launch            */
  uint32_t lds_bitset;      /* 1 if the DDA bitset was staged in LDS          */
  double   phase_ms[4];     /* with tuning.profile: sum over sweeps of the
                               slowest wave's time in halo staging, update
                               rule, line-of-sight walks, commit + bookkeeping */
  double   barrier_ms[4];   /* with tuning.profile: sum over sweeps of the time
                               from a barrier release until the last workgroup
                               started, the longest workgroup span, last
                               arrival until the master had seen all arrivals,
                               and the master's bookkeeping                   */
  uint64_t tile_marks;      /* tiles queued for a next sweep (with repeats)   */
  double   phase_total_ms[4]; /* with tuning.profile: the phases of phase_ms
                                 summed over all waves and sweeps (wave-ms)  */
  uint64_t idle_tiles;      /* tile visits in which no cell needed the rule   */
  uint64_t idle_far_tiles;  /* with tuning.profile: idle visits with no cell
                               changed within two cells of the tile either
                               (queued through a far-dependents box only)  */
} pg_stats;

/* ------------------------------------------------------------ lifecycle */

/* Human-readable message of the last failing call on this thread. */
const char *pg_last_error(void);

/* Version string of the library build. */
const char *pg_version(void);

/* Releases every device resource the library holds outside batching
 * contexts: the propagation workspaces, the cached launch graphs with their
 * events and pinned control blocks, and the device fuel catalog, on every
 * device used so far; a later call recreates what it needs.  Call it with
 * no call of the library in flight (contexts are released by
 * pg_ctx_destroy).  It also runs by itself at process exit, before the HIP
 * runtime's own exit handlers.  No counterpart in the reference (its cust
 * objects are dropped by Rust). */
pg_status pg_shutdown(void);

/* ---------------------------------------------------------------- Stage F */

/* Host-only diagnostic (no GPU needed): writes the f32 catalog constants of
 * fuel `code` (0..13) that the burn kernel uses, in this order: sigma, beta,
 * flux_ratio, slope_k, wind_b, wind_e, wind_k, live_rx_factor,
 * dead_rx_factor, live_area_weight, dead_area_weight, fine_dead_factor,
 * live_ext_factor, bulk_density, residence_time, mext.  Returns the number
 * of particles, 0 for a code without fuel, -1 for a code out of range.
 * Exposes Catalog::STANDARD (crates/firelib/src/lib.rs:125-268). */
int pg_catalog_f32(int code, float out[16]);

/* Per-cell Rothermel/BEHAVE burn over n cells of SoA terrain (TerrainCuda,
 * crates/firelib/src/firelib.rs:95-124: moistures as ratios, wind speed in
 * m/s, wind azimuth / aspect in radians).  Writes speed_max (m/s),
 * azimuth_max (rad) and eccentricity only where the fuel code is a catalog
 * fuel with particles; the caller pre-zeroes the outputs.
 * Replaces the kernel cuda_standard_simple_burn
 * (crates/firelib/src/cuda.rs:57-95) and its launch in
 * crates/propag/src/propag.rs:458-480.  All pointers are device pointers. */
pg_status pg_burn(size_t n, const uint8_t *fuel, const float *d1hr,
                  const float *d10hr, const float *d100hr, const float *herb,
                  const float *wood, const float *wind_speed,
                  const float *wind_azimuth, const float *slope,
                  const float *aspect, float *speed_max, float *azimuth_max,
                  float *eccentricity, void *stream);

/* ------------------------------------------------- moisture pre-pass */

/* One day of weather of a scenario: hourly temperature (C), relative
 * humidity (%) and cloud cover (%), the rain (mm) of the six previous days
 * (index 0 = the day before) and the month (1..12). */
typedef struct pg_day_weather {
  float temperature[24];
  float humidity[24];
  float cloud_cover[24];
  float rain_6_days[6];
  int32_t month;
} pg_day_weather;

/* Dead and live fuel moisture (ratios) of n cells at one hour (0..23) of a
 * day: writes d1hr, d10hr, d100hr, herb and wood for the burn pass from the
 * per-cell fuel code, slope (rise/run) and aspect (rad).  Replaces
 * crates/firelib/src/fuel_moisture.rs:26-63 (create_terrain_with_fuel_moisture)
 * and the dmoist model it calls (crates/dmoist/src/lib.rs:55-145), per cell
 * on the device.  All arrays are device pointers. */
pg_status pg_moisture(size_t n, const uint8_t *fuel, const float *slope, const float *aspect,
                      const pg_day_weather *weather, int hour, float *d1hr, float *d10hr,
                      float *d100hr, float *herb, float *wood, void *stream);

/* ---------------------------------------------------------------- Stage P */

/* Boundaries, relaxation to convergence and post-pass over one raster.
 * Replaces the kernels pre_propagate / propag / post_propagate
 * (crates/propag/src/propag.cu:382-421) and the host loop of
 * crates/propag/src/propag.rs:482-557.
 *   time, refs_x, refs_y  in-out (initial state as propag.rs:433-449 builds it:
 *                         ignited cells hold their time and point at
 *                         themselves, other cells FLT_MAX and 0xFFFF);
 *   boundaries            out, ceil(W*H/32) words (propag.cu:9-13 bit order).
 * On return time holds FLT_MAX where the fire is null or the arrival is
 * later than max_time (propag.cu:110-117).  All arrays are device pointers;
 * the call synchronises the stream.  tuning and stats may be NULL. */
pg_status pg_propagate(const Settings *settings, const float *speed_max,
                       const float *azimuth_max, const float *eccentricity,
                       float *time, uint16_t *refs_x, uint16_t *refs_y,
                       uint32_t *boundaries, const pg_tuning *tuning,
                       pg_stats *stats, void *stream);

/* Allocation-free form of pg_propagate, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  It has no counterpart in the reference.  Same
 * arguments, results and errors as pg_propagate, but every buffer beyond the
 * caller's inputs and outputs comes from a fixed device array of the library
 * (44 MiB, held on each device the library runs on until the process ends),
 * and the call makes no hipMalloc/hipFree, creates no events and caches no
 * graph.  Limits: at most PG_NOALLOC_MAX_CELLS cells (larger rasters return
 * PG_ERR_INVALID), tuning->mode 0 or 1 (2 and profile = 1 return
 * PG_ERR_INVALID), stats->elapsed_ms and relax_ms are reported as 0.  Calls
 * on one device share the array: they are serialised, and each synchronises
 * its stream before returning. */
#define PG_NOALLOC_MAX_CELLS 1048576u
pg_status pg_propagate_noalloc(const Settings *settings, const float *speed_max,
                               const float *azimuth_max, const float *eccentricity,
                               float *time, uint16_t *refs_x, uint16_t *refs_y,
                               uint32_t *boundaries, const pg_tuning *tuning,
                               pg_stats *stats, void *stream);

/* Host-buffer form of the whole device data plane of propagate()
 * (crates/propag/src/propag.rs:375-573): uploads the SoA terrain, runs the
 * burn pass then pg_propagate, copies time / refs / boundaries back.  The
 * terrain columns are W*H host arrays; time/refs_x/refs_y are in-out host
 * arrays; boundaries (out) has ceil(W*H/32) words.  fire_out, if not NULL,
 * receives the 3*W*H burn outputs (speed_max | azimuth_max | eccentricity). */
pg_status pg_propagate_host(const Settings *settings, const uint8_t *fuel,
                            const float *d1hr, const float *d10hr,
                            const float *d100hr, const float *herb,
                            const float *wood, const float *wind_speed,
                            const float *wind_azimuth, const float *slope,
                            const float *aspect, float *time,
                            uint16_t *refs_x, uint16_t *refs_y,
                            uint32_t *boundaries, float *fire_out,
                            const pg_tuning *tuning, pg_stats *stats);

/* ------------------------------------- initial state (host, no GPU) */

/* An ignited feature: a time (s) and its geometry as WKB (caller-owned). */
typedef struct FFITimeFeature {
  float time;
  const uint8_t *geom_wkb;
  uintptr_t geom_wkb_len;
} FFITimeFeature;

/* A fuel feature: an NFFL fuel code and its geometry as WKB. */
typedef struct FFIFuelFeature {
  uint8_t code;
  const uint8_t *geom_wkb;
  uintptr_t geom_wkb_len;
} FFIFuelFeature;

/* Burns n ignition features into time (W*H host array, row-major in the
 * layout of geo): the raster is filled with FLT_MAX, then every feature
 * burns its time, later features replacing earlier ones, by GDAL's default
 * rules (points: the pixel containing them; lines: a Bresenham line between
 * the pixels of consecutive vertices; polygons and collections: the pixels
 * whose centres are inside, scan line by scan line).  Geometries are WKB
 * (OGC or ISO, 2-D/Z/M/ZM, either byte order) in the raster's coordinates;
 * a feature whose WKB does not parse is skipped, as the reference does.
 * *burned (may be NULL) receives the number of features that parsed.
 * Replaces rasterize_times (crates/propag/src/propag.rs:837-890), which
 * calls GDALRasterizeGeometries.  Host code; needs no GPU. */
pg_status pg_rasterize_times(const GeoReference *geo, const FFITimeFeature *features,
                             size_t n, float *time, uint32_t *burned);

/* The same for fuel features into a W*H u8 raster pre-filled with 255:
 * replaces rasterize_fuels (crates/propag/src/propag.rs:929-978). */
pg_status pg_rasterize_fuels(const GeoReference *geo, const FFIFuelFeature *features,
                             size_t n, uint8_t *fuel, uint32_t *burned);

/* Transforms n points in place from one proj4 CRS to another: geographic
 * (longitude, latitude in degrees) and transverse Mercator (+proj=utm,
 * +proj=tmerc) on the common ellipsoids, x = east, y = north.  ok[i] is 0
 * (and the point unchanged) where a point lies outside the target
 * projection's domain.  Two CRSs on different datums are related without a
 * shift (PROJ's default when none is defined); a pair that needs a real
 * shift, or a CRS outside that set, returns PG_ERR_INVALID with the reason.
 * Replaces, for points, the OGR transformation the reference applies to
 * every ignition and fuel geometry before rasterising it
 * (Geometry::transform_to_inplace, crates/propag/src/propag.rs:854-857,
 * 964-969).  Host code; needs no GPU. */
pg_status pg_transform_points(const char *src_proj, const char *dst_proj, size_t n, double *x,
                              double *y, uint8_t *ok);

/* Initial references of propagate() (crates/propag/src/propag.rs:433-449):
 * each cell whose time is not FLT_MAX points at itself, every other cell
 * holds 0xFFFF.  W*H host arrays; needs no GPU. */
pg_status pg_init_refs(const GeoReference *geo, const float *time, uint16_t *refs_x,
                       uint16_t *refs_y);

/* ------------------------------------------- output files (host) */

/* Writes a W*H arrival-time raster (FLT_MAX = not burned) as a GeoTIFF:
 * one f32 band, 256 x 256 tiles, uncompressed, nodata FLT_MAX, the
 * raster's geotransform and CRS (BigTIFF when the file would pass 4 GiB).
 * Replaces write_times (crates/propag/src/propag.rs:576-600). */
pg_status pg_write_times_geotiff(const char *path, const GeoReference *geo, const float *time);

/* Writes a GeoPackage with layer "fire_references": for every cell whose
 * time is below FLT_MAX, visited column by column, a line from the centre
 * of its reference cell to its own centre.  Replaces write_refs
 * (crates/propag/src/propag.rs:699-750).  Uses the system libsqlite3. */
pg_status pg_write_refs_gpkg(const char *path, const GeoReference *geo, const float *time,
                             const uint16_t *refs_x, const uint16_t *refs_y);

/* Writes a GeoPackage with a polygon layer named `layer` of the rectangles
 * of tile_w x tile_h cells that cover the raster, row by row (the launch
 * decomposition: "blocks" are the kernel's tiles, "grids" its launch
 * domains).  Replaces write_boundaries (crates/propag/src/propag.rs:601-698). */
pg_status pg_write_tiles_gpkg(const char *path, const char *layer, const GeoReference *geo,
                              uint32_t tile_w, uint32_t tile_h);

/* ---------------------------------------------- batching and serving */

/* Device-resident context: terrains stay on the device, scenarios on them
 * are admitted while memory allows and advance together in shared
 * persistent launches (several scenarios per launch, one grid barrier per
 * sweep for all of them).  These are the handle variants of the device
 * layer named in crates/propag/src/propag.rs:375-573; the reference has no
 * batching form.  Every call is synchronous and thread-safe per context. */
typedef struct pg_ctx pg_ctx;

/* Weather of one scenario: moisture ratios (spatially constant, as the
 * reference's rasters hold them in config-style runs) and the midflame wind
 * (m/s; azimuth in radians, the direction the wind blows toward).  When the
 * field pointers are not NULL they are host rasters of W*H values that
 * replace the constant wind. */
typedef struct pg_weather {
  float d1hr, d10hr, d100hr, herb, wood;
  float wind_speed, wind_azimuth;
  const float *wind_speed_field, *wind_azimuth_field;
} pg_weather;

/* One ignited cell: column, row, ignition time (s). */
typedef struct pg_ignition {
  uint32_t x, y;
  float time;
} pg_ignition;

/* Scenario states reported by pg_scenario_status.  FAILED: a launch that
 * ran the scenario returned an error (PG_ERR_TIMEOUT, PG_ERR_HIP); it has no
 * result and is reported by pg_ctx_take_finished like a finished one. */
enum { PG_SCENARIO_RUNNING = 0, PG_SCENARIO_DONE = 1, PG_SCENARIO_NOT_CONVERGED = 2, PG_SCENARIO_FAILED = 3 };

/* Creates a context on the current device.  hbm_budget caps the bytes the
 * context allocates for scenarios (0: 90 % of the free device memory). */
pg_status pg_ctx_create(size_t hbm_budget, pg_ctx **ctx);
pg_status pg_ctx_destroy(pg_ctx *ctx);

/* Sets the tile shape (8 x 8 by default; 16 x 4 or 32 x 2) of the scenarios
 * the context admits from now on; PG_ERR_BUSY while any scenario is admitted
 * (queued tiles are numbered in the shape of their scenario).  Results do not
 * depend on it.  No reference counterpart: the reference's tile is its CUDA
 * block (propag.rs:482-557); this is the autotuner's tile-shape axis. */
pg_status pg_ctx_set_tile(pg_ctx *ctx, uint32_t tile_w, uint32_t tile_h);

/* Copies a terrain to the device: fuel codes, slope (rise/run) and aspect
 * (rad), W*H host arrays in the layout of geo.  Returns its id. */
pg_status pg_terrain_add(pg_ctx *ctx, const GeoReference *geo, const uint8_t *fuel,
                         const float *slope, const float *aspect, int *terrain_id);

/* Bytes a scenario on a W x H terrain occupies while resident. */
size_t pg_scenario_bytes(uint32_t width, uint32_t height);

/* Admits a scenario: burn pass with its weather, initial state from the
 * ignitions (duplicates keep the earliest time), boundaries and the queue
 * of its first sweep, all on the device.  It joins the next pg_batch_run.
 * Returns PG_ERR_BUSY when the memory budget has no room for it (retry once
 * scenarios have been released) and PG_ERR_INVALID for an ignition outside
 * the raster.  max_iterations 0 selects the default sweep cap. */
pg_status pg_scenario_admit(pg_ctx *ctx, int terrain_id, const pg_weather *weather,
                            const pg_ignition *ignitions, size_t n_ignitions,
                            float max_time, uint32_t max_iterations, int *scenario_id);

/* Runs every admitted scenario that has not finished, all in one persistent
 * launch, for at most max_sweeps sweeps (0: until all finish).  Scenarios
 * that finish get their post pass; *n_finished (may be NULL) receives how
 * many finished in this call.  stats (may be NULL) receives the launch's
 * counters.  When the launch fails, every scenario it was running becomes
 * PG_SCENARIO_FAILED (queued in the finished list; release it) and the
 * error is returned; finished scenarios and later admissions are not
 * affected. */
pg_status pg_batch_run(pg_ctx *ctx, uint32_t max_sweeps, const pg_tuning *tuning,
                       pg_stats *stats, uint32_t *n_finished);

/* Drains the ids of scenarios that finished (done or capped) in launches
 * since the previous call, oldest first: at most `capacity` ids go to
 * `ids`, *count receives how many; the rest stay for the next call.
 * Released scenarios are dropped from the list.  A serving loop calls this
 * after each pg_batch_run instead of polling every scenario. */
pg_status pg_ctx_take_finished(pg_ctx *ctx, int *ids, uint32_t capacity, uint32_t *count);

/* State of a scenario and the sweeps it has run. */
pg_status pg_scenario_status(pg_ctx *ctx, int scenario_id, int *state, uint64_t *sweeps);

/* Copies a finished scenario's results to host arrays (any may be NULL):
 * time (W*H, FLT_MAX = not burned), refs_x/refs_y (W*H), boundaries
 * (ceil(W*H/32) words) - the outputs of pg_propagate. */
pg_status pg_scenario_result(pg_ctx *ctx, int scenario_id, float *time, uint16_t *refs_x,
                             uint16_t *refs_y, uint32_t *boundaries);

/* Frees a scenario's device memory (finished or not). */
pg_status pg_scenario_release(pg_ctx *ctx, int scenario_id);

#ifdef __cplusplus
}
#endif
#endif
