/*
 * geometry.h -- raster georeference types of the propag C ABI.
 *
 * Same names, field order and #[repr(C)] layout as the reference's types,
 * so a caller passes the same bytes:
 *   GT            crates/geometry/src/lib.rs:20-29
 *   GeoTransform  crates/geometry/src/lib.rs:128-133
 *   GeoReference  crates/geometry/src/lib.rs:180-187
 * GT holds a GDAL-ordered affine transform (x0, dx, rx, y0, ry, dy); `inv` is
 * its inverse.  dy < 0 means north-up (row 0 at the top).  proj is a
 * NUL-padded proj4 string.
 *
 * The GeoReference_* functions are exported by libpropag.so, the plugin
 * library (crates/geometry/src/lib.rs:195-332).
 */
#ifndef GEOMETRY_H
#define GEOMETRY_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GT {
  float x0, dx, rx, y0, ry, dy;
} GT;

typedef struct GeoTransform {
  GT gt;
  GT inv;
} GeoTransform;

typedef struct GeoReference {
  uint32_t width;
  uint32_t height;
  uint8_t proj[1024];
  GeoTransform transform;
} GeoReference;

/* width * height (lib.rs:195-198). */
uint32_t GeoReference_len(const GeoReference *self);
/* Extent of the raster along x and y, lower then upper bound, whatever the
 * sign of dx / dy (lib.rs:200-232). */
float GeoReference_x0(const GeoReference *self);
float GeoReference_x1(const GeoReference *self);
float GeoReference_y0(const GeoReference *self);
float GeoReference_y1(const GeoReference *self);
/* Georeference covering the box (x0, y0)-(x1, y1) with pixels dx by dy:
 * width = ceil((x1 - x0) / dx), height = ceil((y1 - y0) / dy); origin at
 * (x0, y0) with +dy (south-up) or at (x0, y1) with -dy (north-up).  false
 * for a non-invertible transform, a negative or non-finite size, or a proj
 * that is not UTF-8 (lib.rs:234-332). */
bool GeoReference_south_up(float x0, float y0, float x1, float y1, float dx, float dy,
                           const char *proj, GeoReference *result);
bool GeoReference_north_up(float x0, float y0, float x1, float y1, float dx, float dy,
                           const char *proj, GeoReference *result);

#ifdef __cplusplus
}
#endif
#endif
