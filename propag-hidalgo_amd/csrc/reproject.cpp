// pg_transform_points: the CRS transformation of reproject.h over a C ABI
// (include/propag_hip.h).  Host code; needs no GPU.
#include <string>

#include "pg_internal.h"
#include "reproject.h"

using namespace pg;

extern "C" pg_status pg_transform_points(const char *src_proj, const char *dst_proj, size_t n, double *x, double *y,
                                         uint8_t *ok) {
  if (!src_proj || !dst_proj) return set_error(PG_ERR_INVALID, "pg_transform_points: null CRS");
  if (n && (!x || !y || !ok)) return set_error(PG_ERR_INVALID, "pg_transform_points: null coordinate array");
  Transform tr;
  std::string err;
  if (!tr.init(src_proj, dst_proj, err)) return set_error(PG_ERR_INVALID, err);
  for (size_t i = 0; i < n; ++i) {
    double a = x[i], b = y[i];
    ok[i] = tr.apply(a, b) ? 1 : 0;
    if (ok[i]) {
      x[i] = a;
      y[i] = b;
    }
  }
  return PG_OK;
}
