// Device-side layout of a propagation scenario and the launch interface of
// the propagation kernels (propag.hip), shared with the batching context
// (context.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "pg_internal.h"

namespace pg {

constexpr int kTile = 8;  // default tile edge (cells): 8 x 8, one cell per lane of a wave

// Tile shapes: tw x (64 / tw) cells, one cell per lane of a wave.
inline bool tile_shape_ok(uint32_t w, uint32_t h) {
  return (w == 8 && h == 8) || (w == 16 && h == 4) || (w == 32 && h == 2);
}
inline size_t tiles_of(size_t W, size_t H, int tw) {
  const size_t th = static_cast<size_t>(64 / tw);
  return ((W + tw - 1) / tw) * ((H + th - 1) / th);
}
// The most tiles any shape cuts a raster into (buffers sized for every shape).
inline size_t max_tiles(size_t W, size_t H) {
  size_t m = tiles_of(W, H, 8);
  if (tiles_of(W, H, 16) > m) m = tiles_of(W, H, 16);
  if (tiles_of(W, H, 32) > m) m = tiles_of(W, H, 32);
  return m;
}

// Bytes of the device array behind pg_propagate_noalloc: the call and launch
// workspaces of any raster of at most PG_NOALLOC_MAX_CELLS cells (the largest,
// a 17 x 61680 raster, needs 43.5 MiB on a 256-CU device).
constexpr size_t kNoAllocScratchBytes = size_t{44} << 20;

struct Geom {
  int W, H;
  float dx, dy, cell, max_time;
  int north_up, square;
};

// Per-cell fire coefficients of time_from: A = e cos(az), B = e sin(az),
// R = 1 / (speed (1 - e)); null_fire = 1 where speed == 0.
struct alignas(16) Coef {
  float a, b, r, null_fire;
};

// One scenario as the kernels read it (read-only during a launch).  Its
// tile t is bit (t >> nwords_log2) of word t & (2^nwords_log2 - 1) of the
// scenario's mark bitmaps (one per sweep parity).
struct Scen {
  Geom g;
  const Coef *coef;
  uint2 *st[2];          // double-buffered (time bits, reference word)
  int *stamp;            // last sweep in which the cell changed
  const uint32_t *bits;  // boundary bitset (row-major, the output layout)
  const unsigned long long *blk;  // the same bits in 8x8 blocks (walks)
  uint4 *box;            // per tile: bounding box (x0, y0, x1, y1, in cells) of its far dependents
  unsigned long long *isref;  // per tile: its cells that are some cell's far reference (bit x + tw y)
  uint32_t *mark[2];     // tiles queued for a sweep of parity 0 / 1
  int ntx, nty;          // tiles per row / column
  int tw, th;            // tile shape (cells)
  int bw, bh;            // 8x8 blocks of the blocked bitset per row / column
  uint32_t nwords_log2, pad0;
  int k0;                // sweep number of the scenario's first sweep
  unsigned max_iters;    // sweeps allowed from k0
  int cache;             // line-of-sight flags cached in reference words
  int pad;
};

// Per-scenario results written by the kernels.
struct ScenOut {
  int last_change;  // last sweep in which a cell changed (k0 - 1: none)
  unsigned capped;  // 1: work remained when max_iters sweeps had run
  unsigned err;     // set-up: 2 = a burned cell has a reference outside the raster
  unsigned pad;
};

// Per-sweep queues of tiles to visit.  The tiles marked for sweep k sit in
// the nsub sub-lists of list k & 3, sub-list x holding
// list[k & 3][x][0 .. count[k & 3][x]) as scenario << 32 | xcd << 29 | tile,
// each tile once (the scenario's mark bitmap of parity k & 1 has the tile's
// bit set by the first marker; the reader clears it).  With nsub = 8 a
// tile goes to the sub-list of its xcd tag and is run by the workgroups of
// that XCD (workgroup w runs on XCD w % 8); with nsub = 1 every workgroup
// draws from the one list.  Unused list slots hold ~0.  Four lists rotate so
// that a count can be reset two barriers before it is appended to again.
// The queues outlive a launch: a launch that stops at k_end leaves the
// tiles of sweep k_end queued.
struct TileQueue {
  unsigned long long *list;  // sub-list (p, x) at list + (p * nsub + x) * cap
  uint32_t *count;           // its count at count + 32 * (p * nsub + x) (one 128-byte line each)
  size_t cap;                // slots per sub-list (>= tiles of all scenarios)
  uint32_t nsub;             // sub-lists per list: 1, or 8 (one per XCD)
  uint32_t pad;
};
constexpr uint32_t kXcds = 8;  // XCDs of an MI355X; workgroup w runs on XCD w % 8

// Bitmap words a scenario of ntiles tiles uses (a power of two >= ntiles/32).
inline uint32_t need_words_log2(size_t ntiles) {
  uint32_t l = 0;
  while ((static_cast<size_t>(32) << l) < ntiles) ++l;
  return l;
}

// Device buffers of one scenario, carved by the caller.
struct ScenBuffers {
  Coef *coef;
  uint2 *st0, *st1;
  int *stamp;
  uint4 *box;
  unsigned long long *blk;    // blocked boundary bitset
  unsigned long long *isref;  // far-reference cells per tile
};
size_t scen_buffer_bytes(size_t width, size_t height);
ScenBuffers scen_carve(void *base, size_t width, size_t height);

// Fills sc's geometry and its tiling into tiles of tw x (64 / tw) cells from
// a Settings (buffers and bitmap range are the caller's).
pg_status scen_geometry(const Settings *settings, Scen *sc, int tw = kTile);

// Boundaries, coefficients, initial state and first-sweep tiles of scenario
// number si (start sweep sc.k0): writes `boundaries`, the scenario's
// rasters, queues in q the tiles of sweep sc.k0 and resets *out, after
// clearing the scenario's mark bitmaps, dependents boxes and far-reference
// bitmaps (and, with `clear_queue`, every list and count of q: a queue of
// this scenario alone).  With `validate`, synchronises the
// stream and returns PG_ERR_INVALID if a burned cell has a reference outside
// the raster (such cells are dropped either way).
pg_status scen_prepare(const DeviceInfo *dev, const Scen &sc, uint32_t si, ScenOut *out, const TileQueue &q,
                       const float *speed, const float *az, const float *ecc, const float *time,
                       const uint16_t *refs_x, const uint16_t *refs_y, uint32_t *boundaries, bool validate,
                       hipStream_t stream, bool clear_queue = false);

// Admission of scenario si on a resident terrain (fuel, slope, aspect) with
// spatially constant moisture and wind and at most kAdmitIgnitions point
// ignitions (distinct cells): burn pass into speed/az/ecc, boundaries in
// both layouts, time_from's coefficients, initial state (also into the
// result rasters time / refs_x / refs_y), the tiles of sweep sc.k0 queued in
// q and *out reset.  Two launches; nothing else needs to precede it.
constexpr size_t kAdmitIgnitions = 16;
pg_status scen_admit(const DeviceInfo *dev, const Scen &sc, uint32_t si, ScenOut *out, const TileQueue &q,
                     const uint8_t *fuel, const float *slope, const float *aspect, const float moisture[5],
                     float wind_speed, float wind_azimuth, float *speed, float *az, float *ecc, float *time,
                     uint16_t *refs_x, uint16_t *refs_y, uint32_t *boundaries, const pg_ignition *ign, size_t n_ign,
                     hipStream_t stream);

// Initial state from point ignitions (at most one per cell, device array).
pg_status scen_init_state(float *time, uint16_t *refs_x, uint16_t *refs_y, size_t n, uint32_t width,
                          const pg_ignition *ignitions, size_t n_ignitions, hipStream_t stream);

pg_status scen_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t stream);

// post_propagate for one scenario from state buffer `final_buf`, or, when
// `out` (device) is given, from the buffer its launch outcome names.
pg_status scen_finish(const Scen &sc, int final_buf, const float *speed, float *time, uint16_t *refs_x,
                      uint16_t *refs_y, hipStream_t stream, const ScenOut *out = nullptr);

// The post-pass of finished scenarios of a batching context, whose
// descriptors are on the device (d_scen[slot]), kPostBatch per launch.
struct PostItem {
  int slot, final_buf;
  const float *speed;
  float *time;
  uint16_t *rx, *ry;
};
constexpr int kPostBatch = 32;
pg_status scen_finish_batch(const DeviceInfo *dev, const Scen *d_scen, const PostItem *items, size_t n_items,
                            size_t max_cells, hipStream_t stream);

// Buffer that holds a finished scenario's final state.
__host__ __device__ inline int scen_final_buf(const Scen &sc, const ScenOut &o) {
  if (o.capped) return static_cast<int>((static_cast<unsigned>(sc.k0) + sc.max_iters) & 1u);
  return (o.last_change + 2) & 1;
}

// One launch of sweeps [k_begin, k_end) over nscen scenarios (device arrays
// scen and out).
struct Launch {
  const Scen *scen;
  ScenOut *out;
  int nscen;
  TileQueue q;
  int k_begin, k_end;
  int cache;           // all scenarios cache line-of-sight flags
  int tw;              // tile width of all scenarios (their Scen::tw)
  const uint32_t *lds_bits_src;  // nscen == 1: the (row-major) bitset, staged in LDS if it fits
  size_t lds_bits_words;
  char *ws;            // launch workspace of launch_ws_bytes(cu count) bytes, or null for the
                       // per-device one (grown with hipMalloc)
  int untimed;         // 1: no HIP events (relax_ms stays 0)
  // Work to enqueue on the stream right after the relaxation kernel(s)
  // (before the launch's one synchronisation), e.g. the post-pass; it
  // returns PG_OK or an error.  May be empty.
  pg_status (*after)(void *ctx, hipStream_t stream);
  void *after_ctx;
};

// Bytes of a launch workspace (control block, barrier words, profile rows
// when `profile`) on a device with `cu` compute units.
size_t launch_ws_bytes(unsigned cu, bool profile);

struct LaunchResult {
  int last_sweep;      // last sweep run
  int finished;        // 1: a sweep committed nothing (all scenarios done)
};

// Runs the launch (persistent cooperative kernel, or one launch per sweep
// with tuning->mode == 1) and adds its counters to *stats.
pg_status run_sweeps(const DeviceInfo *dev, const Launch &L, const pg_tuning *tuning, pg_stats *stats,
                     LaunchResult *res, hipStream_t stream);

}  // namespace pg
