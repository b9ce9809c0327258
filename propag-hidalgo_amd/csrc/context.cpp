// Device-resident batching context (pg_ctx_*, include/propag_hip.h).
//
// Terrains stay on the device.  A scenario on a terrain owns one device
// block (its rasters, state, mark bitmaps, fire and result buffers) and a
// slot in the context's descriptor array.  All unfinished scenarios advance
// together in one persistent launch per pg_batch_run: they share the grid
// barrier and the per-sweep tile queues, and each starts at the sweep that
// was next when it was admitted, so a scenario admitted while others are
// half-way joins the next launch (continuous batching at launch
// granularity).  A scenario's result does not depend on what else runs:
// its sweeps read only its own state.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "propag_dev.h"

namespace pg {
namespace {

constexpr int kMaxSlots = 4096;

struct Terrain {
  GeoReference geo;
  uint8_t *fuel = nullptr;
  float *slope = nullptr, *aspect = nullptr;
};

struct Scenario {
  bool used = false;
  bool zombie = false;  // released while running: entries may still name the slot
  int state = PG_SCENARIO_RUNNING;
  bool posted = false;
  int terrain = -1;
  Scen sc{};
  void *block = nullptr;
  size_t bytes = 0;
  size_t tiles = 0;
  float *speed = nullptr, *az = nullptr, *ecc = nullptr, *time = nullptr;
  uint16_t *rx = nullptr, *ry = nullptr;
  uint32_t *bits = nullptr;
  uint64_t sweeps = 0;
};

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Layout {
  size_t scen, mark, fire, time, refs, bits, total;
};

Layout layout_for(size_t W, size_t H) {
  const size_t n = W * H;
  const size_t nwords = static_cast<size_t>(1) << need_words_log2(max_tiles(W, H));
  Layout L{};
  size_t off = 0;
  L.scen = off; off = align_up(off + scen_buffer_bytes(W, H), 256);
  L.mark = off; off = align_up(off + 2 * nwords * sizeof(uint32_t), 256);
  L.fire = off; off = align_up(off + 3 * n * sizeof(float), 256);
  L.time = off; off = align_up(off + n * sizeof(float), 256);
  L.refs = off; off = align_up(off + 2 * n * sizeof(uint16_t), 256);
  L.bits = off; off = align_up(off + ((n + 31) / 32) * sizeof(uint32_t), 256);
  L.total = off;
  return L;
}

}  // namespace
}  // namespace pg

struct pg_ctx {
  pg::DeviceInfo *dev = nullptr;
  std::mutex mu;
  hipStream_t stream = nullptr;
  size_t budget = 0, used = 0;
  std::vector<pg::Terrain> terrains;
  std::vector<pg::Scenario> scen;  // index = scenario id = slot
  pg::Scen *d_scen = nullptr;
  pg::ScenOut *d_out = nullptr;
  // Host mirror of the descriptors (pinned); slots [desc_lo, desc_hi] changed
  // since the last upload, which pg_batch_run issues before its launch.
  pg::Scen *h_scen = nullptr;
  int desc_lo = 1 << 30, desc_hi = -1;
  pg::ScenOut *h_out = nullptr;  // pinned: the launch outcomes are copied here
  pg::TileQueue q{};
  void *q_block = nullptr;
  // Ignition lists of admissions since the last synchronisation (the
  // device copies read them asynchronously), one device buffer each.
  std::vector<std::vector<pg_ignition>> host_ign;
  std::vector<void *> dev_ign;
  // Freed scenario blocks by size, reused before allocating.
  std::multimap<size_t, void *> pool;
  // Scenarios finished since the last pg_ctx_take_finished, oldest first.
  std::vector<int> finished;
  int k_next = 0;
  int tile_w = pg::kTile;  // tile shape of the scenarios it admits: tile_w x 64 / tile_w
};

namespace pg {
namespace {

// A dead descriptor: the kernels skip entries of a scenario with no sweeps
// left before touching any of its buffers.
Scen dead_descriptor() {
  Scen s{};
  s.max_iters = 0;
  return s;
}

size_t live_tiles(const pg_ctx *c) {
  size_t t = 0;
  for (const Scenario &s : c->scen)
    if (s.used || s.zombie) t += s.tiles;
  return t;
}

// Grows the tile queues to `cap` slots per list, keeping the entries queued
// for sweep k_next (the only list that holds entries between launches).
pg_status grow_queue(pg_ctx *c, size_t cap) {
  if (cap <= c->q.cap) return PG_OK;
  cap = std::max(cap, 2 * c->q.cap);
  // Eight sub-lists per list, one per XCD (each able to hold every tile).
  const size_t nsub = kXcds;
  const size_t list_bytes = align_up(cap * sizeof(unsigned long long), 256);
  const size_t count_words = 4 * nsub * 32;
  void *blk = nullptr;
  hipError_t e = hipMalloc(&blk, 4 * nsub * list_bytes + count_words * 4);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PG_ERR_OUT_OF_MEMORY, "tile queue allocation failed");
  }
  TileQueue nq{};
  nq.list = reinterpret_cast<unsigned long long *>(blk);
  nq.count = reinterpret_cast<uint32_t *>(static_cast<char *>(blk) + 4 * nsub * list_bytes);
  nq.cap = list_bytes / sizeof(unsigned long long);
  nq.nsub = static_cast<uint32_t>(nsub);
  pg_status st = scen_fill_u32(reinterpret_cast<uint32_t *>(blk), 4 * nsub * list_bytes / 4, 0xFFFFFFFFu, c->stream);
  if (st != PG_OK) return st;
  st = scen_fill_u32(nq.count, count_words, 0u, c->stream);
  if (st != PG_OK) return st;
  if (c->q_block) {
    // Only the list of sweep k_next holds entries between launches.
    PG_HIP(hipMemcpyAsync(nq.count, c->q.count, count_words * 4, hipMemcpyDeviceToDevice, c->stream));
    const size_t p = static_cast<size_t>(c->k_next & 3);
    uint32_t pending[kXcds] = {};
    for (size_t x = 0; x < nsub; ++x)
      PG_HIP(hipMemcpyAsync(&pending[x], c->q.count + 32 * (p * nsub + x), sizeof(uint32_t), hipMemcpyDeviceToHost,
                            c->stream));
    PG_HIP(hipStreamSynchronize(c->stream));
    for (size_t x = 0; x < nsub; ++x)
      if (pending[x])
        PG_HIP(hipMemcpyAsync(nq.list + (p * nsub + x) * nq.cap, c->q.list + (p * nsub + x) * c->q.cap,
                              std::min<size_t>(pending[x], c->q.cap) * sizeof(unsigned long long),
                              hipMemcpyDeviceToDevice, c->stream));
    PG_HIP(hipStreamSynchronize(c->stream));
    PG_HIP(hipFree(c->q_block));
  }
  c->q_block = blk;
  c->q = nq;
  return PG_OK;
}

// Sets a slot's descriptor in the host mirror; the next launch uploads it.
// (The mirror changes only while no upload from it is in flight: every
// pg_batch_run synchronises its stream before returning.)
pg_status put_descriptor(pg_ctx *c, int slot, const Scen *sc) {
  c->h_scen[slot] = *sc;
  c->desc_lo = std::min(c->desc_lo, slot);
  c->desc_hi = std::max(c->desc_hi, slot);
  return PG_OK;
}

// Uploads the descriptors changed since the last upload, in one copy.
pg_status flush_descriptors(pg_ctx *c) {
  if (c->desc_hi < c->desc_lo) return PG_OK;
  const size_t n = static_cast<size_t>(c->desc_hi - c->desc_lo + 1);
  PG_HIP(hipMemcpyAsync(c->d_scen + c->desc_lo, c->h_scen + c->desc_lo, n * sizeof(Scen), hipMemcpyHostToDevice,
                        c->stream));
  c->desc_lo = 1 << 30;
  c->desc_hi = -1;
  return PG_OK;
}

// After the context's stream has been synchronised: host staging is free.
void drop_staging(pg_ctx *c) {
  c->host_ign.clear();
  for (void *p : c->dev_ign) (void)hipFree(p);
  c->dev_ign.clear();
}

pg_status sync_ctx(pg_ctx *c) {
  PG_HIP(hipStreamSynchronize(c->stream));
  drop_staging(c);
  return PG_OK;
}

void *take_block(pg_ctx *c, size_t bytes) {
  auto it = c->pool.find(bytes);
  if (it != c->pool.end()) {
    void *p = it->second;
    c->pool.erase(it);
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    // give pooled blocks back and retry once
    (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->pool) (void)hipFree(kv.second);
    c->pool.clear();
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  return p;
}

const Scen *dead_descriptor_ptr() {
  static const Scen dead = dead_descriptor();
  return &dead;
}

// A launch that returns an error (a barrier wait past its spin limit, a HIP
// failure) has consumed part of a sweep: the tile queues and the state of
// every running scenario are part-way through it.  Those scenarios become
// PG_SCENARIO_FAILED (dead descriptors, so no later launch touches their
// buffers; reported by pg_ctx_take_finished, to be released by the caller),
// the queues are cleared, and `st` is returned with its message.  Finished
// scenarios keep their results; later admissions run normally.
pg_status fail_running(pg_ctx *c, pg_status st) {
  const std::string msg = pg_last_error();
  (void)hipStreamSynchronize(c->stream);
  (void)hipGetLastError();
  for (size_t i = 0; i < c->scen.size(); ++i) {
    Scenario &s = c->scen[i];
    s.zombie = false;  // the cleared queues hold no entry of any slot
    if (!s.used || s.state != PG_SCENARIO_RUNNING) continue;
    s.state = PG_SCENARIO_FAILED;
    (void)put_descriptor(c, static_cast<int>(i), dead_descriptor_ptr());
    c->finished.push_back(static_cast<int>(i));
  }
  if (c->q_block) {
    (void)scen_fill_u32(reinterpret_cast<uint32_t *>(c->q.list), 4 * c->q.nsub * c->q.cap * 2, 0xFFFFFFFFu,
                        c->stream);
    (void)scen_fill_u32(c->q.count, 4 * c->q.nsub * 32, 0u, c->stream);
  }
  (void)sync_ctx(c);
  return set_error(st, msg + " (the running scenarios of the context are marked failed)");
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" {

size_t pg_scenario_bytes(uint32_t width, uint32_t height) { return layout_for(width, height).total; }

pg_status pg_ctx_create(size_t hbm_budget, pg_ctx **out) {
  if (!out) return set_error(PG_ERR_INVALID, "pg_ctx_create: null output");
  *out = nullptr;
  DeviceInfo *dev = nullptr;
  pg_status st = device_info(&dev);
  if (st != PG_OK) return st;
  auto *c = new pg_ctx();
  c->dev = dev;
  if (hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      delete c;
      return set_error(PG_ERR_HIP, "hipMemGetInfo failed");
    }
    hbm_budget = free_b / 10 * 9;
  }
  c->budget = hbm_budget;
  c->scen.reserve(kMaxSlots);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&c->d_scen), kMaxSlots * sizeof(Scen));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&c->d_out), kMaxSlots * sizeof(ScenOut));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&c->h_scen), kMaxSlots * sizeof(Scen));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&c->h_out), kMaxSlots * sizeof(ScenOut));
  if (e != hipSuccess) {
    st = hip_error(e, "pg_ctx_create");
    pg_ctx_destroy(c);
    return st;
  }
  *out = c;
  return PG_OK;
}

pg_status pg_ctx_destroy(pg_ctx *c) {
  if (!c) return PG_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_staging(c);
  for (Scenario &s : c->scen)
    if (s.block) (void)hipFree(s.block);
  for (auto &kv : c->pool) (void)hipFree(kv.second);
  for (Terrain &t : c->terrains) {
    if (t.fuel) (void)hipFree(t.fuel);
    if (t.slope) (void)hipFree(t.slope);
    if (t.aspect) (void)hipFree(t.aspect);
  }
  if (c->q_block) (void)hipFree(c->q_block);
  if (c->d_scen) (void)hipFree(c->d_scen);
  if (c->d_out) (void)hipFree(c->d_out);
  if (c->h_scen) (void)hipHostFree(c->h_scen);
  if (c->h_out) (void)hipHostFree(c->h_out);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PG_OK;
}

pg_status pg_terrain_add(pg_ctx *c, const GeoReference *geo, const uint8_t *fuel, const float *slope,
                         const float *aspect, int *terrain_id) {
  if (!c || !geo || !fuel || !slope || !aspect || !terrain_id)
    return set_error(PG_ERR_INVALID, "pg_terrain_add: null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  const size_t n = static_cast<size_t>(geo->width) * geo->height;
  if (n == 0) return set_error(PG_ERR_INVALID, "terrain has no cells");
  if (geo->width > 0x7fffu || geo->height > 0x7fffu)
    return set_error(PG_ERR_INVALID, "context terrains need sides below 32768");
  Terrain t;
  t.geo = *geo;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&t.fuel), n);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&t.slope), n * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&t.aspect), n * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(t.fuel, fuel, n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t.slope, slope, n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t.aspect, aspect, n * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (t.fuel) (void)hipFree(t.fuel);
    if (t.slope) (void)hipFree(t.slope);
    if (t.aspect) (void)hipFree(t.aspect);
    return hip_error(e, "pg_terrain_add");
  }
  c->terrains.push_back(t);
  *terrain_id = static_cast<int>(c->terrains.size()) - 1;
  return PG_OK;
}

pg_status pg_scenario_admit(pg_ctx *c, int terrain_id, const pg_weather *w, const pg_ignition *ign,
                            size_t n_ign, float max_time, uint32_t max_iterations, int *scenario_id) {
  if (!c || !w || (!ign && n_ign) || !scenario_id)
    return set_error(PG_ERR_INVALID, "pg_scenario_admit: null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  if (terrain_id < 0 || static_cast<size_t>(terrain_id) >= c->terrains.size())
    return set_error(PG_ERR_INVALID, "unknown terrain id");
  const Terrain &t = c->terrains[terrain_id];
  Settings settings{};
  settings.geo_ref = t.geo;
  settings.max_time = max_time;
  Scen sc;
  pg_status st = scen_geometry(&settings, &sc, c->tile_w);
  if (st != PG_OK) return st;
  if (max_iterations) sc.max_iters = max_iterations;
  if (sc.max_iters > 0x3fffffffu) return set_error(PG_ERR_INVALID, "max_iterations too large");
  const uint32_t W = t.geo.width, H = t.geo.height;
  const size_t n = static_cast<size_t>(W) * H;
  // One ignition per cell, the earliest time.
  std::map<std::pair<uint32_t, uint32_t>, float> cells;
  for (size_t i = 0; i < n_ign; ++i) {
    if (ign[i].x >= W || ign[i].y >= H) return set_error(PG_ERR_INVALID, "ignition outside the raster");
    if (!(ign[i].time < FLT_MAX)) return set_error(PG_ERR_INVALID, "ignition time must be below FLT_MAX");
    auto key = std::make_pair(ign[i].y, ign[i].x);
    auto it = cells.find(key);
    if (it == cells.end() || ign[i].time < it->second) cells[key] = ign[i].time;
  }
  std::vector<pg_ignition> uniq;
  uniq.reserve(cells.size());
  for (const auto &kv : cells) uniq.push_back(pg_ignition{kv.first.second, kv.first.first, kv.second});

  const Layout L = layout_for(W, H);
  if (c->used + L.total > c->budget) return set_error(PG_ERR_BUSY, "memory budget full");
  int slot = -1;
  for (size_t i = 0; i < c->scen.size(); ++i)
    if (!c->scen[i].used && !c->scen[i].zombie) {
      slot = static_cast<int>(i);
      break;
    }
  if (slot < 0) {
    if (c->scen.size() >= static_cast<size_t>(kMaxSlots)) return set_error(PG_ERR_BUSY, "all scenario slots in use");
    c->scen.emplace_back();
    slot = static_cast<int>(c->scen.size()) - 1;
  }
  Scenario s;
  s.block = take_block(c, L.total);
  if (!s.block) return set_error(PG_ERR_BUSY, "device allocation failed");
  hipError_t e = hipSuccess;
  char *b = static_cast<char *>(s.block);
  const ScenBuffers sb = scen_carve(b + L.scen, W, H);
  sc.coef = sb.coef;
  sc.st[0] = sb.st0;
  sc.st[1] = sb.st1;
  sc.stamp = sb.stamp;
  sc.box = sb.box;
  sc.blk = sb.blk;
  sc.isref = sb.isref;
  const uint32_t nwords = 1u << sc.nwords_log2;
  sc.mark[0] = reinterpret_cast<uint32_t *>(b + L.mark);
  sc.mark[1] = sc.mark[0] + nwords;
  s.speed = reinterpret_cast<float *>(b + L.fire);
  s.az = s.speed + n;
  s.ecc = s.az + n;
  s.time = reinterpret_cast<float *>(b + L.time);
  s.rx = reinterpret_cast<uint16_t *>(b + L.refs);
  s.ry = s.rx + n;
  s.bits = reinterpret_cast<uint32_t *>(b + L.bits);
  sc.bits = s.bits;
  sc.k0 = c->k_next;
  s.sc = sc;
  s.terrain = terrain_id;
  s.bytes = L.total;
  s.tiles = static_cast<size_t>(sc.ntx) * sc.nty;

  auto fail = [&](pg_status code) {
    (void)sync_ctx(c);
    c->pool.emplace(L.total, s.block);
    return code;
  };
  st = grow_queue(c, live_tiles(c) + s.tiles);
  if (st != PG_OK) return fail(st);
  const float moist[5] = {w->d1hr, w->d10hr, w->d100hr, w->herb, w->wood};
  const size_t n_uniq = uniq.size();
  const bool fields = w->wind_speed_field || w->wind_azimuth_field;
  if (fields && (!w->wind_speed_field || !w->wind_azimuth_field))
    return fail(set_error(PG_ERR_INVALID, "wind fields come in pairs"));
  const float *wsf = nullptr, *wazf = nullptr;
  if (!fields && n_uniq <= kAdmitIgnitions) {
    // constant weather, few ignitions: two launches, ignitions by value
    st = scen_admit(c->dev, sc, static_cast<uint32_t>(slot), c->d_out + slot, c->q, t.fuel, t.slope, t.aspect, moist,
                    w->wind_speed, w->wind_azimuth, s.speed, s.az, s.ecc, s.time, s.rx, s.ry, s.bits, uniq.data(),
                    n_uniq, c->stream);
    if (st != PG_OK) return fail(st);
  } else {
    // Per-cell wind goes through the time / refs buffers (8 bytes per cell)
    // before the initial state overwrites them.
    if (fields) {
      e = hipMemcpyAsync(s.time, w->wind_speed_field, n * sizeof(float), hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(s.rx, w->wind_azimuth_field, n * sizeof(float), hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) return fail(hip_error(e, "wind upload"));
      wsf = s.time;
      wazf = reinterpret_cast<const float *>(s.rx);
    }
    st = launch_burn_weather(c->dev, n, t.fuel, t.slope, t.aspect, moist, w->wind_speed, w->wind_azimuth, wsf, wazf,
                             s.speed, s.az, s.ecc, c->stream);
    if (st != PG_OK) return fail(st);
    pg_ignition *d_ign = nullptr;
    if (n_uniq) {
      e = hipMalloc(reinterpret_cast<void **>(&d_ign), n_uniq * sizeof(pg_ignition));
      if (e != hipSuccess) return fail(hip_error(e, "ignition buffer"));
      c->dev_ign.push_back(d_ign);
      c->host_ign.push_back(std::move(uniq));
      const std::vector<pg_ignition> &hv = c->host_ign.back();
      e = hipMemcpyAsync(d_ign, hv.data(), n_uniq * sizeof(pg_ignition), hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) return fail(hip_error(e, "ignition upload"));
    }
    st = scen_init_state(s.time, s.rx, s.ry, n, W, d_ign, n_uniq, c->stream);
    if (st != PG_OK) return fail(st);
    st = scen_prepare(c->dev, sc, static_cast<uint32_t>(slot), c->d_out + slot, c->q, s.speed, s.az, s.ecc, s.time,
                      s.rx, s.ry, s.bits, false, c->stream);
    if (st != PG_OK) return fail(st);
  }
  s.used = true;
  c->scen[slot] = s;
  st = put_descriptor(c, slot, &c->scen[slot].sc);
  if (st != PG_OK) {
    c->scen[slot] = Scenario{};
    return fail(st);
  }
  // Caller-owned wind rasters are read by the copies above: finish them.
  if (wsf) {
    st = sync_ctx(c);
    if (st != PG_OK) return st;
  }
  c->used += s.bytes;
  *scenario_id = slot;
  return PG_OK;
}

pg_status pg_ctx_set_tile(pg_ctx *c, uint32_t tile_w, uint32_t tile_h) {
  if (!c) return set_error(PG_ERR_INVALID, "pg_ctx_set_tile: null context");
  if (!tile_shape_ok(tile_w, tile_h)) return set_error(PG_ERR_INVALID, "tile shape must be 8x8, 16x4 or 32x2");
  std::lock_guard<std::mutex> lk(c->mu);
  if (static_cast<int>(tile_w) == c->tile_w) return PG_OK;
  for (const Scenario &s : c->scen)
    if (s.used || s.zombie) return set_error(PG_ERR_BUSY, "the tile shape changes only with no scenario admitted");
  c->tile_w = static_cast<int>(tile_w);
  return PG_OK;
}

pg_status pg_batch_run(pg_ctx *c, uint32_t max_sweeps, const pg_tuning *tuning, pg_stats *stats,
                       uint32_t *n_finished) {
  if (!c) return set_error(PG_ERR_INVALID, "pg_batch_run: null context");
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_finished) *n_finished = 0;
  if (stats) *stats = pg_stats{};
  int nslots = 0;
  long long k_end = c->k_next + 1;
  bool any = false;
  int only = -1, live = 0;
  for (size_t i = 0; i < c->scen.size(); ++i) {
    const Scenario &s = c->scen[i];
    if (s.used || s.zombie) nslots = static_cast<int>(i) + 1;
    if (s.used && s.state == PG_SCENARIO_RUNNING) {
      any = true;
      k_end = std::max<long long>(k_end, static_cast<long long>(s.sc.k0) + s.sc.max_iters + 1);
      only = static_cast<int>(i);
      ++live;
    }
  }
  if (!any) return PG_OK;
  if (tuning && (tuning->tile_w || tuning->tile_h) &&
      (static_cast<int>(tuning->tile_w) != c->tile_w || static_cast<int>(tuning->tile_h) != 64 / c->tile_w))
    return set_error(PG_ERR_INVALID, "the tuning's tile shape differs from the context's (pg_ctx_set_tile)");
  if (max_sweeps) k_end = std::min<long long>(k_end, static_cast<long long>(c->k_next) + max_sweeps);
  if (k_end > 0x7fff0000LL) return set_error(PG_ERR_INVALID, "sweep counter exhausted; recreate the context");

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  PG_HIP(hipEventCreate(&ev0));
  PG_HIP(hipEventCreate(&ev1));
  struct EventGuard {
    hipEvent_t *e[2];
    ~EventGuard() {
      for (hipEvent_t *p : e)
        if (*p) (void)hipEventDestroy(*p);
    }
  } guard{{&ev0, &ev1}};
  pg_status fs = flush_descriptors(c);
  if (fs != PG_OK) return fs;
  PG_HIP(hipEventRecord(ev0, c->stream));
  Launch L{};
  L.scen = c->d_scen;
  L.out = c->d_out;
  L.nscen = nslots;
  L.q = c->q;
  L.k_begin = c->k_next;
  L.k_end = static_cast<int>(k_end);
  L.cache = 1;
  L.tw = c->tile_w;
  if (live == 1 && nslots == 1) {
    L.lds_bits_src = c->scen[only].bits;
    L.lds_bits_words = (static_cast<size_t>(c->scen[only].sc.g.W) * c->scen[only].sc.g.H + 31) / 32;
  }
  pg_stats local{};
  LaunchResult res{};
  pg_status st = run_sweeps(c->dev, L, tuning, &local, &res, c->stream);
  if (st != PG_OK) return fail_running(c, st);
  const ScenOut *outs = c->h_out;
  PG_HIP(hipMemcpyAsync(c->h_out, c->d_out, nslots * sizeof(ScenOut), hipMemcpyDeviceToHost, c->stream));
  PG_HIP(hipStreamSynchronize(c->stream));
  uint32_t fin = 0;
  std::vector<PostItem> posts;
  size_t max_cells = 0;
  for (int i = 0; i < nslots; ++i) {
    Scenario &s = c->scen[i];
    if (s.zombie) {  // its queued entries were consumed (and skipped) by this launch
      s.zombie = false;
      continue;
    }
    if (!s.used || s.state != PG_SCENARIO_RUNNING) continue;
    const ScenOut &o = outs[i];
    if (o.capped) {
      s.state = PG_SCENARIO_NOT_CONVERGED;
      s.sweeps = s.sc.max_iters;
    } else if (o.last_change < res.last_sweep) {
      s.state = PG_SCENARIO_DONE;
      s.sweeps = static_cast<uint64_t>(o.last_change - s.sc.k0 + 2);
    } else {
      s.sweeps = static_cast<uint64_t>(res.last_sweep - s.sc.k0 + 1);
      continue;
    }
    posts.push_back(PostItem{i, scen_final_buf(s.sc, o), s.speed, s.time, s.rx, s.ry});
    max_cells = std::max(max_cells, static_cast<size_t>(s.sc.g.W) * s.sc.g.H);
    s.posted = true;
    c->finished.push_back(i);
    ++fin;
  }
  st = scen_finish_batch(c->dev, c->d_scen, posts.data(), posts.size(), max_cells, c->stream);
  if (st != PG_OK) return st;
  c->k_next = res.last_sweep + 1;
  PG_HIP(hipEventRecord(ev1, c->stream));
  st = sync_ctx(c);
  if (st != PG_OK) return st;
  float ms = 0.0f;
  PG_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  if (stats) {
    *stats = local;
    stats->iterations = static_cast<uint64_t>(res.last_sweep - L.k_begin + 1);
    stats->elapsed_ms = ms;
  }
  if (n_finished) *n_finished = fin;
  return PG_OK;
}

pg_status pg_ctx_take_finished(pg_ctx *c, int *ids, uint32_t capacity, uint32_t *count) {
  if (!c || !count || (!ids && capacity)) return set_error(PG_ERR_INVALID, "pg_ctx_take_finished: null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  const size_t n = std::min<size_t>(capacity, c->finished.size());
  std::copy(c->finished.begin(), c->finished.begin() + n, ids);
  c->finished.erase(c->finished.begin(), c->finished.begin() + n);
  *count = static_cast<uint32_t>(n);
  return PG_OK;
}

pg_status pg_scenario_status(pg_ctx *c, int id, int *state, uint64_t *sweeps) {
  if (!c) return set_error(PG_ERR_INVALID, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  if (id < 0 || static_cast<size_t>(id) >= c->scen.size() || !c->scen[id].used)
    return set_error(PG_ERR_INVALID, "unknown scenario id");
  if (state) *state = c->scen[id].state;
  if (sweeps) *sweeps = c->scen[id].sweeps;
  return PG_OK;
}

pg_status pg_scenario_result(pg_ctx *c, int id, float *time, uint16_t *refs_x, uint16_t *refs_y,
                             uint32_t *boundaries) {
  if (!c) return set_error(PG_ERR_INVALID, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  if (id < 0 || static_cast<size_t>(id) >= c->scen.size() || !c->scen[id].used)
    return set_error(PG_ERR_INVALID, "unknown scenario id");
  const Scenario &s = c->scen[id];
  if (s.state == PG_SCENARIO_FAILED)
    return set_error(PG_ERR_INVALID, "scenario failed (its launch returned an error): no result; release it");
  if (s.state == PG_SCENARIO_RUNNING || !s.posted) return set_error(PG_ERR_BUSY, "scenario still running");
  const size_t n = static_cast<size_t>(s.sc.g.W) * s.sc.g.H;
  if (time) PG_HIP(hipMemcpyAsync(time, s.time, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  if (refs_x) PG_HIP(hipMemcpyAsync(refs_x, s.rx, n * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream));
  if (refs_y) PG_HIP(hipMemcpyAsync(refs_y, s.ry, n * sizeof(uint16_t), hipMemcpyDeviceToHost, c->stream));
  if (boundaries)
    PG_HIP(hipMemcpyAsync(boundaries, s.bits, ((n + 31) / 32) * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  return sync_ctx(c);
}

pg_status pg_scenario_release(pg_ctx *c, int id) {
  if (!c) return set_error(PG_ERR_INVALID, "null context");
  std::lock_guard<std::mutex> lk(c->mu);
  if (id < 0 || static_cast<size_t>(id) >= c->scen.size() || !c->scen[id].used)
    return set_error(PG_ERR_INVALID, "unknown scenario id");
  Scenario &s = c->scen[id];
  const bool running = s.state == PG_SCENARIO_RUNNING;
  c->finished.erase(std::remove(c->finished.begin(), c->finished.end(), id), c->finished.end());
  pg_status st = put_descriptor(c, id, dead_descriptor_ptr());
  if (st != PG_OK) return st;
  // Work on the context's stream is ordered: the block can be handed to a
  // later admission without waiting.
  c->pool.emplace(s.bytes, s.block);
  c->used -= s.bytes;
  const size_t tiles = s.tiles;
  s = Scenario{};
  // Tiles of a running scenario may still be queued for the next sweep:
  // keep the slot (and its queue share) out of use until a launch has
  // consumed them.
  if (running) {
    s.zombie = true;
    s.tiles = tiles;
  }
  return PG_OK;
}

}  // extern "C"
