// Stage P: arrival-time propagation on gfx950.
//
// Replaces pre_propagate / propag / post_propagate
// (crates/propag/src/propag.cu:382-421) and the host loop of
// crates/propag/src/propag.rs:482-557.
//
// Schedule.  Inside one cooperative launch the reference runs exact Jacobi
// sweeps: every cell reads the previous sweep's state, then all cells commit
// (propag.cu:62-109).  This file runs the same sweeps over the whole raster
// until a sweep commits nothing, so the result is a pure function of the
// inputs (no dependence on tiling, workgroup count or timing):
//   * state is double-buffered (time, packed reference) in 8-byte words;
//     sweep k reads st[k&1] and writes st[(k+1)&1];
//   * a cell is re-evaluated in sweep k only if one of its 3x3 cells or the
//     reference of one of its 8 neighbours changed in sweep k-1 (per-cell
//     stamps); otherwise its inputs are those of its last evaluation, which
//     committed nothing, so it would commit nothing again.  A cell that
//     changed in sweep k-1 is always re-evaluated and rewritten, which keeps
//     both buffers equal wherever a cell is skipped;
//   * only tiles that hold or touch a burned cell are visited (an append-only
//     worklist; a tile joins when a neighbour tile first changes).
// Each tile of 16x16 cells is staged with its 1-cell halo in LDS (time,
// reference, the reference's time, change flags).  The boundary bitset used
// by the DDA line-of-sight walks lives in LDS when it fits, else in L2.
//
// Arithmetic.  time_from (propag.cu:269-294) is evaluated in the closed form
//   t = t_from + |d| (1 - e cos(az - az_max)) / (speed_max (1 - e))
// with e cos(az - az_max) = (A*vy + B*vx)/r, A = e cos(az_max),
// B = e sin(az_max), R = 1/(speed_max (1 - e)) prepared once per cell in
// double (own Cody-Waite + Taylor sin/cos) and rounded to f32.  Every step is
// an IEEE add/mul/div/sqrt compiled without contraction, so results match the
// CPU oracle bit for bit.
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pg_internal.h"

namespace pg {
namespace {

constexpr int kTile = 8;                // tile edge (cells): one cell per lane of a wave
constexpr int kHalo = kTile + 2;        // staged edge
constexpr int kHaloCells = kHalo * kHalo;
constexpr uint32_t kNoRef = 0xFFFFFFFFu;
constexpr uint32_t kNoTile = 0xFFFFFFFFu;  // work-list slot not written yet
constexpr int kStampNever = INT_MIN / 2;
constexpr int kWalkBatch = 8;           // DDA cells tested per batch
constexpr int kWave = 64;               // lanes per wavefront

// ------------------------------------------------------------ control block

struct alignas(128) Line {
  unsigned v[32];
};

// Grid-wide state of one propagation; each hot word sits on its own line.
struct Control {
  Line gen;          // [0] barrier release: sweep + 1, bit 31 = finished
  Line live;         // [0] worklist length (grows during a sweep)
  Line book;         // [0] live snapshot (plain form), [1] done, [2] capped, [3] error, [4] last sweep
  Line changes;      // [0..1] tiles changed in sweep k (k & 1), plain form only
  unsigned long long stats[8];  // relaxations, improvements, dda_steps, tiles, work, wait,
                                // critical-path ticks
};

__device__ inline unsigned ld_relaxed(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned add_relaxed(unsigned *p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Data that other workgroups write during a persistent launch (cell state,
// stamps, work list, barrier words) is read and written only with these
// device-coherent accesses (agent-scope relaxed atomics: global loads and
// stores with sc1, which are neither served from nor left dirty in an XCD's
// own L2).  A grid barrier then needs no L2 write-back or invalidation, only
// each wave's wait for its own stores (wait_stores), and read-only data
// (fire coefficients, the boundary bitset) stays cached across sweeps.
__device__ inline uint32_t ld_u32(const uint32_t *p) {
  return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_u32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_i32(const int *p) {
  return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_i32(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline unsigned long long ld_u64(const unsigned long long *p) {
  return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_u64(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint2 ld_state(const uint2 *p) {
  const unsigned long long v = ld_u64(reinterpret_cast<const unsigned long long *>(p));
  return make_uint2(static_cast<uint32_t>(v), static_cast<uint32_t>(v >> 32));
}
__device__ inline void st_state(uint2 *p, uint2 v) {
  st_u64(reinterpret_cast<unsigned long long *>(p),
         static_cast<unsigned long long>(v.x) | (static_cast<unsigned long long>(v.y) << 32));
}
// Returns once every store this wave has issued is complete (for the
// coherent ones: visible device-wide); the clobber keeps the compiler from
// moving memory accesses across it.
__device__ inline void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ------------------------------------------------------------------ geometry

struct Geom {
  int W, H;
  float dx, dy, cell, max_time;
  int north_up, square;
};

// Deterministic double sin/cos: reduction by pi/2 with the two-part constant
// of fdlibm (k*P1 exact for |k| < 2^20) and Taylor polynomials in r^2 to
// 1/19! and 1/20!; only IEEE +,-,*,/ and rint.
__device__ inline void sincos_det(double x, double &s, double &c) {
  const double kTwoOverPi = 0.63661977236758134308;
  const double P1 = 1.5707963267341256;
  const double P2 = 6.0771005065061922e-11;
  const double k = rint(x * kTwoOverPi);
  const double r = (x - k * P1) - k * P2;
  const double r2 = r * r;
  double ps = -1.0 / 121645100408832000.0;
  ps = ps * r2 + 1.0 / 355687428096000.0;
  ps = ps * r2 - 1.0 / 1307674368000.0;
  ps = ps * r2 + 1.0 / 6227020800.0;
  ps = ps * r2 - 1.0 / 39916800.0;
  ps = ps * r2 + 1.0 / 362880.0;
  ps = ps * r2 - 1.0 / 5040.0;
  ps = ps * r2 + 1.0 / 120.0;
  ps = ps * r2 - 1.0 / 6.0;
  ps = ps * r2;
  const double sn = r + r * ps;
  double pc = 1.0 / 2432902008176640000.0;
  pc = pc * r2 - 1.0 / 6402373705728000.0;
  pc = pc * r2 + 1.0 / 20922789888000.0;
  pc = pc * r2 - 1.0 / 87178291200.0;
  pc = pc * r2 + 1.0 / 479001600.0;
  pc = pc * r2 - 1.0 / 3628800.0;
  pc = pc * r2 + 1.0 / 40320.0;
  pc = pc * r2 - 1.0 / 720.0;
  pc = pc * r2 + 1.0 / 24.0;
  pc = pc * r2 - 0.5;
  const double cs = 1.0 + r2 * pc;
  switch (static_cast<long>(k) & 3) {
    case 0: s = sn; c = cs; break;
    case 1: s = cs; c = -sn; break;
    case 2: s = -sn; c = -cs; break;
    default: s = -cs; c = sn; break;
  }
}

// time_from (propag.cu:269-294) in closed form; (fx,fy) -> (tx,ty), fire k.
__device__ inline float time_from(float from_t, int fx, int fy, int tx, int ty, const float4 &k,
                                  const Geom &g) {
  const int ix = tx - fx;
  const int iy = ty - fy;
  const float ux = static_cast<float>(ix);
  const float uy = static_cast<float>(g.north_up ? -iy : iy);
  const float r = sqrtf(ux * ux + uy * uy);
  const float proj = k.x * uy + k.y * ux;
  if (g.square) return from_t + ((r - proj) * g.cell) * k.z;
  const float ex = ux * g.dx;
  const float ey = static_cast<float>(iy) * g.dy;
  const float dist = sqrtf(ex * ex + ey * ey);
  return from_t + (dist * (1.0f - proj / r)) * k.z;
}

// Reference words.  With both raster sides below 32768 (the usual case) a
// reference is x | y << 16 with 15-bit coordinates and bit 15 caching
// whether the DDA line from the cell to its reference is free of boundary
// cells.  That walk (propag.cu:164-176) depends only on the pair (cell,
// reference) and the static bitset, so it is computed once when a cell
// commits a new reference instead of by every neighbour that evaluates it.
// Larger rasters use full 16-bit coordinates and walk at evaluation time.
constexpr uint32_t kClearBit = 0x8000u;

template <bool kCache>
__device__ inline int ref_x(uint32_t r) { return static_cast<int>(r & (kCache ? 0x7fffu : 0xffffu)); }
template <bool kCache>
__device__ inline int ref_y(uint32_t r) { return static_cast<int>((r >> 16) & (kCache ? 0x7fffu : 0xffffu)); }
__device__ inline uint32_t make_ref(int x, int y) {
  return static_cast<uint32_t>(x) | (static_cast<uint32_t>(y) << 16);
}

// Dynamic LDS of the persistent kernel: per-wave tile slices, then the
// boundary bitset when it fits.
extern __shared__ uint32_t dyn_lds[];

template <bool kLdsBits>
__device__ inline bool bit_at(const uint32_t *gbits, unsigned lds_off, uint32_t idx) {
  const uint32_t w = kLdsBits ? dyn_lds[lds_off + (idx >> 5)] : gbits[idx >> 5];
  return (w >> (idx & 31u)) & 1u;
}

// (double)d < 1e-9 for a float d >= 0 is exactly d <= this float (the float
// nearest 1e-9 lies below it and the next one above).
constexpr float kTie = 0x1.12e0bep-30f;

// True if a boundary cell lies on the f32 DDA of propag.h:91-135 from
// (fx,fy) to (tx,ty), both included.  ru, rv count the steps left along each
// axis: the reference's "cur is still between from and to" (propag.h:117-120)
// is ru >= 0 && rv >= 0, i.e. (ru | rv) >= 0, and once the walk has visited
// `to` or stepped past it one of them is negative for good.  With
// d = tmax_x - tmax_y (exact whenever |d| is near the tie threshold, the two
// being within a factor of two of each other), the reference's branches
// become: x advances iff d <= kTie (tie or tmax_x < tmax_y), y advances iff
// not d < -kTie (tie or tmax_x >= tmax_y; also for from == to, where d is
// NaN, so that walk visits its one cell and ends).  Cells are tested
// kWalkBatch at a time so a batch's bit loads are in flight together;
// `steps` counts the cells up to and including the first boundary, as a
// serial walk would.
template <bool kLdsBits>
__device__ bool walk_blocked(int fx, int fy, int tx, int ty, const uint32_t *gbits, unsigned lds_off,
                             int W, uint32_t &steps) {
  const int ddx = tx - fx, ddy = ty - fy;
  int ru = abs(ddx), rv = abs(ddy);
  const int ix = (ddx > 0) - (ddx < 0);
  const int iy = ((ddy > 0) - (ddy < 0)) * W;
  // 1.0 / float(abs(..)) is a double division narrowed to float
  const float dlx = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(ru)));
  const float dly = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(rv)));
  float tmx = dlx, tmy = dly;
  int idx = fx + fy * W;
  for (;;) {
    uint32_t id[kWalkBatch];
    int live[kWalkBatch];
#pragma unroll
    for (int b = 0; b < kWalkBatch; ++b) {
      live[b] = ru | rv;
      id[b] = live[b] >= 0 ? static_cast<uint32_t>(idx) : 0u;
      const float d = tmx - tmy;
      const bool xs = d <= kTie, ys = !(d < -kTie);
      ru -= xs ? 1 : 0;
      rv -= ys ? 1 : 0;
      idx += (xs ? ix : 0) + (ys ? iy : 0);
      tmx = xs ? tmx + dlx : tmx;
      tmy = ys ? tmy + dly : tmy;
    }
    uint32_t hits = 0;
#pragma unroll
    for (int b = 0; b < kWalkBatch; ++b)
      hits |= (bit_at<kLdsBits>(gbits, lds_off, id[b]) && live[b] >= 0) ? (1u << b) : 0u;
    if (hits) {
      steps += static_cast<uint32_t>(__builtin_ctz(hits)) + 1u;
      return true;
    }
    if ((ru | rv) < 0) {
#pragma unroll
      for (int b = 0; b < kWalkBatch; ++b) steps += live[b] >= 0 ? 1u : 0u;
      return false;
    }
    steps += kWalkBatch;
  }
}

// Line-of-sight flag for a committed reference: the DDA from the cell to an
// adjacent reference visits exactly those two cells.
template <bool kLdsBits>
__device__ inline bool line_clear(int x, int y, int rx, int ry, const uint32_t *gbits, unsigned lds_off,
                                  int W, uint32_t &steps) {
  if (rx == x && ry == y) return false;  // an ignition's own reference: never walked
  if (abs(rx - x) <= 1 && abs(ry - y) <= 1) {
    steps += 2;
    const uint32_t a = static_cast<uint32_t>(x) + static_cast<uint32_t>(y) * static_cast<uint32_t>(W);
    const uint32_t b = static_cast<uint32_t>(rx) + static_cast<uint32_t>(ry) * static_cast<uint32_t>(W);
    return !bit_at<kLdsBits>(gbits, lds_off, a) && !bit_at<kLdsBits>(gbits, lds_off, b);
  }
  return !walk_blocked<kLdsBits>(x, y, rx, ry, gbits, lds_off, W, steps);
}

__device__ inline uint2 pack_state(float t, uint32_t ref) { return make_uint2(__float_as_uint(t), ref); }

struct Args {
  Geom g;
  const float4 *coef;        // A, B, R, null flag
  uint2 *st[2];              // double-buffered (time bits, reference word)
  int *stamp;                // last sweep in which the cell changed
  const uint32_t *bits;      // boundary bitset (global copy)
  uint32_t *tile_live;       // 1 once a tile is on the worklist
  uint32_t *worklist;        // live tile ids
  Control *ctl;
  struct Slot *slots;        // grid-barrier arrival lines, one per workgroup
  int ntx, nty;              // tiles per row / column
  unsigned max_iters;
  unsigned spin_limit;
  int lds_bits;              // bitset staged in LDS
  int cache;                 // line-of-sight flags cached in reference words
  unsigned long long *prof;  // [sweep][kProfCols] profile ticks, or null
  unsigned prof_sweeps;      // rows of prof
};

struct TileLds {  // one wave's staged tile (16x16 + 1-cell halo)
  float time[kHaloCells];
  uint32_t ref[kHaloCells];
  float rtime[kHaloCells];
  uint8_t chg[kHaloCells];
};
static_assert(sizeof(TileLds) % 4 == 0, "TileLds must keep 4-byte alignment");

// Per-workgroup arrival slot of the grid barrier, one 128-byte line each:
// arrive = (sweep + 1) << 32 | changed tiles, work = ticks spent in the sweep.
struct alignas(128) Slot {
  unsigned long long arrive;
  unsigned long long work;
  unsigned long long ph[4];  // profiling: slowest wave's phase ticks
  unsigned long long t_seen, t_arrive;  // profiling: sweep start / arrival
  unsigned pad[16];
};
static_assert(sizeof(Slot) == 128, "one slot per 128-byte line");

// Profile columns per sweep: slowest wave per phase (staging, update rule,
// line-of-sight walks, commit and work-list), then the barrier: release of
// the previous sweep -> last workgroup starting, the longest workgroup span,
// last arrival -> master has seen all, master bookkeeping.
constexpr int kProfCols = 8;

struct Counters {
  uint32_t relax, improve, steps, tiles;
  uint64_t work_ticks, wait_ticks;  // wall_clock64 ticks (thread 0 of each workgroup)
  uint64_t ph[4];                   // profiling: this wave's phase ticks in the sweep
};

__device__ inline void wave_sync() {
  // Lanes of one wave exchange data through LDS: order the accesses.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Marks the 8 neighbour tiles of (tx,ty) live, appending newcomers.
__device__ inline void grow_worklist(const Args &a, int tx, int ty, int lane) {
  if (lane < 9 && lane != 4) {
    const int nx = tx + lane % 3 - 1, ny = ty + lane / 3 - 1;
    if (nx >= 0 && ny >= 0 && nx < a.ntx && ny < a.nty) {
      const int t = nx + ny * a.ntx;
      if (__hip_atomic_load(&a.tile_live[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
        unsigned expected = 0u;
        if (__hip_atomic_compare_exchange_strong(&a.tile_live[t], &expected, 1u, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          const unsigned slot = add_relaxed(&a.ctl->live.v[0], 1u);
          st_u32(&a.worklist[slot], static_cast<uint32_t>(t));
        }
      }
    }
  }
}

// One sweep over one tile by one wave: lane l owns cell (l % 8, l / 8) of
// the 8x8 tile.  Returns, uniformly over the wave, whether any cell
// committed an improvement.
template <bool kCache, bool kLdsBits>
__device__ bool wave_tile(const Args &a, unsigned lds_off, int tile, int k, TileLds &s, Counters &cnt) {
  const Geom &g = a.g;
  const uint2 *__restrict__ st_in = a.st[k & 1];
  uint2 *__restrict__ st_out = a.st[(k + 1) & 1];
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  const int ox = tx * kTile, oy = ty * kTile;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t tp0 = a.prof ? wall_clock64() : 0;

  // Stage the halo tile (load_points_into_shared_memory, propag.cu:231-261)
  // in two rounds of independent loads: the cells' state words and stamps,
  // then their references' times and stamps (load_point, propag.cu:209-229).
  constexpr int kPer = (kHaloCells + kWave - 1) / kWave;  // halo cells per lane
  uint2 w[kPer];
  int stp[kPer];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int h = lane + it * kWave;
    const int gx = ox + h % kHalo - 1, gy = oy + h / kHalo - 1;
    const bool inb = h < kHaloCells && gx >= 0 && gy >= 0 && gx < g.W && gy < g.H;
    const size_t gi = inb ? static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W : 0;
    w[it] = inb ? ld_state(&st_in[gi]) : make_uint2(__float_as_uint(FLT_MAX), kNoRef);
    stp[it] = inb ? ld_i32(&a.stamp[gi]) : kStampNever;
  }
  float rt[kPer];
  int rstp[kPer];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const float t = __uint_as_float(w[it].x);
    // A neighbour only feeds the update rule while it burns before
    // max_time; only then is its reference's live time needed.
    const bool gather = t < g.max_time && t < FLT_MAX;
    const size_t ri =
        gather ? static_cast<size_t>(ref_x<kCache>(w[it].y)) + static_cast<size_t>(ref_y<kCache>(w[it].y)) * g.W : 0;
    rt[it] = gather ? __uint_as_float(ld_u32(&st_in[ri].x)) : FLT_MAX;
    rstp[it] = gather ? ld_i32(&a.stamp[ri]) : kStampNever;
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int h = lane + it * kWave;
    if (h < kHaloCells) {
      s.time[h] = __uint_as_float(w[it].x);
      s.ref[h] = w[it].y;
      s.rtime[h] = rt[it];
      s.chg[h] = (stp[it] >= k - 1 || rstp[it] >= k - 1) ? 1 : 0;
    }
  }
  wave_sync();
  uint64_t tp1 = a.prof ? wall_clock64() : 0;
  uint64_t walk_ticks = 0;

  const int lx = (lane & (kTile - 1)) + 1, ly = lane / kTile + 1;
  const int gx = ox + lx - 1, gy = oy + ly - 1;
  bool need = false;
  if (gx < g.W && gy < g.H) {
#pragma unroll
    for (int j = -1; j < 2; ++j)
#pragma unroll
      for (int i = -1; i < 2; ++i) need = need || s.chg[(lx + i) + (ly + j) * kHalo];
  }
  bool improved = false;
  if (need) {
    ++cnt.relax;
    float4 win[9];  // fire coefficients of the 3x3 neighbourhood
#pragma unroll
    for (int j = -1; j < 2; ++j) {
#pragma unroll
      for (int i = -1; i < 2; ++i) {
        const int X = gx + i, Y = gy + j;
        win[(j + 1) * 3 + (i + 1)] = (X >= 0 && Y >= 0 && X < g.W && Y < g.H)
                                         ? a.coef[static_cast<size_t>(X) + static_cast<size_t>(Y) * g.W]
                                         : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
      }
    }
    const int me = lx + ly * kHalo;
    const float own_t = s.time[me];
    float best_t = own_t;
    uint32_t best_ref = s.ref[me];
    // find_neighbor_which_reaches_first (propag.cu:124-190): j outer, i
    // inner, strict < so the first of equal candidates wins.
#pragma unroll
    for (int j = -1; j < 2; ++j) {
#pragma unroll
      for (int i = -1; i < 2; ++i) {
        if (i == 0 && j == 0) continue;
        const int nx = gx + i, ny = gy + j;
        if (nx < 0 || ny < 0 || nx >= g.W || ny >= g.H) continue;
        const int nl = (lx + i) + (ly + j) * kHalo;
        const float n_t = s.time[nl];
        if (!(n_t < g.max_time)) continue;
        const float4 kf = win[(j + 1) * 3 + (i + 1)];
        if (kf.w != 0.0f) continue;  // fire_is_null
        const uint32_t ref = s.ref[nl];
        const int rfx = ref_x<kCache>(ref), rfy = ref_y<kCache>(ref);
        if (rfx == gx && rfy == gy) continue;
        const float nt = time_from(n_t, nx, ny, gx, gy, kf, g);
        const float ct = time_from(s.rtime[nl], rfx, rfy, gx, gy, kf, g);
        if (ct < nt && ct < best_t) {
          bool blocked;
          if (kCache)
            blocked = (ref & kClearBit) == 0u;
          else
            blocked = walk_blocked<kLdsBits>(nx, ny, rfx, rfy, a.bits, lds_off, g.W, cnt.steps);
          const uint32_t cand_ref = blocked ? make_ref(nx, ny) : make_ref(rfx, rfy);
          const float cand_t = blocked ? nt : ct;
          if (cand_t < best_t) {
            best_t = cand_t;
            best_ref = cand_ref;
          }
        } else if (nt < best_t) {
          best_t = nt;
          best_ref = make_ref(nx, ny);
        }
      }
    }
    const size_t gi = static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W;
    improved = best_t < own_t;  // update_point (propag.cu:192-200)
    if (improved) {
      const uint64_t tw = a.prof ? wall_clock64() : 0;
      if (kCache) {
        // The flag depends only on (cell, reference): an improvement that
        // keeps the reference keeps its flag.
        const uint32_t old = s.ref[me];
        if ((old & ~kClearBit) == best_ref)
          best_ref = old;
        else if (line_clear<kLdsBits>(gx, gy, ref_x<true>(best_ref), ref_y<true>(best_ref), a.bits, lds_off, g.W,
                                      cnt.steps))
          best_ref |= kClearBit;
      }
      if (a.prof) walk_ticks += wall_clock64() - tw;
      st_state(&st_out[gi], pack_state(best_t, best_ref));
      st_i32(&a.stamp[gi], k);
      ++cnt.improve;
    } else {
      st_state(&st_out[gi], pack_state(own_t, s.ref[me]));
    }
  }
  uint64_t tp2 = a.prof ? wall_clock64() : 0;
  const bool changed = __ballot(improved) != 0ull;
  if (changed) grow_worklist(a, tx, ty, lane);
  if (lane == 0) ++cnt.tiles;
  wave_sync();  // the next tile reuses this wave's LDS slice
  if (a.prof) {
    // per wave: staging, rule (minus walks), walks (slowest lane), tail
    unsigned long long wk = walk_ticks;
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(wk, off);
      wk = o > wk ? o : wk;
    }
    const uint64_t tp3 = wall_clock64();
    cnt.ph[0] += tp1 - tp0;
    cnt.ph[1] += (tp2 - tp1) > wk ? (tp2 - tp1) - wk : 0;
    cnt.ph[2] += wk;
    cnt.ph[3] += tp3 - tp2;
  }
  return changed;
}

// Grid barrier.  Callers have waited for their stores (wait_stores) and
// passed __syncthreads.  Every workgroup publishes one 64-bit arrival word in
// its own line: (sweep + 1) << 32 | changed << 31 | work ticks.  Wave 0 of
// workgroup 0 polls all lines, a batch of loads at a time, folds the
// quiet-sweep test, the iteration cap and the critical-path time, and
// releases the generation word, (sweep + 1) | finished << 31, which thread 0
// of every other workgroup polls.  All words are device-coherent, so no
// cache maintenance is needed.  Returns 1 when the propagation is finished.
__device__ int grid_barrier(const Args &a, int k, bool wg_changed, uint64_t sweep_t0, Counters &cnt,
                            const unsigned long long *wg_ph) {
  __shared__ int s_done;
  __shared__ unsigned long long s_release;  // master: time of the last release
  const unsigned target = static_cast<unsigned>(k) + 1u;
  Slot *slots = a.slots;
  Control *c = a.ctl;
  uint64_t t_arrive = 0;
  if (threadIdx.x == 0) {
    t_arrive = wall_clock64();
    const uint64_t work = t_arrive - sweep_t0;
    cnt.work_ticks += work;
    if (a.prof) {
      for (int q = 0; q < 4; ++q) st_u64(&slots[blockIdx.x].ph[q], wg_ph[q]);
      st_u64(&slots[blockIdx.x].t_seen, sweep_t0);
      st_u64(&slots[blockIdx.x].t_arrive, t_arrive);
      wait_stores();
    }
    const uint64_t w31 = work < 0x7fffffffull ? work : 0x7fffffffull;
    st_u64(&slots[blockIdx.x].arrive,
           (static_cast<unsigned long long>(target) << 32) | (wg_changed ? (1ull << 31) : 0ull) | w31);
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x < kWave) {
      const unsigned lane = threadIdx.x;
      const unsigned G = gridDim.x;
      const unsigned nq = (G + kWave - 1) / kWave;  // slots per lane (at most 64)
      unsigned long long pending = 0;
      for (unsigned q = 0; q < nq; ++q)
        if (lane + q * kWave < G) pending |= 1ull << q;
      bool changed = false, timeout = false;
      unsigned long long maxwork = 0;
      unsigned spins = 0;
      while (pending) {
        for (unsigned q0 = 0; q0 < nq; q0 += 4) {
          unsigned long long v[4];
#pragma unroll
          for (unsigned j = 0; j < 4; ++j) {
            const unsigned q = q0 + j;
            v[j] = (q < nq && ((pending >> q) & 1ull)) ? ld_u64(&slots[lane + q * kWave].arrive) : 0ull;
          }
#pragma unroll
          for (unsigned j = 0; j < 4; ++j) {
            const unsigned q = q0 + j;
            if (q < nq && ((pending >> q) & 1ull) && (v[j] >> 32) == target) {
              pending &= ~(1ull << q);
              changed = changed || ((v[j] >> 31) & 1ull);
              const unsigned long long w = v[j] & 0x7fffffffull;
              maxwork = w > maxwork ? w : maxwork;
            }
          }
        }
        if (pending) {
          __builtin_amdgcn_s_sleep(1);
          if (++spins > a.spin_limit) {
            timeout = true;
            break;
          }
        }
      }
      asm volatile("" ::: "memory");
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(maxwork, off);
        maxwork = o > maxwork ? o : maxwork;
      }
      changed = __ballot(changed) != 0ull;
      timeout = __ballot(timeout) != 0ull;
      if (a.prof && static_cast<unsigned>(k) < a.prof_sweeps) {
        const unsigned long long t_all = wall_clock64();
        // k == 0: the kernel start of workgroup 0 stands in for a release
        const unsigned long long rel = k == 0 ? __shfl(sweep_t0, 0) : s_release;
        unsigned long long m[7] = {0, 0, 0, 0, 0, 0, 0};  // ph[4], start, span, arrival
        for (unsigned i = lane; i < G; i += kWave) {
          for (int q = 0; q < 4; ++q) m[q] = max(m[q], ld_u64(&slots[i].ph[q]));
          const unsigned long long ts = ld_u64(&slots[i].t_seen), ta = ld_u64(&slots[i].t_arrive);
          m[4] = max(m[4], ts > rel ? ts - rel : 0ull);
          m[5] = max(m[5], ta - ts);
          m[6] = max(m[6], ta);
        }
        for (int q = 0; q < 7; ++q)
          for (int off = 32; off > 0; off >>= 1) m[q] = max(m[q], __shfl_xor(m[q], off));
        if (lane == 0) {
          unsigned long long *row = a.prof + static_cast<size_t>(k) * kProfCols;
          for (int q = 0; q < 6; ++q) row[q] = m[q];
          row[6] = t_all > m[6] ? t_all - m[6] : 0ull;
          row[7] = t_all;  // completed below with the release time
        }
      }
      if (lane == 0) {
        const bool cap = target >= a.max_iters;
        const bool done = !changed || cap || timeout;
        if (cap && changed) st_relaxed(&c->book.v[2], 1u);
        if (timeout) st_relaxed(&c->book.v[3], 1u);
        if (done) st_relaxed(&c->book.v[4], static_cast<unsigned>(k));
        c->stats[6] += maxwork;  // only this lane ever writes it
        if (a.prof) {
          s_release = wall_clock64();
          if (static_cast<unsigned>(k) < a.prof_sweeps) {
            unsigned long long *row = a.prof + static_cast<size_t>(k) * kProfCols;
            row[7] = s_release - row[7];
          }
        }
        st_relaxed(&c->gen.v[0], target | (done ? 0x80000000u : 0u));
        s_done = done ? 1 : 0;
      }
    }
  } else if (threadIdx.x == 0) {
    unsigned spins = 0, g;
    bool timeout = false;
    while (((g = ld_relaxed(&c->gen.v[0])) & 0x7fffffffu) != target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > a.spin_limit) {
        timeout = true;
        st_relaxed(&c->book.v[3], 1u);
        break;
      }
    }
    asm volatile("" ::: "memory");
    s_done = (timeout || (g >> 31) != 0u) ? 1 : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) cnt.wait_ticks += wall_clock64() - t_arrive;
  return s_done;
}

__device__ inline void flush_counters(const Args &a, const Counters &cnt) {
  // One 64-bit add per counter per wave.
  unsigned long long v[6] = {cnt.relax, cnt.improve, cnt.steps, cnt.tiles, cnt.work_ticks,
                             cnt.wait_ticks};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    unsigned long long x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(&a.ctl->stats[q], x);
  }
}

// Persistent form: one launch runs every sweep; workgroups meet at the grid
// barrier after each sweep.  Work-list entry w belongs to wave w / gridDim.x
// of workgroup w % gridDim.x, so tiles found together (neighbours on the
// fire front) land on different CUs.  Slots fill in order and never change
// once written; a wave keeps its first 64 tile ids in one register (lane i:
// its i-th entry) and each sweep probes the slots after them.  A slot still
// unwritten when probed belongs to a tile added during this sweep, which
// this sweep need not visit (visiting it anyway would only recompute
// unchanged cells).  Dynamic LDS: one TileLds per wave, then (optionally)
// the boundary bitset.
template <bool kCache, bool kLdsBits>
__global__ __launch_bounds__(1024) void propagate_persistent(Args a) {
  // Per-sweep workgroup flags, double-buffered by sweep parity: sweep k
  // writes [k & 1], thread 0 clears it after barrier k, and it is written
  // again only in sweep k + 2, after barrier k + 1.
  __shared__ unsigned s_changed[2];
  __shared__ unsigned long long s_ph[2][4];
  const int nw = blockDim.x / kWave, wave = threadIdx.x / kWave;
  const unsigned lane = threadIdx.x & (kWave - 1);
  TileLds *slices = reinterpret_cast<TileLds *>(dyn_lds);
  TileLds &s = slices[wave];
  const unsigned lds_off = nw * static_cast<unsigned>(sizeof(TileLds) / 4);
  if (threadIdx.x < 2) s_changed[threadIdx.x] = 0;
  if (threadIdx.x < 8) s_ph[threadIdx.x / 4][threadIdx.x % 4] = 0;
  if (kLdsBits) {
    const unsigned nwords = (static_cast<unsigned>(a.g.W) * static_cast<unsigned>(a.g.H) + 31u) / 32u;
    for (unsigned i = threadIdx.x; i < nwords; i += blockDim.x) dyn_lds[lds_off + i] = a.bits[i];
  }
  __syncthreads();
  Counters cnt{0, 0, 0, 0, 0, 0, {0, 0, 0, 0}};
  uint64_t t0 = threadIdx.x == 0 ? wall_clock64() : 0;
  const unsigned gwave = wave * gridDim.x + blockIdx.x, nwaves = gridDim.x * nw;
  const size_t ntiles = static_cast<size_t>(a.ntx) * a.nty;
  uint32_t ids = kNoTile;  // lane i: this wave's i-th work-list entry
  unsigned known = 0;      // entries found so far (at most 64 cached)
  for (int k = 0;; ++k) {
    uint32_t probe = kNoTile;
    const size_t wl = gwave + static_cast<size_t>(lane) * nwaves;
    if (lane >= known && wl < ntiles) probe = ld_u32(&a.worklist[wl]);
    bool changed = false;
    for (unsigned i = 0; i < known; ++i)
      changed |= wave_tile<kCache, kLdsBits>(a, lds_off, static_cast<int>(__shfl(ids, static_cast<int>(i))), k, s,
                                             cnt);
    if (known < 64u) {
      // new entries: the run of written slots right after the known ones
      const unsigned long long r = ~(__ballot(probe != kNoTile) >> known);
      unsigned nnew = r ? static_cast<unsigned>(__builtin_ctzll(r)) : 64u - known;
      if (nnew > 64u - known) nnew = 64u - known;
      if (lane >= known && lane < known + nnew) ids = probe;
      for (unsigned i = known; i < known + nnew; ++i)
        changed |= wave_tile<kCache, kLdsBits>(a, lds_off, static_cast<int>(__shfl(ids, static_cast<int>(i))), k,
                                               s, cnt);
      known += nnew;
    }
    if (known == 64u) {
      // entries beyond the cached ones are read every sweep
      for (size_t w = gwave + 64ull * nwaves; w < ntiles; w += nwaves) {
        const uint32_t t = ld_u32(&a.worklist[w]);
        if (t == kNoTile) break;
        changed |= wave_tile<kCache, kLdsBits>(a, lds_off, static_cast<int>(t), k, s, cnt);
      }
    }
    if (lane == 0 && changed) s_changed[k & 1] = 1u;
    if (a.prof && lane == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        atomicMax(&s_ph[k & 1][q], static_cast<unsigned long long>(cnt.ph[q]));
        cnt.ph[q] = 0;
      }
    }
    wait_stores();
    __syncthreads();
    const int done = grid_barrier(a, k, s_changed[k & 1] != 0u, t0, cnt, s_ph[k & 1]);
    if (threadIdx.x == 0) {
      s_changed[k & 1] = 0;
      t0 = wall_clock64();
    }
    if (threadIdx.x < 4) s_ph[k & 1][threadIdx.x] = 0;
    if (done) break;
  }
  flush_counters(a, cnt);
}

// Plain form: one launch per sweep (the reference's relaunch structure).
template <bool kCache>
__global__ __launch_bounds__(256) void propagate_sweep(Args a, int k) {
  __shared__ TileLds slices[256 / kWave];
  __shared__ unsigned s_changes;
  const int wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_changes = 0;
  __syncthreads();
  Counters cnt{0, 0, 0, 0, 0, 0, {0, 0, 0, 0}};
  const unsigned n_live = ld_relaxed(&a.ctl->book.v[0]);
  const unsigned nw = blockDim.x / kWave;
  unsigned my_changes = 0;
  for (unsigned w = blockIdx.x * nw + wave; w < n_live; w += gridDim.x * nw)
    my_changes += wave_tile<kCache, false>(a, 0, static_cast<int>(a.worklist[w]), k, slices[wave], cnt) ? 1u : 0u;
  if ((threadIdx.x & (kWave - 1)) == 0 && my_changes) atomicAdd(&s_changes, my_changes);
  __syncthreads();
  if (threadIdx.x == 0 && s_changes) add_relaxed(&a.ctl->changes.v[k & 1], s_changes);
  flush_counters(a, cnt);
}

// Between plain sweeps: the bookkeeping the persistent barrier does.
__global__ void sweep_book(Args a, int k) {
  Control *c = a.ctl;
  const unsigned ch = c->changes.v[k & 1];
  c->changes.v[(k + 1) & 1] = 0u;
  c->book.v[0] = c->live.v[0];
  const bool cap = static_cast<unsigned>(k + 1) >= a.max_iters;
  const bool done = ch == 0u || cap;
  if (cap && ch != 0u) c->book.v[2] = 1u;
  if (done) c->book.v[4] = static_cast<unsigned>(k);
  c->book.v[1] = done ? 1u : 0u;
}

// ---------------------------------------------------------------- set-up

__global__ void fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Fire coefficients, packed double-buffered state, stamps, reference
// validation, the line-of-sight flags of the initial references (needs the
// boundary bitset, so it runs after boundaries_kernel) and the initial
// worklist (tiles holding or touching a burned cell).
__global__ void prep(Args a, const float *speed, const float *az, const float *ecc,
                     const float *time, const uint16_t *rx, const uint16_t *ry) {
  const size_t n = static_cast<size_t>(a.g.W) * a.g.H;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const float s = speed[i], e = ecc[i];
    double sn, cs;
    sincos_det(static_cast<double>(az[i]), sn, cs);
    float4 k;
    k.x = static_cast<float>(static_cast<double>(e) * cs);
    k.y = static_cast<float>(static_cast<double>(e) * sn);
    k.z = static_cast<float>(1.0 / (static_cast<double>(s) * (1.0 - static_cast<double>(e))));
    k.w = s == 0.0f ? 1.0f : 0.0f;
    const_cast<float4 *>(a.coef)[i] = k;
    const float t = time[i];
    uint32_t ref = static_cast<uint32_t>(rx[i]) | (static_cast<uint32_t>(ry[i]) << 16);
    const bool burned = t < FLT_MAX;
    const int x = static_cast<int>(i % a.g.W), y = static_cast<int>(i / a.g.W);
    if (burned && (rx[i] >= a.g.W || ry[i] >= a.g.H)) {
      atomicOr(&a.ctl->book.v[3], 2u);  // invalid reference
      a.st[0][i] = pack_state(FLT_MAX, kNoRef);
      a.st[1][i] = pack_state(FLT_MAX, kNoRef);
      a.stamp[i] = kStampNever;
      continue;
    }
    if (burned && a.cache) {
      uint32_t steps = 0;
      if (line_clear<false>(x, y, rx[i], ry[i], a.bits, 0, a.g.W, steps)) ref |= kClearBit;
    }
    a.st[0][i] = pack_state(t, ref);
    a.st[1][i] = pack_state(t, ref);
    a.stamp[i] = burned ? -1 : kStampNever;
    if (burned) {
      const int tx = x / kTile, ty = y / kTile;
      for (int q = 0; q < 9; ++q) {
        const int nx = tx + q % 3 - 1, ny = ty + q / 3 - 1;
        if (nx < 0 || ny < 0 || nx >= a.ntx || ny >= a.nty) continue;
        const int t2 = nx + ny * a.ntx;
        if (atomicCAS(&a.tile_live[t2], 0u, 1u) == 0u) a.worklist[atomicAdd(&a.ctl->live.v[0], 1u)] = t2;
      }
    }
  }
}

__global__ void init_control(Control *c) {
  c->book.v[0] = c->live.v[0];
}

// pre_propagate's find_boundaries (propag.cu:345-375): a cell is a boundary
// if an in-bounds neighbour has a dissimilar fire that is faster.  One wave
// covers 64 consecutive cells = two bitset words, so no atomics are needed.
__global__ __launch_bounds__(256) void boundaries_kernel(Geom g, const float *speed, const float *az,
                                                        const float *ecc, uint32_t *bits) {
  const size_t n = static_cast<size_t>(g.W) * g.H;
  const size_t nwords = (n + 31) / 32;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t base = static_cast<size_t>(blockIdx.x) * blockDim.x; base < nwords * 32; base += stride) {
    const size_t i = base + threadIdx.x;
    bool changed = false;
    if (i < n) {
      const int x = static_cast<int>(i % g.W), y = static_cast<int>(i / g.W);
      const float s0 = speed[i], a0 = az[i], e0 = ecc[i];
      for (int di = -1; di < 2; ++di) {
        for (int dj = -1; dj < 2; ++dj) {
          if (di == 0 && dj == 0) continue;
          const int nx = x + di, ny = y + dj;
          if (nx < 0 || ny < 0 || nx >= g.W || ny >= g.H) continue;
          const size_t j = static_cast<size_t>(nx) + static_cast<size_t>(ny) * g.W;
          const float s1 = speed[j], a1 = az[j], e1 = ecc[j];
          const bool similar = static_cast<double>(fabsf(s1 - s0)) < 0.1 &&
                               static_cast<double>(fabsf(a1 - a0)) < (5.0 * (2 * 3.14159265358979323846) / 360.0) &&
                               static_cast<double>(fabsf(e1 - e0)) < 0.05;
          changed = changed || (!similar && s0 < s1);
        }
      }
    }
    const unsigned long long m = __ballot(changed);
    const int lane = threadIdx.x & 63;
    const size_t w0 = (base + (threadIdx.x & ~63u)) / 32;
    if (lane == 0 && w0 < nwords) bits[w0] = static_cast<uint32_t>(m);
    if (lane == 32 && w0 + 1 < nwords) bits[w0 + 1] = static_cast<uint32_t>(m >> 32);
  }
}

// post_propagate (propag.cu:110-117) plus unpacking into the caller's rasters.
__global__ void post(Args a, int final_buf, const float *speed, float *time, uint16_t *rx, uint16_t *ry) {
  const size_t n = static_cast<size_t>(a.g.W) * a.g.H;
  const uint2 *st = a.st[final_buf];
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const uint2 w = st[i];
    float t = __uint_as_float(w.x);
    // A cell never reached keeps the caller's references untouched.
    if (t < FLT_MAX) {
      rx[i] = static_cast<uint16_t>(a.cache ? ref_x<true>(w.y) : ref_x<false>(w.y));
      ry[i] = static_cast<uint16_t>(a.cache ? ref_y<true>(w.y) : ref_y<false>(w.y));
    }
    if (speed[i] == 0.0f || t > a.g.max_time) t = FLT_MAX;
    time[i] = t;
  }
}

// ------------------------------------------------------------- workspace

struct Workspace {
  void *base = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::vector<Workspace> g_ws;  // per device

pg_status workspace(int device, size_t bytes, void **out) {
  std::lock_guard<std::mutex> lk(g_ws_mutex);
  if (g_ws.size() <= static_cast<size_t>(device)) g_ws.resize(device + 1);
  Workspace &w = g_ws[device];
  if (w.bytes < bytes) {
    if (w.base) PG_HIP(hipFree(w.base));
    w.base = nullptr;
    w.bytes = 0;
    hipError_t e = hipMalloc(&w.base, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return set_error(PG_ERR_OUT_OF_MEMORY, "workspace allocation of " + std::to_string(bytes) + " bytes failed");
    }
    w.bytes = bytes;
  }
  *out = w.base;
  return PG_OK;
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

pg_status run_propagate(const DeviceInfo *dev, const Settings *settings, const float *speed,
                        const float *az, const float *ecc, float *time, uint16_t *refs_x,
                        uint16_t *refs_y, uint32_t *boundaries, const pg_tuning *tuning,
                        pg_stats *stats, hipStream_t stream) {
  if (!settings) return set_error(PG_ERR_INVALID, "settings is NULL");
  const uint32_t W = settings->geo_ref.width, H = settings->geo_ref.height;
  if (W == 0 || H == 0) return set_error(PG_ERR_INVALID, "raster has no cells");
  if (W > 0xFFFFu || H > 0xFFFFu)
    return set_error(PG_ERR_INVALID, "raster sides must be below 65536 (16-bit references)");
  if (!speed || !az || !ecc || !time || !refs_x || !refs_y || !boundaries)
    return set_error(PG_ERR_INVALID, "null raster pointer");
  // A non-finite max_time would let unburned cells (FLT_MAX, no reference)
  // count as burning neighbours.
  if (!std::isfinite(settings->max_time))
    return set_error(PG_ERR_INVALID, "max_time must be finite");
  pg_tuning tn{};
  if (tuning) tn = *tuning;
  if (tn.tile_w && tn.tile_w != kTile) return set_error(PG_ERR_INVALID, "tile_w must be 8");
  if (tn.tile_h && tn.tile_h != kTile) return set_error(PG_ERR_INVALID, "tile_h must be 8");
  const float gdx = settings->geo_ref.transform.gt.dx, gdy = settings->geo_ref.transform.gt.dy;

  Args a{};
  a.g.W = static_cast<int>(W);
  a.g.H = static_cast<int>(H);
  a.g.dx = gdx;
  a.g.dy = gdy;
  a.g.cell = fabsf(gdx);
  a.g.max_time = settings->max_time;
  a.g.north_up = gdy < 0.0f;
  a.g.square = fabsf(gdx) == fabsf(gdy);
  a.cache = (W <= 0x7fffu && H <= 0x7fffu) ? 1 : 0;
  a.ntx = static_cast<int>((W + kTile - 1) / kTile);
  a.nty = static_cast<int>((H + kTile - 1) / kTile);
  a.max_iters = tn.max_iterations ? tn.max_iterations : 64u * (W + H) + 1024u;
  a.spin_limit = 1u << (tn.spin_limit_log2 ? (tn.spin_limit_log2 > 30 ? 30 : tn.spin_limit_log2) : 22);
  const size_t n = static_cast<size_t>(W) * H;
  const size_t ntiles = static_cast<size_t>(a.ntx) * a.nty;
  const size_t nwords = (n + 31) / 32;

  // Workspace: coef | st0 | st1 | stamp | tile_live | worklist | control
  size_t off = 0;
  const size_t o_coef = off; off = align_up(off + n * sizeof(float4), 256);
  const size_t o_st0 = off; off = align_up(off + n * sizeof(uint2), 256);
  const size_t o_st1 = off; off = align_up(off + n * sizeof(uint2), 256);
  const size_t o_stamp = off; off = align_up(off + n * sizeof(int), 256);
  const size_t o_live = off; off = align_up(off + ntiles * sizeof(uint32_t), 256);
  const size_t o_work = off; off = align_up(off + ntiles * sizeof(uint32_t), 256);
  const size_t o_ctl = off; off = align_up(off + sizeof(Control), 256);
  const size_t max_groups = static_cast<size_t>(dev->cu_count) * 32;
  const size_t o_slots = off; off = align_up(off + max_groups * sizeof(Slot), 256);
  const unsigned prof_sweeps = tn.profile ? 8192u : 0u;
  const size_t o_prof = off; off = align_up(off + static_cast<size_t>(prof_sweeps) * kProfCols * sizeof(unsigned long long), 256);
  void *ws = nullptr;
  pg_status st = workspace(dev->device, off, &ws);
  if (st != PG_OK) return st;
  char *base = static_cast<char *>(ws);
  a.coef = reinterpret_cast<const float4 *>(base + o_coef);
  a.st[0] = reinterpret_cast<uint2 *>(base + o_st0);
  a.st[1] = reinterpret_cast<uint2 *>(base + o_st1);
  a.stamp = reinterpret_cast<int *>(base + o_stamp);
  a.tile_live = reinterpret_cast<uint32_t *>(base + o_live);
  a.worklist = reinterpret_cast<uint32_t *>(base + o_work);
  a.ctl = reinterpret_cast<Control *>(base + o_ctl);
  a.slots = reinterpret_cast<Slot *>(base + o_slots);
  a.prof = prof_sweeps ? reinterpret_cast<unsigned long long *>(base + o_prof) : nullptr;
  a.prof_sweeps = prof_sweeps;
  if (prof_sweeps) PG_HIP(hipMemsetAsync(a.prof, 0, static_cast<size_t>(prof_sweeps) * kProfCols * 8, stream));
  a.bits = boundaries;

  hipEvent_t ev0 = nullptr, ev1 = nullptr, evr0 = nullptr, evr1 = nullptr;
  PG_HIP(hipEventCreate(&ev0));
  PG_HIP(hipEventCreate(&ev1));
  PG_HIP(hipEventCreate(&evr0));
  PG_HIP(hipEventCreate(&evr1));
  struct EventGuard {
    hipEvent_t *e[4];
    ~EventGuard() {
      for (hipEvent_t *p : e)
        if (*p) (void)hipEventDestroy(*p);
    }
  } guard{{&ev0, &ev1, &evr0, &evr1}};
  PG_HIP(hipEventRecord(ev0, stream));

  const unsigned cu = static_cast<unsigned>(dev->cu_count);
  auto grid_for = [&](size_t items) {
    size_t b = (items + 255) / 256;
    if (b > cu * 8) b = cu * 8;
    return dim3(static_cast<unsigned>(b ? b : 1));
  };
  hipLaunchKernelGGL(fill_u32, grid_for(ntiles), dim3(256), 0, stream, a.tile_live, ntiles, 0u);
  hipLaunchKernelGGL(fill_u32, grid_for(ntiles), dim3(256), 0, stream, a.worklist, ntiles, kNoTile);
  hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, stream, reinterpret_cast<uint32_t *>(a.ctl),
                     sizeof(Control) / 4, 0u);
  hipLaunchKernelGGL(fill_u32, grid_for(max_groups * sizeof(Slot) / 4), dim3(256), 0, stream,
                     reinterpret_cast<uint32_t *>(a.slots), max_groups * sizeof(Slot) / 4, 0u);
  hipLaunchKernelGGL(boundaries_kernel, grid_for(nwords * 32), dim3(256), 0, stream, a.g, speed, az, ecc,
                     boundaries);
  hipLaunchKernelGGL(prep, grid_for(n), dim3(256), 0, stream, a, speed, az, ecc, time, refs_x, refs_y);
  hipLaunchKernelGGL(init_control, dim3(1), dim3(1), 0, stream, a.ctl);
  PG_HIP(hipGetLastError());

  // Validation result before any sweep gathers through the references.
  unsigned book[8];
  PG_HIP(hipMemcpyAsync(book, &a.ctl->book, sizeof(book), hipMemcpyDeviceToHost, stream));
  PG_HIP(hipStreamSynchronize(stream));
  if (book[3] & 2u) {
    return set_error(PG_ERR_INVALID, "a burned cell (time < FLT_MAX) has a reference outside the raster");
  }

  PG_HIP(hipEventRecord(evr0, stream));
  unsigned launched_groups = 0;
  if (tn.mode == 1) {
    const dim3 grid(cu * 4);
    launched_groups = grid.x;
    for (int k = 0;; ++k) {
      if (a.cache)
        hipLaunchKernelGGL(propagate_sweep<true>, grid, dim3(256), 0, stream, a, k);
      else
        hipLaunchKernelGGL(propagate_sweep<false>, grid, dim3(256), 0, stream, a, k);
      hipLaunchKernelGGL(sweep_book, dim3(1), dim3(1), 0, stream, a, k);
      PG_HIP(hipGetLastError());
      PG_HIP(hipMemcpyAsync(book, &a.ctl->book, sizeof(book), hipMemcpyDeviceToHost, stream));
      PG_HIP(hipStreamSynchronize(stream));
      if (book[1]) break;
    }
    PG_HIP(hipEventRecord(evr1, stream));
  } else {
    // Waves per workgroup: as many tile slices as fit beside the bitset when
    // the bitset can live in LDS (one workgroup per CU), else 16 waves.
    const size_t bits_bytes = nwords * sizeof(uint32_t);
    const size_t static_lds = 256;
    int waves = 16;
    a.lds_bits = 0;
    if (bits_bytes + static_lds + sizeof(TileLds) * 4 <= dev->lds_per_block) {
      a.lds_bits = 1;
      waves = static_cast<int>((dev->lds_per_block - bits_bytes - static_lds) / sizeof(TileLds));
    }
    const int max_waves = 16;  // the kernel's launch bound: 1024 threads
    if (tn.waves_per_group) waves = static_cast<int>(tn.waves_per_group);
    if (waves > max_waves) waves = max_waves;
    const size_t shmem = waves * sizeof(TileLds) + (a.lds_bits ? bits_bytes : 0);
    const int block = waves * kWave;
    void (*kern)(Args) = a.cache ? (a.lds_bits ? propagate_persistent<true, true> : propagate_persistent<true, false>)
                                 : (a.lds_bits ? propagate_persistent<false, true> : propagate_persistent<false, false>);
    int per_cu = 0;
    PG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, shmem));
    if (per_cu < 1) return set_error(PG_ERR_INVALID, "persistent kernel does not fit on a CU");
    unsigned groups = cu * static_cast<unsigned>(per_cu);
    if (groups > max_groups) groups = static_cast<unsigned>(max_groups);
    if (groups > 64u * kWave) groups = 64u * kWave;  // the barrier master polls <= 64 slots per lane
    if (tn.workgroups && tn.workgroups < groups) groups = tn.workgroups;
    launched_groups = groups;
    void *kargs[] = {&a};
    PG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kern), dim3(groups),
                                      dim3(block), kargs, static_cast<unsigned>(shmem), stream));
    PG_HIP(hipEventRecord(evr1, stream));
    PG_HIP(hipMemcpyAsync(book, &a.ctl->book, sizeof(book), hipMemcpyDeviceToHost, stream));
    PG_HIP(hipStreamSynchronize(stream));
  }
  if (book[3] & 1u) {
    return set_error(PG_ERR_TIMEOUT, "a grid barrier wait exceeded its spin limit");
  }
  const int last = static_cast<int>(book[4]);
  hipLaunchKernelGGL(post, grid_for(n), dim3(256), 0, stream, a, (last + 1) & 1, speed, time, refs_x, refs_y);
  PG_HIP(hipGetLastError());
  PG_HIP(hipEventRecord(ev1, stream));
  unsigned long long counters[8];
  PG_HIP(hipMemcpyAsync(counters, a.ctl->stats, sizeof(counters), hipMemcpyDeviceToHost, stream));
  PG_HIP(hipStreamSynchronize(stream));
  float ms = 0.0f, rms = 0.0f;
  PG_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  PG_HIP(hipEventElapsedTime(&rms, evr0, evr1));
  if (stats) {
    stats->iterations = static_cast<uint64_t>(last) + 1;
    stats->relaxations = counters[0];
    stats->improvements = counters[1];
    stats->dda_steps = counters[2];
    stats->tiles_visited = counters[3];
    stats->elapsed_ms = ms;
    stats->relax_ms = rms;
    int clock_khz = 0;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev->device) != hipSuccess ||
        clock_khz <= 0)
      clock_khz = 100000;  // 100 MHz constant clock
    const double groups_d = launched_groups ? static_cast<double>(launched_groups) : 1.0;
    stats->work_ms = static_cast<double>(counters[4]) / clock_khz / groups_d;
    stats->wait_ms = static_cast<double>(counters[5]) / clock_khz / groups_d;
    stats->crit_ms = static_cast<double>(counters[6]) / clock_khz;
    stats->workgroups = launched_groups;
    stats->lds_bitset = static_cast<uint32_t>(a.lds_bits);
    for (int q = 0; q < 4; ++q) stats->phase_ms[q] = stats->barrier_ms[q] = 0.0;
    if (prof_sweeps) {
      std::vector<unsigned long long> prof(static_cast<size_t>(prof_sweeps) * kProfCols);
      PG_HIP(hipMemcpy(prof.data(), a.prof, prof.size() * 8, hipMemcpyDeviceToHost));
      for (size_t r = 0; r < prof.size(); ++r) {
        const double ms = static_cast<double>(prof[r]) / clock_khz;
        if (r % kProfCols < 4)
          stats->phase_ms[r % kProfCols] += ms;
        else
          stats->barrier_ms[r % kProfCols - 4] += ms;
      }
    }
  }
  if (book[2]) return set_error(PG_ERR_NOT_CONVERGED, "iteration cap reached before convergence");
  return PG_OK;
}

}  // namespace pg
