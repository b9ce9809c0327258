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

constexpr int kTile = 16;               // tile edge (cells)
constexpr int kHalo = kTile + 2;        // staged edge
constexpr int kHaloCells = kHalo * kHalo;
constexpr int kThreads = kTile * kTile; // one thread per tile cell
constexpr uint32_t kNoRef = 0xFFFFFFFFu;
constexpr int kStampNever = INT_MIN / 2;
constexpr int kWalkBatch = 8;           // DDA cells tested per batch

// ------------------------------------------------------------ control block

struct alignas(128) Line {
  unsigned v[32];
};

// Grid-wide state of one propagation; each hot word sits on its own line.
struct Control {
  Line arrive;       // [0] barrier arrivals
  Line gen;          // [0] barrier generation
  Line live;         // [0] worklist length (grows during a sweep)
  Line book;         // [0] live snapshot, [1] done, [2] capped, [3] error, [4] last sweep
  Line changes;      // [0..1] tiles changed in sweep k (k & 1)
  unsigned long long stats[8];  // relaxations, improvements, dda_steps, tiles
};

__device__ inline unsigned ld_acquire(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned ld_relaxed(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_relaxed(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_release(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned add_acq_rel(unsigned *p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned add_relaxed(unsigned *p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// f32 DDA of propag.h:91-135.
struct Dda {
  int tx, ty, sx, sy, cx, cy;
  float tmx, tmy, dlx, dly;
  __device__ inline void init(int fx, int fy, int tx_, int ty_) {
    tx = tx_;
    ty = ty_;
    sx = (0 < tx - fx) - (tx - fx < 0);
    sy = (0 < ty - fy) - (ty - fy < 0);
    cx = fx;
    cy = fy;
    tmx = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(abs(tx - fx))));
    tmy = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(abs(ty - fy))));
    dlx = tmx;
    dly = tmy;
  }
  __device__ inline bool next(int &rx, int &ry) {
    rx = cx;
    ry = cy;
    if (cx == tx && cy == ty) {
      cx += sx;
      cy += sy;
      return true;
    }
    const bool vx = sx > 0 ? cx <= tx : cx >= tx;
    const bool vy = sy > 0 ? cy <= ty : cy >= ty;
    if (!(vx && vy)) return false;
    if (static_cast<double>(fabsf(tmx - tmy)) < 1e-9) {
      cx += sx;
      cy += sy;
      tmx += dlx;
      tmy += dly;
    } else if (tmx < tmy) {
      cx += sx;
      tmx += dlx;
    } else {
      cy += sy;
      tmy += dly;
    }
    return true;
  }
};

__device__ inline bool bit_of(const uint32_t *words, uint32_t idx) {
  return (words[idx >> 5] >> (idx & 31u)) & 1u;
}

// True if a boundary cell lies on the DDA from (fx,fy) to (tx,ty), both
// included (propag.cu:164-176).  Cells are tested kWalkBatch at a time so the
// bit loads of a batch are in flight together; `steps` counts the cells up to
// and including the first boundary, as a serial walk would.
__device__ inline bool walk_blocked(int fx, int fy, int tx, int ty, const uint32_t *bits, int W,
                                    uint32_t &steps) {
  Dda d;
  d.init(fx, fy, tx, ty);
  bool alive = true;
  while (alive) {
    uint32_t idx[kWalkBatch];
    bool valid[kWalkBatch];
#pragma unroll
    for (int b = 0; b < kWalkBatch; ++b) {
      int x = 0, y = 0;
      valid[b] = alive && d.next(x, y);
      alive = valid[b];
      idx[b] = valid[b] ? static_cast<uint32_t>(x) + static_cast<uint32_t>(y) * static_cast<uint32_t>(W) : 0u;
    }
    bool hit_bits[kWalkBatch];
#pragma unroll
    for (int b = 0; b < kWalkBatch; ++b) hit_bits[b] = bit_of(bits, idx[b]);
#pragma unroll
    for (int b = 0; b < kWalkBatch; ++b) {
      if (valid[b]) {
        ++steps;
        if (hit_bits[b]) return true;
      }
    }
  }
  return false;
}

__device__ inline uint2 pack_state(float t, uint32_t ref) { return make_uint2(__float_as_uint(t), ref); }

struct Args {
  Geom g;
  const float4 *coef;        // A, B, R, null flag
  uint2 *st[2];              // double-buffered (time bits, refx | refy << 16)
  int *stamp;                // last sweep in which the cell changed
  const uint32_t *bits;      // boundary bitset (global copy)
  uint32_t *tile_live;       // 1 once a tile is on the worklist
  uint32_t *worklist;        // live tile ids
  Control *ctl;
  int ntx, nty;              // tiles per row / column
  unsigned max_iters;
  unsigned spin_limit;
  int lds_bits;              // bitset staged in LDS
};

struct TileLds {
  float time[kHaloCells];
  uint32_t ref[kHaloCells];
  float rtime[kHaloCells];
  uint8_t chg[kHaloCells];
};

struct Counters {
  uint32_t relax, improve, steps, tiles;
};

// Marks the 8 neighbour tiles of (tx,ty) live, appending newcomers.
__device__ inline void grow_worklist(const Args &a, int tx, int ty) {
  const int lane = threadIdx.x;
  if (lane < 9 && lane != 4) {
    const int nx = tx + lane % 3 - 1, ny = ty + lane / 3 - 1;
    if (nx >= 0 && ny >= 0 && nx < a.ntx && ny < a.nty) {
      const int t = nx + ny * a.ntx;
      if (__hip_atomic_load(&a.tile_live[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
        unsigned expected = 0u;
        if (__hip_atomic_compare_exchange_strong(&a.tile_live[t], &expected, 1u, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          const unsigned slot = add_relaxed(&a.ctl->live.v[0], 1u);
          a.worklist[slot] = static_cast<uint32_t>(t);
        }
      }
    }
  }
}

// One sweep over one tile.  Returns, uniformly over the workgroup, whether
// any cell committed an improvement.
__device__ int process_tile(const Args &a, const uint32_t *bits, int tile, int k, TileLds &s,
                             Counters &cnt) {
  const Geom &g = a.g;
  const uint2 *__restrict__ st_in = a.st[k & 1];
  uint2 *__restrict__ st_out = a.st[(k + 1) & 1];
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  const int ox = tx * kTile, oy = ty * kTile;
  const int tid = threadIdx.x;

  // Stage the halo tile (load_points_into_shared_memory, propag.cu:231-261).
  for (int h = tid; h < kHaloCells; h += kThreads) {
    const int gx = ox + h % kHalo - 1, gy = oy + h / kHalo - 1;
    float t = FLT_MAX, rt = FLT_MAX;
    uint32_t ref = kNoRef;
    bool chg = false;
    if (gx >= 0 && gy >= 0 && gx < g.W && gy < g.H) {
      const size_t gi = static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W;
      const uint2 w = st_in[gi];
      t = __uint_as_float(w.x);
      ref = w.y;
      chg = a.stamp[gi] >= k - 1;
      // A neighbour only feeds the update rule while it burns before
      // max_time; only then is its reference's live time needed.
      if (t < g.max_time && t < FLT_MAX) {
        const size_t ri = static_cast<size_t>(ref & 0xffffu) + static_cast<size_t>(ref >> 16) * g.W;
        rt = __uint_as_float(st_in[ri].x);
        chg = chg || a.stamp[ri] >= k - 1;
      }
    }
    s.time[h] = t;
    s.ref[h] = ref;
    s.rtime[h] = rt;
    s.chg[h] = chg ? 1 : 0;
  }
  __syncthreads();

  const int lx = tid % kTile + 1, ly = tid / kTile + 1;
  const int gx = ox + lx - 1, gy = oy + ly - 1;
  const bool inside = gx < g.W && gy < g.H;
  bool need = false;
  if (inside) {
#pragma unroll
    for (int j = -1; j < 2; ++j)
#pragma unroll
      for (int i = -1; i < 2; ++i) need = need || s.chg[(lx + i) + (ly + j) * kHalo];
  }
  bool improved = false;
  if (need) {
    ++cnt.relax;
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
        const float4 kf = a.coef[static_cast<size_t>(nx) + static_cast<size_t>(ny) * g.W];
        if (kf.w != 0.0f) continue;  // fire_is_null
        const uint32_t ref = s.ref[nl];
        const int rfx = static_cast<int>(ref & 0xffffu), rfy = static_cast<int>(ref >> 16);
        if (rfx == gx && rfy == gy) continue;
        const float nt = time_from(n_t, nx, ny, gx, gy, kf, g);
        const float ct = time_from(s.rtime[nl], rfx, rfy, gx, gy, kf, g);
        if (ct < nt && ct < best_t) {
          uint32_t cand_ref = ref;
          float cand_t = ct;
          if (walk_blocked(nx, ny, rfx, rfy, bits, g.W, cnt.steps)) {
            cand_ref = static_cast<uint32_t>(nx) | (static_cast<uint32_t>(ny) << 16);
            cand_t = nt;
          }
          if (cand_t < best_t) {
            best_t = cand_t;
            best_ref = cand_ref;
          }
        } else if (nt < best_t) {
          best_t = nt;
          best_ref = static_cast<uint32_t>(nx) | (static_cast<uint32_t>(ny) << 16);
        }
      }
    }
    const size_t gi = static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W;
    improved = best_t < own_t;  // update_point (propag.cu:192-200)
    st_out[gi] = improved ? pack_state(best_t, best_ref) : pack_state(own_t, s.ref[me]);
    if (improved) {
      a.stamp[gi] = k;
      ++cnt.improve;
    }
  }
  // Note: the barrier below also orders the LDS reuse by the next tile.
  const int changed = __syncthreads_or(improved ? 1 : 0);
  if (changed) grow_worklist(a, tx, ty);
  if (tid == 0) ++cnt.tiles;
  return changed;
}

// Grid barrier with bookkeeping by the last arriver.  Returns 1 when the
// propagation is finished (quiet sweep, iteration cap or timeout).
__device__ int grid_sync(const Args &a, unsigned &gen, int k, unsigned my_changes) {
  __shared__ int s_done;
  __syncthreads();
  if (threadIdx.x == 0) {
    Control *c = a.ctl;
    if (my_changes) add_relaxed(&c->changes.v[k & 1], my_changes);
    __threadfence();
    const unsigned prev = add_acq_rel(&c->arrive.v[0], 1u);
    int done = 0;
    if (prev == gridDim.x - 1) {
      const unsigned ch = ld_relaxed(&c->changes.v[k & 1]);
      st_relaxed(&c->changes.v[(k + 1) & 1], 0u);
      st_relaxed(&c->book.v[0], ld_relaxed(&c->live.v[0]));
      const bool cap = static_cast<unsigned>(k + 1) >= a.max_iters;
      done = (ch == 0u) || cap;
      if (cap && ch != 0u) st_relaxed(&c->book.v[2], 1u);
      if (done) st_relaxed(&c->book.v[4], static_cast<unsigned>(k));
      st_relaxed(&c->book.v[1], done ? 1u : 0u);
      st_relaxed(&c->arrive.v[0], 0u);
      st_release(&c->gen.v[0], gen + 1u);
    } else {
      unsigned spins = 0;
      while (ld_acquire(&c->gen.v[0]) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > a.spin_limit) {
          st_relaxed(&c->book.v[3], 1u);
          break;
        }
      }
      done = ld_relaxed(&c->book.v[1]) != 0u || ld_relaxed(&c->book.v[3]) != 0u;
    }
    __threadfence();
    gen += 1u;
    s_done = done;
  }
  __syncthreads();
  return s_done;
}

__device__ inline void flush_counters(const Args &a, const Counters &cnt) {
  // One 64-bit add per counter per wave.
  unsigned long long v[4] = {cnt.relax, cnt.improve, cnt.steps, cnt.tiles};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned long long x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(&a.ctl->stats[q], x);
  }
}

// Persistent form: one launch runs every sweep; workgroups stride over the
// live worklist and meet at the grid barrier after each sweep.
__global__ __launch_bounds__(kThreads) void propagate_persistent(Args a) {
  extern __shared__ uint32_t dyn_lds[];
  __shared__ TileLds s;
  const uint32_t *bits = a.bits;
  if (a.lds_bits) {
    const unsigned nwords = (static_cast<unsigned>(a.g.W) * static_cast<unsigned>(a.g.H) + 31u) / 32u;
    for (unsigned i = threadIdx.x; i < nwords; i += blockDim.x) dyn_lds[i] = a.bits[i];
    __syncthreads();
    bits = dyn_lds;
  }
  Counters cnt{0, 0, 0, 0};
  unsigned gen = 0;
  for (int k = 0;; ++k) {
    const unsigned n_live = ld_relaxed(&a.ctl->book.v[0]);
    unsigned my_changes = 0;
    for (unsigned w = blockIdx.x; w < n_live; w += gridDim.x) {
      const int tile = static_cast<int>(a.worklist[w]);
      my_changes += process_tile(a, bits, tile, k, s, cnt) ? 1u : 0u;
      __syncthreads();
    }
    if (grid_sync(a, gen, k, my_changes)) break;
  }
  flush_counters(a, cnt);
}

// Plain form: one launch per sweep (the reference's relaunch structure).
__global__ __launch_bounds__(kThreads) void propagate_sweep(Args a, int k) {
  __shared__ TileLds s;
  Counters cnt{0, 0, 0, 0};
  const unsigned n_live = ld_relaxed(&a.ctl->book.v[0]);
  unsigned my_changes = 0;
  for (unsigned w = blockIdx.x; w < n_live; w += gridDim.x) {
    my_changes += process_tile(a, a.bits, static_cast<int>(a.worklist[w]), k, s, cnt) ? 1u : 0u;
    __syncthreads();
  }
  if (threadIdx.x == 0 && my_changes) add_relaxed(&a.ctl->changes.v[k & 1], my_changes);
  flush_counters(a, cnt);
}

// Between plain sweeps: the last-arriver bookkeeping of grid_sync.
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
// validation and the initial worklist (tiles holding or touching a burned
// cell).
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
    const uint32_t ref = static_cast<uint32_t>(rx[i]) | (static_cast<uint32_t>(ry[i]) << 16);
    a.st[0][i] = pack_state(t, ref);
    a.st[1][i] = pack_state(t, ref);
    const bool burned = t < FLT_MAX;
    a.stamp[i] = burned ? -1 : kStampNever;
    if (burned) {
      if (rx[i] >= a.g.W || ry[i] >= a.g.H) {
        atomicOr(&a.ctl->book.v[3], 2u);  // invalid reference
        continue;
      }
      const int x = static_cast<int>(i % a.g.W), y = static_cast<int>(i / a.g.W);
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
    if (speed[i] == 0.0f || t > a.g.max_time) t = FLT_MAX;
    time[i] = t;
    rx[i] = static_cast<uint16_t>(w.y & 0xffffu);
    ry[i] = static_cast<uint16_t>(w.y >> 16);
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
  if (tn.tile_w && tn.tile_w != kTile) return set_error(PG_ERR_INVALID, "tile_w must be 16");
  if (tn.tile_h && tn.tile_h != kTile) return set_error(PG_ERR_INVALID, "tile_h must be 16");
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
  hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, stream, reinterpret_cast<uint32_t *>(a.ctl),
                     sizeof(Control) / 4, 0u);
  hipLaunchKernelGGL(prep, grid_for(n), dim3(256), 0, stream, a, speed, az, ecc, time, refs_x, refs_y);
  hipLaunchKernelGGL(boundaries_kernel, grid_for(nwords * 32), dim3(256), 0, stream, a.g, speed, az, ecc,
                     boundaries);
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
  if (tn.mode == 1) {
    const dim3 grid(cu * 4);
    for (int k = 0;; ++k) {
      hipLaunchKernelGGL(propagate_sweep, grid, dim3(kThreads), 0, stream, a, k);
      hipLaunchKernelGGL(sweep_book, dim3(1), dim3(1), 0, stream, a, k);
      PG_HIP(hipGetLastError());
      PG_HIP(hipMemcpyAsync(book, &a.ctl->book, sizeof(book), hipMemcpyDeviceToHost, stream));
      PG_HIP(hipStreamSynchronize(stream));
      if (book[1]) break;
    }
    PG_HIP(hipEventRecord(evr1, stream));
  } else {
    const size_t bits_bytes = nwords * sizeof(uint32_t);
    a.lds_bits = bits_bytes + sizeof(TileLds) + 1024 <= dev->lds_per_block ? 1 : 0;
    const size_t shmem = a.lds_bits ? bits_bytes : 0;
    int per_cu = 0;
    PG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, propagate_persistent, kThreads, shmem));
    if (per_cu < 1) return set_error(PG_ERR_INVALID, "persistent kernel does not fit on a CU");
    unsigned groups = cu * static_cast<unsigned>(per_cu);
    if (tn.workgroups && tn.workgroups < groups) groups = tn.workgroups;
    void *kargs[] = {&a};
    PG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(propagate_persistent), dim3(groups),
                                      dim3(kThreads), kargs, static_cast<unsigned>(shmem), stream));
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
  }
  if (book[2]) return set_error(PG_ERR_NOT_CONVERGED, "iteration cap reached before convergence");
  return PG_OK;
}

}  // namespace pg
