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
//   * only tiles that can hold such a cell are visited.  A tile in which a
//     cell changed in sweep k marks, for sweep k+1, its 3x3 tile
//     neighbourhood and every tile within one tile of a cell that takes its
//     reference from this tile from two or more cells away (each tile keeps
//     the tile box of those dependents).  A reference at most one cell away
//     from its dependent lies in the dependent's tile neighbourhood, and a
//     cell that adopts a reference in sweep k is itself a change of sweep k,
//     so the marks cover every cell the rule above re-evaluates.
// Marks are bits of a per-sweep bitmap (double-buffered by sweep parity,
// cleared by the wave that reads them).  Several scenarios can share one
// launch: each owns a range of bitmap words and its own rasters, starts at
// its own sweep, and drops out when a sweep of it commits nothing.
// Each tile of 8x8 cells is staged with its 1-cell halo in LDS (time,
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
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "firelib_dev.h"
#include "pg_internal.h"
#include "propag_dev.h"

namespace pg {
namespace {

constexpr uint32_t kNoRef = 0xFFFFFFFFu;
constexpr int kStampNever = INT_MIN / 2;
constexpr int kWave = 64;               // lanes per wavefront

// ------------------------------------------------------------ control block

struct alignas(128) Line {
  unsigned v[32];
};

// Grid-wide state of one launch; each hot word sits on its own line.
struct Control {
  Line book;         // [1] done, [3] error bits, [4] last sweep run
  Line changes;      // [0..1] tiles changed in sweep k (k & 1), plain form only
  unsigned long long stats[16];  // relaxations, improvements, dda_steps, tiles, work, wait,
                                 // critical-path ticks, marks, phase ticks summed over waves [4],
                                 // tile visits without work
};

// Every pointer the kernels use points to device memory; pointers read from
// a scenario descriptor are generic to the compiler, which would emit FLAT
// accesses (counted against the LDS wait counter too).  G() restates them as
// global.
#define PG_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline PG_GLOBAL T *G(T *p) {
  return (PG_GLOBAL T *)(p);
}

__device__ inline void st_relaxed(unsigned *p, unsigned v) {
  __hip_atomic_store(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned add_relaxed(unsigned *p, unsigned v) {
  return __hip_atomic_fetch_add(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t or_fetch(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_or(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void or_relaxed(uint32_t *p, uint32_t v) {
  (void)__hip_atomic_fetch_or(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void and_relaxed(uint32_t *p, uint32_t v) {
  (void)__hip_atomic_fetch_and(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void min_relaxed(uint32_t *p, uint32_t v) {
  (void)__hip_atomic_fetch_min(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void max_relaxed(uint32_t *p, uint32_t v) {
  (void)__hip_atomic_fetch_max(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Data that other workgroups write during a persistent launch (cell state,
// stamps, work list, barrier words) is read and written only with these
// device-coherent accesses (agent-scope relaxed atomics: global loads and
// stores with sc1, which are neither served from nor left dirty in an XCD's
// own L2).  A grid barrier then needs no L2 write-back or invalidation, only
// each wave's wait for its own stores (wait_stores), and read-only data
// (fire coefficients, the boundary bitset) stays cached across sweeps.
__device__ inline uint32_t ld_u32(const uint32_t *p) {
  return __hip_atomic_load(G(const_cast<uint32_t *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_u32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_i32(const int *p) {
  return __hip_atomic_load(G(const_cast<int *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_i32(int *p, int v) {
  __hip_atomic_store(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_u64(const unsigned long long *p) {
  return __hip_atomic_load(G(const_cast<unsigned long long *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_u64(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(G(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint2 ld_state(const uint2 *p) {
  const unsigned long long v = ld_u64(reinterpret_cast<const unsigned long long *>(p));
  return make_uint2(static_cast<uint32_t>(v), static_cast<uint32_t>(v >> 32));
}
// Cacheable forms (workgroup-scope relaxed atomics: plain loads that may hit
// the CU's L1 and the XCD's L2) for launches that run an agent-scope acquire
// after every grid barrier, which drops the lines a previous sweep left.
__device__ inline uint2 ld_state_c(const uint2 *p) {
  const unsigned long long v = __hip_atomic_load(
      G(reinterpret_cast<unsigned long long *>(const_cast<uint2 *>(p))), __ATOMIC_RELAXED,
      __HIP_MEMORY_SCOPE_WORKGROUP);
  return make_uint2(static_cast<uint32_t>(v), static_cast<uint32_t>(v >> 32));
}
__device__ inline int ld_i32_c(const int *p) {
  return __hip_atomic_load(G(const_cast<int *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline uint32_t ld_u32_c(const uint32_t *p) {
  return __hip_atomic_load(G(const_cast<uint32_t *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void st_state(uint2 *p, uint2 v) {
  st_u64(reinterpret_cast<unsigned long long *>(p),
         static_cast<unsigned long long>(v.x) | (static_cast<unsigned long long>(v.y) << 32));
}
// Returns once every store this wave has issued is complete (for the
// coherent ones: visible device-wide); the clobber keeps the compiler from
// moving memory accesses across it.
__device__ inline void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Lane l's value for the whole wave, l the same in every lane (a constant
// or a value read off a ballot): v_readlane into a scalar register, with no
// LDS round trip as a shuffle takes.
__device__ inline uint32_t lane_value(uint32_t v, int l) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
}
__device__ inline unsigned long long lane_value(unsigned long long v, int l) {
  return static_cast<unsigned long long>(lane_value(static_cast<uint32_t>(v), l)) |
         (static_cast<unsigned long long>(lane_value(static_cast<uint32_t>(v >> 32), l)) << 32);
}

// OR of v over the wave's 64 lanes, every lane active: prefix ORs within
// each row of 16 lanes by DPP row shifts (row_shr 1, 2, 4, 8; a lane shifted
// in from outside the row reads 0), after which lanes 15, 31, 47 and 63 hold
// their rows' ORs.  VALU and scalar only, no LDS round trip.
__device__ inline uint32_t wave_or(uint32_t v) {
  int x = static_cast<int>(v);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(x, 15) | __builtin_amdgcn_readlane(x, 31) |
                               __builtin_amdgcn_readlane(x, 47) | __builtin_amdgcn_readlane(x, 63));
}

__device__ inline void wave_sync() {
  // Lanes of one wave exchange data through LDS: order the accesses.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ geometry

// A tile of TW x (64 / TW) cells, one per lane (lane l: cell (l % TW, l / TW)),
// staged with a one-cell halo of HW x HH cells.
template <int TW>
struct Tile {
  static constexpr int W = TW, H = kWave / TW;
  static constexpr int HW = W + 2, HH = H + 2;
  static constexpr int Cells = HW * HH;
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

// sqrtf(x), correctly rounded, for x in [1, 2^126): the hardware root s
// (within one ulp) corrected by a Tuckerman-style rounding test, the
// products of s with its neighbours s -+ ulp standing in for the squared
// midpoints, each residual one fused multiply-add: the root is s - ulp when
// x <= (s - ulp) s, s + ulp when x > (s + ulp) s, else s.  Checked on the host against the correctly rounded root
// for every float in [1, 4) and every s within one ulp of it (x -> 4x scales
// s by 2 exactly).  No scaling for tiny x and no special values, which
// sqrtf's general expansion spends about half its instructions on.
__device__ inline float sqrt_rn_ge1(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const uint32_t b = __float_as_uint(s);
  const float sd = __uint_as_float(b - 1u), su = __uint_as_float(b + 1u);
  float r = fmaf(-sd, s, x) <= 0.0f ? sd : s;
  r = fmaf(-su, s, x) > 0.0f ? su : r;
  return r;
}

// time_from (propag.cu:269-294) in closed form; (fx,fy) -> (tx,ty), fire k,
// (fx,fy) != (tx,ty).
__device__ inline float time_from(float from_t, int fx, int fy, int tx, int ty, const float4 &k,
                                  const Geom &g) {
  const int ix = tx - fx;
  const int iy = ty - fy;
  const float ux = static_cast<float>(ix);
  const float uy = static_cast<float>(g.north_up ? -iy : iy);
  const float r = sqrt_rn_ge1(ux * ux + uy * uy);  // integer cells apart: ux^2 + uy^2 >= 1
  const float proj = k.x * uy + k.y * ux;
  if (g.square) return from_t + ((r - proj) * g.cell) * k.z;
  const float ex = ux * g.dx;
  const float ey = static_cast<float>(iy) * g.dy;
  const float dist = sqrtf(ex * ex + ey * ey);
  return from_t + (dist * (1.0f - proj / r)) * k.z;
}

// time_from from the neighbour (tx + i, ty + j), |i|, |j| <= 1, to (tx, ty):
// the same operations as time_from, whose radius sqrtf(ux^2 + uy^2) is then
// sqrtf(1) = 1 or sqrtf(2), both exact constants, on square cells.
__device__ inline float time_from_adjacent(float from_t, int i, int j, int tx, int ty, const float4 &k,
                                           const Geom &g) {
  if (!g.square) return time_from(from_t, tx + i, ty + j, tx, ty, k, g);
  const float ux = static_cast<float>(-i);
  const float uy = static_cast<float>(g.north_up ? j : -j);
  const float r = (i != 0 && j != 0) ? 0x1.6a09e6p+0f : 1.0f;  // sqrtf(2.0f) correctly rounded
  const float proj = k.x * uy + k.y * ux;
  return from_t + ((r - proj) * g.cell) * k.z;
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

// Dynamic LDS of the persistent kernel: the boundary bitset when it fits
// (first, right after the kernel's few static words), then the per-wave
// tile slices.
extern __shared__ __attribute__((aligned(16))) uint32_t dyn_lds[];

// Where a walk reads the boundary bits.  From global memory (L2) it reads a
// copy of the bitset in 8x8 blocks: bit (x, y) is bit (x & 7) | (y & 7) << 3
// of the 64-bit word of block (x >> 3, y >> 3), blocks row-major (bw per
// row), so eight consecutive DDA cells lie in at most 2 x 2 blocks, loaded
// together once per batch of eight instead of one dependent load per cell.
// From LDS (the whole bitset staged there) it reads the row-major bitset,
// one 32-bit word per cell visited, which LDS latency makes cheaper than
// selecting among blocks.
struct BitView {
  const unsigned long long *blk;  // blocked copy (global form)
  unsigned lds_off;               // unused: the LDS form's row-major bitset starts dyn_lds
  int W, bw, bh;
  int ncell;                      // W x H (the LDS form's clamp)
};

template <bool kLdsBits>
__device__ inline unsigned long long blk_at(const BitView &bv, uint32_t b) {
  return G(bv.blk)[b];
}

template <bool kLdsBits>
__device__ inline bool bit_at(const BitView &bv, int x, int y) {
  if (kLdsBits) {
    const uint32_t i = static_cast<uint32_t>(x) + static_cast<uint32_t>(y) * static_cast<uint32_t>(bv.W);
    return (dyn_lds[i >> 5] >> (i & 31u)) & 1u;
  }
  const unsigned long long w = G(bv.blk)[static_cast<uint32_t>((x >> 3) + (y >> 3) * bv.bw)];
  return (w >> ((x & 7) | ((y & 7) << 3))) & 1ull;
}

// The walks address the LDS bitset by absolute LDS byte address: a cell
// index plus lds_bit_base() (the bitset's LDS address in bits, a multiple of
// 32) gives the word's byte address as (id >> 3) & ~3 with the cell's bit
// at id & 31, so no base is added per cell.
typedef const __attribute__((address_space(3))) uint32_t lds_cu32;
__device__ inline uint32_t lds_bit_base() {
  return static_cast<uint32_t>(reinterpret_cast<size_t>((__attribute__((address_space(3))) uint32_t *)dyn_lds)) << 3;
}
__device__ inline uint32_t lds_bit_word(uint32_t id) { return *(lds_cu32 *)(size_t)((id >> 3) & ~3u); }

// (double)d < 1e-9 for a float d >= 0 is exactly d <= this float (the float
// nearest 1e-9 lies below it and the next one above).
constexpr float kTie = 0x1.12e0bep-30f;

// True if a boundary cell lies on the f32 DDA of propag.h:91-135 from
// (fx,fy) to (tx,ty), both included.  ru, rv count the steps left along each
// axis: the reference's "cur is still between from and to" (propag.h:117-120)
// is ru >= 0 && rv >= 0, i.e. (ru | rv) >= 0, and once the walk has visited
// `to` or stepped past it one of them is negative for good; the cell visited
// with ru, rv left is (tx - sx*ru, ty - sy*rv).  With d = tmax_x - tmax_y
// (exact whenever |d| is near the tie threshold, the two being within a
// factor of two of each other), the reference's branches become: x advances
// iff d <= kTie (tie or tmax_x < tmax_y), y advances iff not d < -kTie (tie
// or tmax_x >= tmax_y; also for from == to, where d is NaN, so that walk
// visits its one cell and ends).  Cells are tested eight at a time; `steps`
// counts the cells up to and including the first boundary, as a serial
// walk would.
template <bool kLdsBits>
__device__ bool walk_blocked(int fx, int fy, int tx, int ty, const BitView &bv, uint32_t &steps) {
  const int ddx = tx - fx, ddy = ty - fy;
  int ru = abs(ddx), rv = abs(ddy);
  const int sx = (ddx > 0) - (ddx < 0), sy = (ddy > 0) - (ddy < 0);
  // 1.0 / float(abs(..)) is a double division narrowed to float
  const float dlx = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(ru)));
  const float dly = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(rv)));
  float tmx = dlx, tmy = dly;
  // An axis without displacement never steps (its tmax is infinite), so it
  // cannot end the walk; pu / pv lift it out of the all-live test below.
  const int pu = ddx == 0 ? (1 << 24) : 0, pv = ddy == 0 ? (1 << 24) : 0;
  if constexpr (kLdsBits) {
    // The cell visited with ru, rv steps left has the index
    // tc - sx*ru - sy*W*rv (24-bit products: |ru|, |rv|, W < 2^15 and the
    // raster below 2^30 cells), here biased by the bitset's LDS address.  Taking it from the counts keeps them
    // current at every step (one subtract-with-borrow each) instead of
    // carrying a running index beside them.
    const uint32_t base = lds_bit_base();
    const int tc = tx + ty * bv.W + static_cast<int>(base);
    const int nsx = -sx, nsw = -sy * bv.W;
    const uint32_t last = static_cast<uint32_t>(bv.ncell - 1) + base;
    for (;;) {
      uint32_t id[8];
      // ru and rv drop by at most one per cell, so with both at 7 or more all
      // eight cells of the batch lie on the walk: no per-cell end test, and
      // the bits are only OR-ed (bit 0 of each shifted word) until one is
      // set.  Taken only when every walking lane of the wave qualifies, so
      // the wave never runs both forms of the batch.
      if (__ballot(min(ru + pu, rv + pv) >= 7) == __activemask()) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          id[b] = static_cast<uint32_t>(tc + __mul24(nsx, ru) + __mul24(nsw, rv));
          const float d = tmx - tmy;
          const bool xs = d <= kTie, ys = !(d < -kTie);
          ru -= xs ? 1 : 0;
          rv -= ys ? 1 : 0;
          tmx = xs ? tmx + dlx : tmx;
          tmy = ys ? tmy + dly : tmy;
        }
        uint32_t acc = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) acc |= lds_bit_word(id[b]) >> (id[b] & 31u);
        if (acc & 1u) {
          uint32_t hits = 0;
#pragma unroll
          for (int b = 0; b < 8; ++b) hits |= ((lds_bit_word(id[b]) >> (id[b] & 31u)) & 1u) << b;
          steps += static_cast<uint32_t>(__builtin_ctz(hits)) + 1u;
          return true;
        }
        steps += 8;
        continue;
      }
      // The walk may end within the batch.  A cell past its end (at most
      // seven) reads an in-raster index (the unsigned clamp) and is dropped
      // by the dead mask: bit 7 - b is set when cell b's counts are past the
      // end, and since the counts never grow the dead cells are a suffix.
      uint32_t dead = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        id[b] = min(static_cast<uint32_t>(tc + __mul24(nsx, ru) + __mul24(nsw, rv)), last);
        dead = __builtin_amdgcn_alignbit(dead, static_cast<uint32_t>(ru | rv), 31);
        const float d = tmx - tmy;
        const bool xs = d <= kTie, ys = !(d < -kTie);
        ru -= xs ? 1 : 0;
        rv -= ys ? 1 : 0;
        tmx = xs ? tmx + dlx : tmx;
        tmy = ys ? tmy + dly : tmy;
      }
      uint32_t acc = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) acc |= lds_bit_word(id[b]) >> (id[b] & 31u);
      if (!(acc & 1u) && (ru | rv) >= 0) {
        steps += 8;
        continue;
      }
      const uint32_t live = 8u - static_cast<uint32_t>(__builtin_popcount(dead & 0xffu));
      uint32_t hits = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) hits |= ((lds_bit_word(id[b]) >> (id[b] & 31u)) & 1u) << b;
      hits &= (1u << live) - 1u;
      if (hits) {
        steps += static_cast<uint32_t>(__builtin_ctz(hits)) + 1u;
        return true;
      }
      steps += live;  // the walk ended without a boundary cell
      return false;
    }
  } else {
    for (;;) {
      // The batch's cells lie between (x0, y0) and (x0 + 7 sx, y0 + 7 sy).
      const int x0 = tx - sx * ru, y0 = ty - sy * rv;
      const int xl = max(min(x0, x0 + 7 * sx), 0) >> 3, yl = max(min(y0, y0 + 7 * sy), 0) >> 3;
      const int xh = min(xl + 1, bv.bw - 1), yh = min(yl + 1, bv.bh - 1);
      const unsigned long long w00 = blk_at<false>(bv, static_cast<uint32_t>(xl + yl * bv.bw));
      const unsigned long long w10 = blk_at<false>(bv, static_cast<uint32_t>(xh + yl * bv.bw));
      const unsigned long long w01 = blk_at<false>(bv, static_cast<uint32_t>(xl + yh * bv.bw));
      const unsigned long long w11 = blk_at<false>(bv, static_cast<uint32_t>(xh + yh * bv.bw));
      int lx = x0 - (xl << 3), ly = y0 - (yl << 3);  // in [0, 16) while live
      uint32_t hits = 0, live_n = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool live = (ru | rv) >= 0;
        const unsigned long long w = (ly & 8) ? ((lx & 8) ? w11 : w01) : ((lx & 8) ? w10 : w00);
        const bool bit = (w >> ((lx & 7) | ((ly & 7) << 3))) & 1ull;
        hits |= (live && bit) ? (1u << b) : 0u;
        live_n += live ? 1u : 0u;
        const float d = tmx - tmy;
        const bool xs = d <= kTie, ys = !(d < -kTie);
        ru -= xs ? 1 : 0;
        rv -= ys ? 1 : 0;
        lx += xs ? sx : 0;
        ly += ys ? sy : 0;
        tmx = xs ? tmx + dlx : tmx;
        tmy = ys ? tmy + dly : tmy;
      }
      if (hits) {
        steps += static_cast<uint32_t>(__builtin_ctz(hits)) + 1u;
        return true;
      }
      steps += live_n;
      if ((ru | rv) < 0) return false;
    }
  }
}

// walk_blocked over the blocked bitset in global memory, for the deferred
// walks: each batch of eight cells loads its 2 x 2 blocks once and spreads
// them over a 16 x 16 window of this lane's LDS (`win`, 32 bytes; row r of
// the window is bits 16 (r & 1) .. +15 of word r >> 1), so each cell is one
// LDS read at window bit 16 ly + lx, as in the LDS form.  The batch's cells,
// also the up to seven past the walk's end, lie within 7 cells of its first
// cell in the walk's direction, so the window from block (xb, yb) =
// floor(min(x0, x0 + 7 sx) / 8), floor(min(y0, y0 + 7 sy) / 8) holds them
// all; the loads are clamped into the block grid, which changes only window
// cells off the raster (past the walk's end, dropped by the dead mask).
__device__ bool walk_window(int fx, int fy, int tx, int ty, const BitView &bv, uint32_t *win, uint32_t &steps) {
  const int ddx = tx - fx, ddy = ty - fy;
  int ru = abs(ddx), rv = abs(ddy);
  const int sx = (ddx > 0) - (ddx < 0), sy = (ddy > 0) - (ddy < 0);
  const float dlx = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(ru)));
  const float dly = static_cast<float>(1.0 / static_cast<double>(static_cast<float>(rv)));
  float tmx = dlx, tmy = dly;
  const uint32_t wbits =
      static_cast<uint32_t>(reinterpret_cast<size_t>((__attribute__((address_space(3))) uint32_t *)win)) << 3;
  const int nsx = -sx, nsy = -16 * sy;
  unsigned long long *w64 = reinterpret_cast<unsigned long long *>(win);
  for (;;) {
    const int x0 = tx + __mul24(nsx, ru), y0 = ty + __mul24(-sy, rv);
    const int xb = min(x0, x0 + 7 * sx) >> 3, yb = min(y0, y0 + 7 * sy) >> 3;
    const int bx0 = max(xb, 0), bx1 = min(xb + 1, bv.bw - 1);
    const int by0 = max(yb, 0), by1 = min(yb + 1, bv.bh - 1);
    const unsigned long long w00 = blk_at<false>(bv, static_cast<uint32_t>(bx0 + by0 * bv.bw));
    const unsigned long long w10 = blk_at<false>(bv, static_cast<uint32_t>(bx1 + by0 * bv.bw));
    const unsigned long long w01 = blk_at<false>(bv, static_cast<uint32_t>(bx0 + by1 * bv.bw));
    const unsigned long long w11 = blk_at<false>(bv, static_cast<uint32_t>(bx1 + by1 * bv.bw));
    // v_perm_b32 picks bytes of {S0, S1} (selector 0-3: S1's bytes, 4-7: S0's):
    // rows 2j and 2j + 1 as (left, right, left, right) bytes.
    const uint32_t a0 = static_cast<uint32_t>(w00), a1 = static_cast<uint32_t>(w00 >> 32);
    const uint32_t b0 = static_cast<uint32_t>(w10), b1 = static_cast<uint32_t>(w10 >> 32);
    const uint32_t c0 = static_cast<uint32_t>(w01), c1 = static_cast<uint32_t>(w01 >> 32);
    const uint32_t e0 = static_cast<uint32_t>(w11), e1 = static_cast<uint32_t>(w11 >> 32);
    constexpr uint32_t kLo = 0x05010400u, kHi = 0x07030602u;
    w64[0] = __builtin_amdgcn_perm(b0, a0, kLo) | (static_cast<unsigned long long>(__builtin_amdgcn_perm(b0, a0, kHi)) << 32);
    w64[1] = __builtin_amdgcn_perm(b1, a1, kLo) | (static_cast<unsigned long long>(__builtin_amdgcn_perm(b1, a1, kHi)) << 32);
    w64[2] = __builtin_amdgcn_perm(e0, c0, kLo) | (static_cast<unsigned long long>(__builtin_amdgcn_perm(e0, c0, kHi)) << 32);
    w64[3] = __builtin_amdgcn_perm(e1, c1, kLo) | (static_cast<unsigned long long>(__builtin_amdgcn_perm(e1, c1, kHi)) << 32);
    // The cells are read by absolute LDS address (lds_bit_word), which the
    // compiler cannot relate to `win`: keep the window's stores before these
    // reads and the next batch's stores after them.
    asm volatile("" ::: "memory");
    // window bit of the cell with ru, rv steps left: 16 (y - 8 yb) + (x - 8 xb)
    const int kb = static_cast<int>(wbits) + (tx - 8 * xb) + 16 * (ty - 8 * yb);
    uint32_t id[8];
    uint32_t dead = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      id[b] = static_cast<uint32_t>(kb + __mul24(nsx, ru) + __mul24(nsy, rv));
      dead = __builtin_amdgcn_alignbit(dead, static_cast<uint32_t>(ru | rv), 31);
      const float d = tmx - tmy;
      const bool xs = d <= kTie, ys = !(d < -kTie);
      ru -= xs ? 1 : 0;
      rv -= ys ? 1 : 0;
      tmx = xs ? tmx + dlx : tmx;
      tmy = ys ? tmy + dly : tmy;
    }
    uint32_t acc = 0, wd[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      wd[b] = lds_bit_word(id[b]) >> (id[b] & 31u);
      acc |= wd[b];
    }
    asm volatile("" ::: "memory");
    if (!(acc & 1u) && (ru | rv) >= 0) {
      steps += 8;
      continue;
    }
    const uint32_t live = 8u - static_cast<uint32_t>(__builtin_popcount(dead & 0xffu));
    uint32_t hits = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) hits |= (wd[b] & 1u) << b;
    hits &= (1u << live) - 1u;
    if (hits) {
      steps += static_cast<uint32_t>(__builtin_ctz(hits)) + 1u;
      return true;
    }
    steps += live;  // the walk ended without a boundary cell
    return false;
  }
}

// Line-of-sight flag for a committed reference: the DDA from the cell to an
// adjacent reference visits exactly those two cells.
template <bool kLdsBits>
__device__ inline bool line_clear(int x, int y, int rx, int ry, const BitView &bv, uint32_t &steps) {
  if (rx == x && ry == y) return false;  // an ignition's own reference: never walked
  if (abs(rx - x) <= 1 && abs(ry - y) <= 1) {
    steps += 2;
    return !bit_at<kLdsBits>(bv, x, y) && !bit_at<kLdsBits>(bv, rx, ry);
  }
  return !walk_blocked<kLdsBits>(x, y, rx, ry, bv, steps);
}

__device__ inline BitView scen_bits(const Scen &sc, unsigned lds_off) {
  return BitView{sc.blk, lds_off, static_cast<int>(sc.g.W), sc.bw, sc.bh,
                 static_cast<int>(sc.g.W * sc.g.H)};
}


__device__ inline uint2 pack_state(float t, uint32_t ref) { return make_uint2(__float_as_uint(t), ref); }


struct Slot;

struct Args {
  const Scen *scen;          // nscen descriptors
  ScenOut *out;
  TileQueue q;               // tiles to visit, by sweep parity
  int nscen;
  int k_begin, k_end;
  Control *ctl;
  Slot *slots;               // per-workgroup profiling lines of the grid barrier
  unsigned long long *arrive;  // grid-barrier arrival words: [2][arrive_stride]
  unsigned arrive_stride;
  unsigned spin_limit;
  int lds_bits;              // bitset of scenario 0 staged in LDS
  int cached_stage;          // stage halos with cacheable loads (acquire after each barrier)
  const uint32_t *lds_src;   // that bitset (row-major)
  unsigned lds_words;
  unsigned long long *prof;  // [sweep - k_begin][kProfCols] profile ticks, or null
  unsigned prof_sweeps;      // rows of prof
};

constexpr int kPairsMax = kWave * 8;  // (cell, neighbour) pairs of one tile

template <int TW>
struct TileLds {  // one wave's staged tile (TW x 64 / TW cells + 1-cell halo)
  static constexpr int kCells = Tile<TW>::Cells;
  unsigned long long key[kWave];  // per cell: least candidate key (packed rule)
  float time[kCells];
  uint32_t ref[kCells];
  float rtime[kCells];
  float ca[kCells], cb[kCells], cr[kCells];  // fire coefficients A, B, R (R = inf: null fire)
  uint16_t pair[kPairsMax];  // packed rule: cell << 4 | neighbour
  uint8_t chg[kCells];       // bit 0: changed in sweep k - 1; bit 1: also feeds the rule
};
static_assert(sizeof(TileLds<8>) % 8 == 0 && sizeof(TileLds<16>) % 8 == 0 && sizeof(TileLds<32>) % 8 == 0,
              "TileLds must keep 8-byte alignment");


// Per-workgroup profiling line of the grid barrier, 128 bytes each, one
// half per parity of the barrier's target: a workgroup that has passed the
// barrier may write its next sweep's line while workgroup 0 still reads
// this sweep's, but not the one after (that needs workgroup 0's arrival).
struct alignas(128) Slot {
  unsigned long long ph[2][4];              // profiling: slowest wave's phase ticks
  unsigned long long t_seen[2], t_arrive[2];  // profiling: sweep start / arrival
  unsigned pad[8];
};
static_assert(sizeof(Slot) == 128, "one slot per 128-byte line");

// The barrier's workspace: the profiling lines, then the arrival words of
// both sweep parities.
constexpr size_t slot_bytes(size_t max_groups) {
  return max_groups * sizeof(Slot) + 2 * max_groups * sizeof(unsigned long long);
}

// Profile columns per sweep: slowest wave per phase (staging, update rule,
// line-of-sight walks, commit and marks), then the barrier: workgroup 0's
// books of the previous sweep done -> last workgroup starting, the longest
// workgroup span, last arrival -> workgroup 0 has seen all, its books.
constexpr int kProfCols = 8;

struct Counters {
  uint32_t relax, improve, steps, tiles, marks, idle;
  uint32_t idle_far;                // profiling: idle visits with no change within two cells
  uint64_t work_ticks, wait_ticks;  // wall_clock64 ticks (wave 0 of each workgroup)
  uint64_t crit_ticks;              // workgroup 0: slowest workgroup's work per sweep, summed
  uint64_t ph[4];                   // profiling: this wave's phase ticks in the sweep
  uint64_t ph_total[4];             // profiling: this wave's phase ticks, all sweeps
};


// Deferred line-of-sight walks.  When the boundary bitset is read from L2,
// a tile's commit walks would keep only its few improving lanes busy for
// the length of the longest walk; instead each improving cell with a new
// reference leaves its state here and the wave runs up to kWalkQ walks at
// once (one per lane), writing each cell's state word when its flag is
// known - still within the sweep, so the next sweep sees the same state.
struct WalkReq {
  uint32_t si;    // scenario
  uint32_t cell;  // x | y << 16
  uint32_t ref;   // reference word without the flag
  float time;
};
constexpr int kWalkQ = kWave;
static_assert(kWalkQ >= kWave, "a tile defers at most one walk per lane, after at most one flush");

// In the same launches a wave also holds back the list entries of the tiles
// it queues for the next sweep (their marks are set at once, so each tile
// is still queued once) and appends them with one atomic add per batch
// instead of one round trip per changed tile.
constexpr int kEntryQ = 2 * kWave;
static_assert(kEntryQ >= kWave, "one round of marks fits after at most one flush");

struct DeferLds {  // one wave's deferred work (LDS)
  WalkReq walk[kWalkQ];
  unsigned long long entry[kEntryQ];
  uint32_t win[kWave][8];  // per lane: the 16 x 16 bitset window of a walk batch
};

template <bool kCache, bool kLdsBits>
constexpr bool defer_walks() { return kCache && !kLdsBits; }

struct WalkQueue {
  WalkReq *q;  // this wave's kWalkQ walk slots (LDS)
  unsigned n;  // queued walks (wave-uniform)
  unsigned long long *e;  // this wave's kEntryQ list entries for the next sweep (LDS)
  unsigned ne;            // held entries (wave-uniform)
  uint32_t (*win)[8];     // this wave's walk windows, one per lane (LDS)
  uint32_t changed_si = 0xFFFFFFFFu;  // the scenario whose last_change this wave set in this sweep
  unsigned long long *lc = nullptr;  // the workgroup's (scenario, sweep) pairs stored (LDS), or null
  uint32_t lc_mask = 0;              // its entries - 1 (direct-mapped by scenario)
};
// Entries of that table: the one-scenario LDS-bitset kernel has little LDS
// to spare; a batched launch keeps one per scenario slot it is likely to meet.
template <bool kLdsBits>
constexpr int last_change_entries() { return kLdsBits ? 16 : 256; }

// Runs the queued walks, one per lane, and writes those cells' states.
__device__ inline void run_walks(const Args &a, int k, WalkQueue &wq, Counters &cnt) {
  wave_sync();
  const unsigned lane = threadIdx.x & (kWave - 1);
  if (lane < wq.n) {
    const WalkReq e = wq.q[lane];
    const Scen &sc = a.scen[e.si];
    const int x = static_cast<int>(e.cell & 0xffffu), y = static_cast<int>(e.cell >> 16);
    uint32_t ref = e.ref;
    const int rx = ref_x<true>(ref), ry = ref_y<true>(ref);
    const BitView bv = scen_bits(sc, 0);
    bool clear;
    if (rx == x && ry == y) {
      clear = false;  // an ignition's own reference: never walked
    } else if (abs(rx - x) <= 1 && abs(ry - y) <= 1) {
      cnt.steps += 2;
      clear = !bit_at<false>(bv, x, y) && !bit_at<false>(bv, rx, ry);
    } else {
      clear = !walk_window(x, y, rx, ry, bv, wq.win[lane], cnt.steps);
    }
    if (clear) ref |= kClearBit;
    uint2 *st_out = (k & 1) ? sc.st[0] : sc.st[1];
    st_state(&st_out[static_cast<size_t>(x) + static_cast<size_t>(y) * sc.g.W], pack_state(e.time, ref));
  }
  wq.n = 0;
  wave_sync();
}

__device__ inline void flush_walks(const Args &a, int k, WalkQueue &wq, Counters &cnt) {
  if (wq.n != 0) run_walks(a, k, wq, cnt);
}

constexpr unsigned long long kNoEntry = ~0ull;

// (base + stride, not an array of pointers: a dynamic index into the
// kernel argument would move the whole argument block to scratch)
__device__ inline unsigned long long *list_of(const TileQueue &q, int k, uint32_t x) {
  return q.list + (static_cast<size_t>(k & 3) * q.nsub + x) * q.cap;
}
__device__ inline uint32_t *count_of(const TileQueue &q, int k, uint32_t x) {
  return q.count + 32 * (static_cast<uint32_t>(k & 3) * q.nsub + x);
}
__device__ inline uint32_t *mark_of(const Scen &sc, int k) { return (k & 1) ? sc.mark[1] : sc.mark[0]; }

// A list entry is scenario << 32 | xcd << 29 | tile: xcd (0..7) names the
// XCD whose workgroups run the tile in launches that split work by XCD,
// the same for all tiles of one 64 x 64-cell block of a scenario (the tile
// with origin (cx, cy) lies in block (cx >> 6, cy >> 6)), so a tile's halo
// and its neighbours' stay in one L2 while a fire front spans several XCDs.
constexpr uint32_t kTileMask = 0x1FFFFFFFu;
__device__ inline uint32_t entry_low(uint32_t si, int cx, int cy, uint32_t t) {
  const uint32_t tag = (static_cast<uint32_t>(cx >> 6) + static_cast<uint32_t>(cy >> 6) * 3u + si * 5u) & 7u;
  return t | (tag << 29);
}
// The sub-list an entry goes to.
__device__ inline uint32_t sub_of(const TileQueue &q, uint32_t low) { return q.nsub > 1 ? (low >> 29) & 7u : 0u; }

// Appends the entries of the lanes with `fresh` to the sub-lists of sweep
// k1 that their tags name: the first lane of each sub-list's group adds the
// group's size to that sub-list's count (all groups in one round trip),
// then each lane stores its entry in its slot.  Called by all lanes.
__device__ inline void append_entries(const TileQueue &q, int k1, bool fresh, unsigned long long ev, int lane) {
  if (__ballot(fresh) == 0ull) return;
  const uint32_t x = sub_of(q, static_cast<uint32_t>(ev));
  unsigned long long mine = 0;
  for (uint32_t t = 0; t < q.nsub; ++t) {
    const unsigned long long g = __ballot(fresh && x == t);
    if (x == t) mine = g;
  }
  const int leader = mine ? __builtin_ctzll(mine) : lane;
  uint32_t base = 0;
  if (fresh && lane == leader) base = add_relaxed(count_of(q, k1, x), static_cast<unsigned>(__builtin_popcountll(mine)));
  base = __shfl(base, leader);
  if (fresh) {
    const size_t slot = base + static_cast<uint32_t>(__builtin_popcountll(mine & ((1ull << lane) - 1ull)));
    if (slot < q.cap) st_u64(&list_of(q, k1, x)[slot], ev);
  }
}

// Queuing a tile for sweep k1 is split in two so that the first round trip
// can overlap other work: mark_issue sets the tile's bit (lanes with
// `valid`), mark_commit gives the lanes that set it first consecutive list
// slots from one atomic add.  Both are called by all lanes of a wave.
struct Mark {
  uint32_t t, bit, old;
  bool valid;
  uint32_t low;  // the list entry's low word (tile and XCD tag)
};

template <int TW>
__device__ inline Mark mark_issue(const TileQueue &q, const Scen &sc, uint32_t si, int k1, bool valid, int X, int Y) {
  Mark m{0u, 0u, 0u, valid, 0u};
  if (valid) {
    m.t = static_cast<uint32_t>(X) + static_cast<uint32_t>(Y) * static_cast<uint32_t>(sc.ntx);
    m.low = entry_low(si, X * TW, Y * Tile<TW>::H, m.t);
    m.bit = 1u << (m.t >> sc.nwords_log2);
    m.old = or_fetch(&mark_of(sc, k1)[m.t & ((1u << sc.nwords_log2) - 1u)], m.bit);
  }
  return m;
}

__device__ inline void mark_commit(const TileQueue &q, uint32_t si, int k1, const Mark &mk, int lane) {
  const bool fresh = mk.valid && (mk.old & mk.bit) == 0u;
  append_entries(q, k1, fresh, (static_cast<unsigned long long>(si) << 32) | mk.low, lane);
}

// Appends the wave's held entries to the list of sweep k1: one atomic add
// for the batch, then the stores.  Called by all lanes of a wave.
__device__ inline void flush_entries(const TileQueue &q, WalkQueue &wq, int k1, int lane) {
  if (wq.ne == 0) return;
  wave_sync();
  for (unsigned i0 = 0; i0 < wq.ne; i0 += kWave) {
    const unsigned i = i0 + static_cast<unsigned>(lane);
    const bool has = i < wq.ne;
    append_entries(q, k1, has, has ? wq.e[i] : 0ull, lane);
  }
  wq.ne = 0;
  wave_sync();
}

// mark_commit for a wave that holds its entries back: the lanes that set
// their tile's mark first add the tile to the wave's entries.
__device__ inline void mark_hold(const TileQueue &q, WalkQueue &wq, uint32_t si, int k1, const Mark &mk, int lane) {
  const bool fresh = mk.valid && (mk.old & mk.bit) == 0u;
  const unsigned long long m = __ballot(fresh);
  if (!m) return;
  const unsigned nf = static_cast<unsigned>(__builtin_popcountll(m));
  if (wq.ne + nf > static_cast<unsigned>(kEntryQ)) flush_entries(q, wq, k1, lane);
  if (fresh)
    wq.e[wq.ne + static_cast<unsigned>(__builtin_popcountll(m & ((1ull << lane) - 1ull)))] =
        (static_cast<unsigned long long>(si) << 32) | mk.low;
  wq.ne += nf;
}

// Single-thread form for set-up kernels.
__device__ inline void queue_tile(const TileQueue &q, const Scen &sc, uint32_t si, int k1, int X, int Y) {
  const uint32_t t = static_cast<uint32_t>(X) + static_cast<uint32_t>(Y) * static_cast<uint32_t>(sc.ntx);
  const uint32_t bit = 1u << (t >> sc.nwords_log2);
  const uint32_t old = atomicOr(&mark_of(sc, k1)[t & ((1u << sc.nwords_log2) - 1u)], bit);
  if (old & bit) return;
  const uint32_t low = entry_low(si, X * sc.tw, Y * sc.th, t);
  const uint32_t x = sub_of(q, low);
  const size_t slot = atomicAdd(count_of(q, k1, x), 1u);
  if (slot < q.cap) list_of(q, k1, x)[slot] = (static_cast<unsigned long long>(si) << 32) | low;
}

// The far dependents a change of a tile's far-reference cell in sweep k
// queues for sweep k + 1: the tiles that the tile's dependents box (the
// bounding box, in cells, of every cell that took a far reference from the
// tile; bx: x0, y0, x1, y1 in lanes 0..3) covers once grown by one cell, as
// the dependents' neighbours take candidates through them.
struct NextTiles {
  int X0, Y0, w, n;  // tile range (n = 0: none)
};

template <int TW>
__device__ inline NextTiles next_tiles(const Scen &sc, uint32_t bx) {
  NextTiles nt{0, 0, 1, 0};
  const uint32_t x0 = lane_value(bx, 0), y0 = lane_value(bx, 1), x1 = lane_value(bx, 2), y1 = lane_value(bx, 3);
  if (x0 <= x1 && y0 <= y1) {  // an empty box is x0 = y0 = 0xFFFFFFFF, x1 = y1 = 0
    nt.X0 = max(static_cast<int>(x0) - 1, 0) / TW;
    nt.Y0 = max(static_cast<int>(y0) - 1, 0) / Tile<TW>::H;
    const int X1 = min(static_cast<int>(x1) + 1, sc.g.W - 1) / TW;
    const int Y1 = min(static_cast<int>(y1) + 1, sc.g.H - 1) / Tile<TW>::H;
    nt.w = X1 - nt.X0 + 1;
    nt.n = nt.w * (Y1 - nt.Y0 + 1);
  }
  return nt;
}

// The 3x3 tile neighbourhood a changed tile queues for sweep k + 1.  A
// change of cell r in sweep k makes sweep k + 1 evaluate the cells next to
// r and the cells next to each cell c that references r (c's neighbours
// take a candidate through c's reference).  Those c two or more cells from
// r are reached through the dependents box; the others are r's neighbours,
// in the tile or its one-cell ring.  So with F = the changed cells and the
// unchanged cells of the tile and ring whose reference is a changed
// neighbour, a neighbour tile is queued exactly when some cell of F lies
// next to it (tile-local x <= 0 for the tile on the left, >= TW - 1 for the
// one on the right, the same for y).  Returns the 9-bit set (bit
// (dy + 1) * 3 + dx + 1), uniform over the wave; called by all lanes with
// `imp` the ballot of the tile's changed cells (lane l = cell (l % TW, l / TW)).
template <bool kCache, int TW>
__device__ inline uint32_t next_neighbours(const TileLds<TW> &s, int ox, int oy, int lx, int ly, bool improved,
                                           unsigned long long imp, int lane) {
  constexpr int TH = Tile<TW>::H, HW = Tile<TW>::HW;
  // A cell at tile-local (cx, cy) whose reference is a changed neighbour in
  // the tile.
  auto near_dep = [&](int cx, int cy, uint32_t ref) {
    const int rx = ref_x<kCache>(ref) - ox, ry = ref_y<kCache>(ref) - oy;
    const bool next = abs(rx - cx) <= 1 && abs(ry - cy) <= 1 && (rx != cx || ry != cy);
    return next && rx >= 0 && ry >= 0 && rx < TW && ry < TH && ((imp >> (rx + ry * TW)) & 1ull);
  };
  // Bits of the neighbour tiles next to tile-local (cx, cy).
  auto sides = [](int cx, int cy) {
    const uint32_t xs = 2u | (cx <= 0 ? 1u : 0u) | (cx >= TW - 1 ? 4u : 0u);
    const uint32_t ys = 2u | (cy <= 0 ? 1u : 0u) | (cy >= TH - 1 ? 4u : 0u);
    return xs * (ys & 1u) | ((xs * ((ys >> 1) & 1u)) << 3) | ((xs * ((ys >> 2) & 1u)) << 6);
  };
  uint32_t m = 0;
  const int cx = lx - 1, cy = ly - 1;
  if (improved || near_dep(cx, cy, s.ref[lx + ly * HW])) m = sides(cx, cy);
  // the one-cell ring, a cell per lane per round: the top and bottom rows
  // (HW cells each), then the left and right columns (TH cells each)
  constexpr int kRing = 2 * HW + 2 * TH;
#pragma unroll
  for (int r0 = 0; r0 < kRing; r0 += kWave) {
    const int l = r0 + lane;
    if (l < kRing) {
      int rx, ry;
      if (l < 2 * HW) {
        rx = l % HW - 1;
        ry = l < HW ? -1 : TH;
      } else {
        const int c = l - 2 * HW;
        rx = c < TH ? -1 : TW;
        ry = c % TH;
      }
      if (near_dep(rx, ry, s.ref[(rx + 1) + (ry + 1) * HW])) m |= sides(rx, ry);
    }
  }
  uint32_t q = 0;
#pragma unroll
  for (int b = 0; b < 9; ++b) q |= (__ballot((m >> b) & 1u) != 0ull) ? (1u << b) : 0u;
  return q;
}

// Ballot masks of a tile's two first / last columns and rows.
template <int TW>
constexpr unsigned long long two_cols() {
  unsigned long long m = 0;
  for (int y = 0; y < Tile<TW>::H; ++y) m |= 3ull << (y * TW);
  return m;
}
template <int TW>
constexpr unsigned long long two_rows() {
  return 2 * TW >= 64 ? ~0ull : (1ull << (2 * TW)) - 1ull;
}

// The same set widened to what the tile's ballot alone shows: a neighbour
// tile is queued if a changed cell lies within two cells of it (a superset
// of next_neighbours, without its reads and ballots: the one-scenario
// kernel with the bitset in LDS runs its sweeps as one dependent chain, where
// those cost more than the extra visits they save).
template <int TW>
__device__ inline uint32_t near_band(unsigned long long imp) {
  constexpr unsigned long long kLeft = two_cols<TW>(), kRight = two_cols<TW>() << (TW - 2);
  constexpr unsigned long long kTop = two_rows<TW>(), kBottom = two_rows<TW>() << (TW * (Tile<TW>::H - 2));
  uint32_t q = 0;
#pragma unroll
  for (int b = 0; b < 9; ++b) {
    const int dx = b % 3 - 1, dy = b / 3 - 1;
    const unsigned long long selx = dx < 0 ? kLeft : (dx > 0 ? kRight : ~0ull);
    const unsigned long long sely = dy < 0 ? kTop : (dy > 0 ? kBottom : ~0ull);
    q |= (imp & selx & sely) != 0ull ? 1u << b : 0u;
  }
  return q;
}

// Issues the marks of the queued neighbour tiles `nb` (lanes 0..8) and of
// the box of far dependents (box tile q in lane 9 + q of round 0 and lane
// (q - 55) % 64 of round 1 + (q - 55) / 64).
template <int TW>
__device__ inline Mark next_round0_issue(const TileQueue &q, const Scen &sc, uint32_t si, int k, int tx, int ty,
                                         const NextTiles &nt, uint32_t nb, int lane) {
  if (lane < 9) {
    const int dx = lane % 3 - 1, dy = lane / 3 - 1;
    const int X = tx + dx, Y = ty + dy;
    return mark_issue<TW>(q, sc, si, k + 1, ((nb >> lane) & 1u) && X >= 0 && Y >= 0 && X < sc.ntx && Y < sc.nty, X,
                          Y);
  }
  const int b = lane - 9;
  return mark_issue<TW>(q, sc, si, k + 1, b < nt.n, nt.X0 + b % nt.w, nt.Y0 + b / nt.w);
}

// Rounds 1.. of the box; returns the number of tiles marked in all rounds.
template <bool kHold, int TW>
__device__ inline uint32_t next_rest(const TileQueue &q, const Scen &sc, uint32_t si, int k, const NextTiles &nt,
                                     WalkQueue &wq, int lane) {
  for (int b0 = kWave - 9; b0 < nt.n; b0 += kWave) {
    const int b = b0 + lane;
    const Mark mk = mark_issue<TW>(q, sc, si, k + 1, b < nt.n, nt.X0 + b % nt.w, nt.Y0 + b / nt.w);
    if (kHold)
      mark_hold(q, wq, si, k + 1, mk, lane);
    else
      mark_commit(q, si, k + 1, mk, lane);
  }
  return 9u + static_cast<uint32_t>(nt.n);
}

// Order-preserving map of a float to 32 bits (negative values below
// positive ones) and back.
__device__ inline uint32_t order_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float from_order_bits(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// The update rule (find_neighbor_which_reaches_first, propag.cu:124-190)
// with line-of-sight flags cached in the references.
//
// A neighbour n offers the candidate cand(n) = ct if ct < nt and the line
// from n to its reference is clear, else nt; the rule keeps the first
// neighbour in its order (j outer, i inner) whose candidate is below the
// running best, i.e. the least (cand, order) pair if that candidate beats
// the cell's own time.  cand(n) depends only on n's state, the live time of
// n's reference and static data, so a neighbour whose state and reference
// time are unchanged since the cell's previous evaluation offers the same
// candidate as then, which was no better than the cell's time then and
// cannot be better than its (never larger) time now.  Only neighbours
// changed in sweep k - 1 (chg bit 0, stamps) that can feed the rule (bit 1)
// are evaluated: their (cell, neighbour) pairs are packed over the wave's
// lanes and each cell's least key (order-mapped time, neighbour index,
// reference-or-neighbour bit) is an LDS atomic minimum.  Neighbour (i, j)
// has the index (i + 1) + 4 (j + 1): ordered as the reference's loops (j
// outer, i inner) and decoded with a mask and a shift.
template <int TW>
__device__ inline void packed_rule(TileLds<TW> &s, const Geom &g, int ox, int oy, int lx, int ly, int gx, int gy,
                                   int lane, bool need, Counters &cnt, bool &improved, float &best_t_out,
                                   uint32_t &best_ref_out) {
  constexpr int HW = Tile<TW>::HW;
  uint32_t m = 0;
  cnt.relax += static_cast<uint32_t>(__builtin_popcountll(__ballot(need)));
  if (need) {
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      if (d == 4) continue;
      const int i = d % 3 - 1, j = d / 3 - 1;
      const int nx = gx + i, ny = gy + j;
      const bool inb = nx >= 0 && ny >= 0 && nx < g.W && ny < g.H;
      m |= (inb && s.chg[(lx + i) + (ly + j) * HW] == 3) ? (1u << ((i + 1) + 4 * (j + 1))) : 0u;
    }
  }
  // Exclusive wave scan of the pair counts (0..8), one ballot per bit.
  const uint32_t c = static_cast<uint32_t>(__builtin_popcount(m));
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t excl = 0, total = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const unsigned long long bb = __ballot((c >> b) & 1u);
    excl += static_cast<uint32_t>(__builtin_popcountll(bb & below)) << b;
    total += static_cast<uint32_t>(__builtin_popcountll(bb)) << b;
  }
  s.key[lane] = ~0ull;
  for (uint32_t mm = m, pos = excl; mm; mm &= mm - 1u, ++pos)
    s.pair[pos] = static_cast<uint16_t>((lane << 4) | __builtin_ctz(mm));
  wave_sync();
  for (uint32_t p0 = 0; p0 < total; p0 += kWave) {
    const uint32_t e = p0 + static_cast<uint32_t>(lane);
    if (e < total) {
      const uint32_t pr = s.pair[e];
      const int cl = static_cast<int>(pr >> 4), d = static_cast<int>(pr & 15u);
      const int cx = cl % TW + 1, cy = cl / TW + 1;
      const int tx = ox + cx - 1, ty = oy + cy - 1;
      const int i = (d & 3) - 1, j = (d >> 2) - 1;
      const int nl = (cx + i) + (cy + j) * HW;
      const uint32_t ref = s.ref[nl];
      const int rfx = ref_x<true>(ref), rfy = ref_y<true>(ref);
      if (rfx != tx || rfy != ty) {  // a neighbour referencing this cell offers nothing
        const float4 kf = make_float4(s.ca[nl], s.cb[nl], s.cr[nl], 0.0f);
        const float nt = time_from_adjacent(s.time[nl], i, j, tx, ty, kf, g);
        const float ct = time_from(s.rtime[nl], rfx, rfy, tx, ty, kf, g);
        const bool use_ref = ct < nt && (ref & kClearBit) != 0u;
        const float cand = use_ref ? ct : nt;
        const unsigned long long key = (static_cast<unsigned long long>(order_bits(cand)) << 5) |
                                       static_cast<unsigned long long>(d << 1) | (use_ref ? 1ull : 0ull);
        atomicMin(&s.key[cl], key);
      }
    }
  }
  wave_sync();
  if (need) {
    const int me = lx + ly * HW;
    const float own_t = s.time[me];
    const unsigned long long key = s.key[lane];
    const float cand = from_order_bits(static_cast<uint32_t>(key >> 5));
    improved = key != ~0ull && cand < own_t;  // update_point (propag.cu:192-200)
    if (improved) {
      const int d = static_cast<int>((key >> 1) & 15u);
      const int i = (d & 3) - 1, j = (d >> 2) - 1;
      const uint32_t ref = s.ref[(lx + i) + (ly + j) * HW];
      best_t_out = cand;
      best_ref_out = (key & 1u) ? make_ref(ref_x<true>(ref), ref_y<true>(ref)) : make_ref(gx + i, gy + j);
    }
  }
}

// One sweep over one tile by one wave: lane l owns cell (l % TW, l / TW) of
// the TW x (64 / TW) tile.  Returns, uniformly over the wave, whether any
// cell committed an improvement.
template <bool kCache, bool kLdsBits, int TW>
__device__ bool wave_tile(const Args &a, const Scen &sc, uint32_t si, unsigned lds_off, int tile, int k,
                          TileLds<TW> &s, WalkQueue &wq, Counters &cnt) {
  constexpr bool kDefer = defer_walks<kCache, kLdsBits>();
  constexpr int TH = Tile<TW>::H, HW = Tile<TW>::HW, kHaloCells = Tile<TW>::Cells;
  // Descriptor fields in registers once (sc lives in global memory, so each
  // use of a field after a store would otherwise reload it).
  const Geom g = sc.g;
  // one 16-byte load per coefficient record
  typedef float f4v __attribute__((ext_vector_type(4)));
  const PG_GLOBAL f4v *__restrict__ coef = G(reinterpret_cast<const f4v *>(sc.coef));
  int *__restrict__ stamp = sc.stamp;
  const BitView bv = scen_bits(sc, lds_off);
  // (selects, not sc.st[k & 1]: a dynamic index would put sc in scratch)
  const uint2 *__restrict__ st_in = (k & 1) ? sc.st[1] : sc.st[0];
  uint2 *__restrict__ st_out = (k & 1) ? sc.st[0] : sc.st[1];
  const int tx = tile % sc.ntx, ty = tile / sc.ntx;
  const int ox = tx * TW, oy = ty * TH;
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t tp0 = a.prof ? wall_clock64() : 0;

  // The tile's dependents box (used only if the tile changes).
  // Lanes 0..3: the tile's dependents box; lanes 4, 5: its far-reference cells.
  const uint32_t bx = lane < 4   ? ld_u32(reinterpret_cast<const uint32_t *>(&sc.box[tile]) + lane)
                      : lane < 6 ? ld_u32(reinterpret_cast<const uint32_t *>(&sc.isref[tile]) + (lane - 4))
                                 : 0u;

  // Stage the halo tile (load_points_into_shared_memory, propag.cu:231-261)
  // in two rounds of independent loads: the cells' state words and stamps,
  // then their references' times and stamps (load_point, propag.cu:209-229).
  constexpr int kPer = (kHaloCells + kWave - 1) / kWave;  // halo cells per lane
  uint2 w[kPer];
  int stp[kPer];
  f4v cf[kPer];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int h = lane + it * kWave;
    const int gx = ox + h % HW - 1, gy = oy + h / HW - 1;
    const bool inb = h < kHaloCells && gx >= 0 && gy >= 0 && gx < g.W && gy < g.H;
    const size_t gi = inb ? static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W : 0;
    if (a.cached_stage) {
      w[it] = inb ? ld_state_c(&st_in[gi]) : make_uint2(__float_as_uint(FLT_MAX), kNoRef);
      stp[it] = inb ? ld_i32_c(&stamp[gi]) : kStampNever;
    } else {
      w[it] = inb ? ld_state(&st_in[gi]) : make_uint2(__float_as_uint(FLT_MAX), kNoRef);
      stp[it] = inb ? ld_i32(&stamp[gi]) : kStampNever;
    }
    cf[it] = coef[gi];
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
    if (a.cached_stage) {
      rt[it] = gather ? __uint_as_float(ld_u32_c(&st_in[ri].x)) : FLT_MAX;
      rstp[it] = gather ? ld_i32_c(&stamp[ri]) : kStampNever;
    } else {
      rt[it] = gather ? __uint_as_float(ld_u32(&st_in[ri].x)) : FLT_MAX;
      rstp[it] = gather ? ld_i32(&stamp[ri]) : kStampNever;
    }
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int h = lane + it * kWave;
    if (h < kHaloCells) {
      s.time[h] = __uint_as_float(w[it].x);
      s.ref[h] = w[it].y;
      s.rtime[h] = rt[it];
      s.ca[h] = cf[it].x;  // Coef: a, b, r, null_fire
      s.cb[h] = cf[it].y;
      s.cr[h] = cf[it].w != 0.0f ? __builtin_inff() : cf[it].z;
      const float t = __uint_as_float(w[it].x);
      const bool chg = stp[it] >= k - 1 || rstp[it] >= k - 1;
      const bool src = t < g.max_time && cf[it].w == 0.0f;  // burns before max_time, fire not null
      s.chg[h] = static_cast<uint8_t>((chg ? 1 : 0) | (chg && src ? 2 : 0));
    }
  }
  wave_sync();
  uint64_t tp1 = a.prof ? wall_clock64() : 0;
  uint64_t walk_ticks = 0;

  const int lx = lane % TW + 1, ly = lane / TW + 1;
  const int gx = ox + lx - 1, gy = oy + ly - 1;
  bool need = false;
  if (gx < g.W && gy < g.H) {
#pragma unroll
    for (int j = -1; j < 2; ++j)
#pragma unroll
      for (int i = -1; i < 2; ++i) need = need || (s.chg[(lx + i) + (ly + j) * HW] & 1);
  }
  bool improved = false;
  bool deferred = false;  // improved with a new reference whose walk is queued
  bool far_dep = false;  // improved with a new reference two or more cells away
  uint32_t rtile = 0;    // that reference's tile
  uint32_t rbit = 0;     // and its cell's bit in the tile
  float best_t_out = 0.0f;
  uint32_t best_ref_out = 0u;
  if constexpr (kCache) {
    packed_rule<TW>(s, g, ox, oy, lx, ly, gx, gy, lane, need, cnt, improved, best_t_out, best_ref_out);
  } else if (cnt.relax += static_cast<uint32_t>(__builtin_popcountll(__ballot(need))); need) {
    const int me = lx + ly * HW;
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
        const int nl = (lx + i) + (ly + j) * HW;
        const float n_t = s.time[nl];
        if (!(n_t < g.max_time)) continue;
        const float kr = s.cr[nl];
        if (__builtin_isinf(kr)) continue;  // fire_is_null (propag.h:16-80)
        const float4 kf = make_float4(s.ca[nl], s.cb[nl], kr, 0.0f);
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
            blocked = walk_blocked<kLdsBits>(nx, ny, rfx, rfy, bv, cnt.steps);
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
    improved = best_t < own_t;  // update_point (propag.cu:192-200)
    best_t_out = best_t;
    best_ref_out = best_ref;
  }
  // A changed tile issues the marks of its next-sweep tiles first, so their
  // round trip overlaps the line-of-sight walks below.
  const unsigned long long imp = __ballot(improved);
  const bool changed = imp != 0ull;
  cnt.improve += static_cast<uint32_t>(__builtin_popcountll(imp));
  const bool idle = __ballot(need) == 0ull;  // (a ballot of every lane)
  if (idle) ++cnt.idle;
  if (TW == kTile && a.prof && idle) {
    // Diagnostics: nothing in an idle tile's halo changed; if nothing two
    // cells from the tile changed either, only a far-dependents box queued
    // it.  Lanes 0..43 read the ring of cells two cells from the tile.
    bool near = false;
    if (lane < 4 * kTile + 12) {
      int rx, ry;
      if (lane < kTile + 4) {
        rx = lane - 2, ry = -2;
      } else if (lane < 2 * kTile + 8) {
        rx = lane - (kTile + 4) - 2, ry = kTile + 1;
      } else if (lane < 3 * kTile + 10) {
        rx = -2, ry = lane - (2 * kTile + 8) - 1;
      } else {
        rx = kTile + 1, ry = lane - (3 * kTile + 10) - 1;
      }
      const int rgx = ox + rx, rgy = oy + ry;
      if (rgx >= 0 && rgy >= 0 && rgx < g.W && rgy < g.H)
        near = ld_i32(&stamp[static_cast<size_t>(rgx) + static_cast<size_t>(rgy) * g.W]) >= k - 1;
    }
    if (__ballot(near) == 0ull) ++cnt.idle_far;
  }
  NextTiles nt{0, 0, 1, 0};
  Mark mk0{0u, 0u, 0u, false};
  if (changed) {
    // The far dependents need the next sweep only if a changed cell is
    // some cell's far reference.
    const unsigned long long refs =
        static_cast<unsigned long long>(lane_value(bx, 4)) | (static_cast<unsigned long long>(lane_value(bx, 5)) << 32);
    if (imp & refs) nt = next_tiles<TW>(sc, bx);
    const uint32_t nb =
        kLdsBits ? near_band<TW>(imp) : next_neighbours<kCache, TW>(s, ox, oy, lx, ly, improved, imp, lane);
    mk0 = next_round0_issue<TW>(a.q, sc, si, k, tx, ty, nt, nb, lane);
  }
  if (need) {
    const int me = lx + ly * HW;
    const size_t gi = static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W;
    float best_t = best_t_out;
    uint32_t best_ref = best_ref_out;
    if (improved) {
      const uint64_t tw = a.prof ? wall_clock64() : 0;
      const uint32_t old = s.ref[me];
      const bool same_ref = (old & (kCache ? ~kClearBit : ~0u)) == best_ref;
      if (kCache) {
        // The flag depends only on (cell, reference): an improvement that
        // keeps the reference keeps its flag.
        if (same_ref)
          best_ref = old;
        else if (kDefer)
          deferred = true;
        else if (line_clear<kLdsBits>(gx, gy, ref_x<true>(best_ref), ref_y<true>(best_ref), bv, cnt.steps))
          best_ref |= kClearBit;
      }
      if constexpr (!kDefer) {
        if (a.prof) walk_ticks += wall_clock64() - tw;
      }
      const int brx = ref_x<kCache>(best_ref), bry = ref_y<kCache>(best_ref);
      far_dep = !same_ref && (abs(brx - gx) > 1 || abs(bry - gy) > 1);
      rtile = static_cast<uint32_t>(brx / TW) + static_cast<uint32_t>(bry / TH) * static_cast<uint32_t>(sc.ntx);
      rbit = static_cast<uint32_t>(brx % TW + (bry % TH) * TW);
      if (!deferred) st_state(&st_out[gi], pack_state(best_t, best_ref));
      st_i32(&stamp[gi], k);
    } else {
      st_state(&st_out[gi], pack_state(s.time[me], s.ref[me]));
    }
  }
  if (kDefer) {
    const unsigned long long dm = __ballot(deferred);
    if (dm) {
      const unsigned nd = static_cast<unsigned>(__builtin_popcountll(dm));
      if (wq.n + nd > static_cast<unsigned>(kWalkQ)) {
        const uint64_t tw = a.prof ? wall_clock64() : 0;
        flush_walks(a, k, wq, cnt);
        if (a.prof) walk_ticks += wall_clock64() - tw;
      }
      if (deferred) {
        const unsigned slot = wq.n + static_cast<unsigned>(__builtin_popcountll(dm & ((1ull << lane) - 1ull)));
        wq.q[slot] = WalkReq{si, static_cast<uint32_t>(gx) | (static_cast<uint32_t>(gy) << 16), best_ref_out,
                             best_t_out};
      }
      wq.n += nd;
    }
  }
  uint64_t tp2 = a.prof ? wall_clock64() : 0;
  if (changed) {
    if (kDefer)
      mark_hold(a.q, wq, si, k + 1, mk0, lane);
    else
      mark_commit(a.q, si, k + 1, mk0, lane);
    const uint32_t marks = next_rest<kDefer, TW>(a.q, sc, si, k, nt, wq, lane);
    cnt.marks += __builtin_amdgcn_readfirstlane(marks);
    // Register this tile as a dependent of each new far reference's tile
    // and mark the reference cells: each distinct reference tile's group is
    // gathered in its first lane, which then reads that tile's box and
    // far-reference bits and issues only the atomics that change them.
    // Boxes only grow and bits are only set during a launch, so any value
    // read (also a stale one) lies within the current one: skipping an
    // update it already covers is exact, and the many tiles whose cells take
    // references in the same few tiles (ignitions) no longer queue atomics
    // on the same words.
    unsigned long long m = __ballot(far_dep);
    bool lead = false;
    uint32_t g_rt = 0, g_lo = 0, g_hi = 0;
    uint4 g_box = make_uint4(0u, 0u, 0u, 0u);
    while (m) {
      const int l = __builtin_ctzll(m);
      const uint32_t rt0 = lane_value(rtile, l);
      const bool grp = far_dep && rtile == rt0;
      const unsigned long long gm = __ballot(grp);
      m &= ~gm;
      const uint32_t lo = wave_or((grp && rbit < 32u) ? 1u << rbit : 0u);
      const uint32_t hi = wave_or((grp && rbit >= 32u) ? 1u << (rbit - 32u) : 0u);
      if (lane == l) {
        lead = true;
        g_rt = rt0;
        g_lo = lo;
        g_hi = hi;
        // the group's cells' bounding box, from the ballot (lane = x + TW y)
        unsigned long long cols = gm;
#pragma unroll
        for (int sh = 32; sh >= TW; sh >>= 1) cols |= cols >> sh;
        cols &= TW == 64 ? ~0ull : (1ull << TW) - 1ull;
        g_box = make_uint4(static_cast<uint32_t>(ox + __builtin_ctzll(cols)),
                           static_cast<uint32_t>(oy + __builtin_ctzll(gm) / TW),
                           static_cast<uint32_t>(ox + 63 - __builtin_clzll(cols)),
                           static_cast<uint32_t>(oy + (63 - __builtin_clzll(gm)) / TW));
      }
    }
    if (lead) {
      uint32_t *r = reinterpret_cast<uint32_t *>(&sc.isref[g_rt]);
      uint32_t *b = reinterpret_cast<uint32_t *>(&sc.box[g_rt]);
      // (The one-scenario LDS-bitset kernel runs its sweeps as one chain of
      // round trips and takes the atomics without the read.)
      uint32_t cur[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, 0u};
      if (!kLdsBits) {
#pragma unroll
        for (int q = 0; q < 4; ++q) cur[q] = a.cached_stage ? ld_u32_c(b + q) : ld_u32(b + q);
        cur[4] = a.cached_stage ? ld_u32_c(r) : ld_u32(r);
        cur[5] = a.cached_stage ? ld_u32_c(r + 1) : ld_u32(r + 1);
      }
      if (g_lo & ~cur[4]) or_relaxed(r, g_lo);
      if (g_hi & ~cur[5]) or_relaxed(r + 1, g_hi);
      if (g_box.x < cur[0]) min_relaxed(b + 0, g_box.x);
      if (g_box.y < cur[1]) min_relaxed(b + 1, g_box.y);
      if (g_box.z > cur[2]) max_relaxed(b + 2, g_box.z);
      if (g_box.w > cur[3]) max_relaxed(b + 3, g_box.w);
    }
  }
  ++cnt.tiles;
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

// One list entry (scenario << 32 | tile, uniform over the wave): clears the
// tile's mark, skips released scenarios and those past their sweep budget
// (flagging the latter as capped), runs the tile.  Returns whether it
// changed.
template <bool kCache, bool kLdsBits, int TW>
__device__ inline bool run_entry(const Args &a, unsigned long long ev, unsigned lds_off, int k, TileLds<TW> &s,
                                 WalkQueue &wq, Counters &cnt) {
  const unsigned lane = threadIdx.x & (kWave - 1);
  const uint32_t si = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(ev >> 32));
  const uint32_t tile = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(ev) & kTileMask);
  if (si >= static_cast<uint32_t>(a.nscen)) return false;
  // The descriptors do not change during a launch, and si is wave-uniform:
  // read through the constant address space, the fields come in scalar loads
  // and stay in scalar registers instead of taking vector registers in every
  // lane.
  Scen sc;
  {
    static_assert(sizeof(Scen) % 4 == 0, "copied by words");
    const __attribute__((address_space(4))) uint32_t *src =
        reinterpret_cast<const __attribute__((address_space(4))) uint32_t *>(reinterpret_cast<uintptr_t>(a.scen + si));
    uint32_t *dst = reinterpret_cast<uint32_t *>(&sc);
#pragma unroll
    for (unsigned i = 0; i < sizeof(Scen) / 4; ++i) dst[i] = src[i];
  }
  if (sc.max_iters == 0u) return false;  // released scenario: its buffers are gone
  if (lane == 0) {
    const uint32_t bit = 1u << (tile >> sc.nwords_log2);
    and_relaxed(&mark_of(sc, k)[tile & ((1u << sc.nwords_log2) - 1u)], ~bit);
  }
  if (static_cast<unsigned>(k - sc.k0) >= sc.max_iters) {
    if (lane == 0) st_u32(&a.out[si].capped, 1u);
    return false;
  }
  if (tile >= static_cast<uint32_t>(sc.ntx) * static_cast<uint32_t>(sc.nty)) return false;
  const bool tc = wave_tile<kCache, kLdsBits, TW>(a, sc, si, lds_off, static_cast<int>(tile), k, s, wq, cnt);
  // last_change is stored once per wave, scenario and sweep, and in the
  // persistent kernel once per workgroup as far as its direct-mapped table
  // of (scenario, sweep) pairs tells: in a launch of one large scenario
  // every changed tile would store to the same word, and such stores queue
  // up behind each other at that word's memory channel.
  if (tc && wq.changed_si != si) {
    if (lane == 0) {
      const unsigned long long key = (static_cast<unsigned long long>(si) << 32) | static_cast<uint32_t>(k);
      unsigned long long *slot = wq.lc ? &wq.lc[si & wq.lc_mask] : nullptr;
      if (!slot || *slot != key) {
        if (slot) *slot = key;
        st_i32(&a.out[si].last_change, k);
      }
    }
    wq.changed_si = si;
  }
  return tc;
}

// Persistent form's share of one sweep: this workgroup's entries
// blockIdx.x, blockIdx.x + gridDim.x, ... are handed to its waves one at a
// time from an LDS counter (tiles differ widely in cost, so a fixed split
// leaves waves idle at the barrier); the next entry's load is issued before
// the current tile runs.  Entries fill slots from 0, so a wave stops at the
// first empty slot it draws.
template <bool kCache, bool kLdsBits, int TW>
__device__ bool sweep_wave_dyn(const Args &a, unsigned *next, unsigned lds_off, int k, TileLds<TW> &s, DeferLds *dq,
                               Counters &cnt, unsigned long long *last_change) {
  WalkQueue wq{dq->walk, 0u, dq->entry, 0u, dq->win};
  wq.lc = last_change;
  wq.lc_mask = static_cast<uint32_t>(last_change_entries<kLdsBits>() - 1);
  const unsigned lane = threadIdx.x & (kWave - 1);
  // Workgroups fall in nx classes, w in class w % nx; class c takes the
  // sub-lists c, c + nx, ... and its workgroups share them, workgroup w
  // taking entries w / nx, w / nx + G / nx, ... of each.  With one sub-list
  // per XCD and G a multiple of 8, nx = 8: workgroup w (on XCD w % 8) runs
  // sub-list w % 8.  next[x] counts the draws from sub-list x.
  const unsigned G = gridDim.x;
  const unsigned nx = a.q.nsub > 1 ? (G < kXcds ? G : kXcds) : 1u;
  const unsigned r = blockIdx.x / nx, nr = G / nx;
  bool changed = false;
  for (uint32_t x = blockIdx.x % nx; x < a.q.nsub; x += nx) {
    unsigned long long *list = list_of(a.q, k, x);
    auto draw = [&](size_t &e) {
      unsigned i = 0;
      if (lane == 0) i = atomicAdd(&next[x], 1u);
      i = lane_value(i, 0);
      e = r + static_cast<size_t>(nr) * i;
      return e < a.q.cap ? ld_u64(&list[e]) : kNoEntry;
    };
    size_t e;
    unsigned long long v = draw(e);
    while (v != kNoEntry) {
      if (lane == 0) st_u64(&list[e], kNoEntry);  // written again only four sweeps later
      size_t e2;
      const unsigned long long v2 = draw(e2);  // in flight while this tile runs
      changed |= run_entry<kCache, kLdsBits, TW>(a, v, lds_off, k, s, wq, cnt);
      v = v2;
      e = e2;
    }
  }
  if (defer_walks<kCache, kLdsBits>()) {
    const uint64_t tw = a.prof ? wall_clock64() : 0;
    flush_walks(a, k, wq, cnt);
    if (a.prof) cnt.ph[2] += wall_clock64() - tw;
    flush_entries(a.q, wq, k + 1, static_cast<int>(lane));
  }
  return changed;
}

// One sweep's share of one wave: list entries gwave, gwave + nwaves, ...
// (64 per round, one per lane), each slot emptied and its tile's mark
// cleared after reading.  Entries fill slots from 0, so the first empty slot
// ends the wave's share.  Tiles of a scenario past its sweep budget are
// skipped and the scenario flagged as capped.  Returns whether any tile of
// this wave changed.
template <bool kCache, bool kLdsBits, int TW>
__device__ bool sweep_wave(const Args &a, unsigned gwave, unsigned nwaves, unsigned lds_off, int k, TileLds<TW> &s,
                           DeferLds *dq, Counters &cnt) {
  WalkQueue wq{dq->walk, 0u, dq->entry, 0u, dq->win};
  const unsigned lane = threadIdx.x & (kWave - 1);
  bool changed = false;
  for (uint32_t x = 0; x < a.q.nsub; ++x) {
    unsigned long long *list = list_of(a.q, k, x);
    for (size_t e0 = gwave; e0 < a.q.cap; e0 += static_cast<size_t>(nwaves) * kWave) {
      const size_t e = e0 + static_cast<size_t>(lane) * nwaves;
      unsigned long long v = kNoEntry;
      if (e < a.q.cap) v = ld_u64(&list[e]);
      if (v != kNoEntry) st_u64(&list[e], kNoEntry);  // written again only four sweeps later
      const unsigned long long m = __ballot(v != kNoEntry);
      for (unsigned long long r = m; r;) {
        const int l = __builtin_ctzll(r);
        r &= r - 1;
        changed |= run_entry<kCache, kLdsBits, TW>(a, lane_value(v, l), lds_off, k, s, wq, cnt);
      }
      if (m != ~0ull) break;
    }
  }
  if (defer_walks<kCache, kLdsBits>()) {
    const uint64_t tw = a.prof ? wall_clock64() : 0;
    flush_walks(a, k, wq, cnt);
    if (a.prof) cnt.ph[2] += wall_clock64() - tw;
    flush_entries(a.q, wq, k + 1, static_cast<int>(lane));
  }
  return changed;
}

// Grid barrier.  Callers have waited for their stores (wait_stores) and
// passed __syncthreads.  Every workgroup publishes one 64-bit arrival word,
// (sweep + 1) << 32 | changed << 31 | work ticks, in an array by sweep
// parity (a workgroup that has passed this barrier writes its next arrival
// into the other array, so a slower reader still finds this sweep's word).
// Wave 0 of every workgroup polls all words, a batch of loads at a time and
// only those not yet seen, and folds the quiet-sweep test and the end of the
// launch itself: no release word, so a sweep starts one memory round trip
// after the last arrival.  Workgroup 0 also keeps the books (critical-path
// time, the list counter reset, the launch outcome).  All words are
// device-coherent, so no cache maintenance is needed.  Returns 1 when the
// launch is finished.
__device__ int grid_barrier(const Args &a, int k, bool wg_changed, uint64_t sweep_t0, Counters &cnt,
                            const unsigned long long *wg_ph) {
  __shared__ int s_done;
  __shared__ unsigned long long s_release;  // workgroup 0: time its books were done
  const unsigned target = static_cast<unsigned>(k - a.k_begin) + 1u;
  const unsigned row = static_cast<unsigned>(k - a.k_begin);
  Slot *slots = a.slots;
  Control *c = a.ctl;
  unsigned long long *arrive = a.arrive + (target & 1u) * static_cast<size_t>(a.arrive_stride);
  uint64_t t_arrive = 0;
  if (threadIdx.x < kWave) {  // wave 0 keeps the workgroup's times (wave-uniform)
    t_arrive = wall_clock64();
    cnt.work_ticks += t_arrive - sweep_t0;
  }
  if (threadIdx.x == 0) {
    const uint64_t work = t_arrive - sweep_t0;
    if (a.prof) {
      const unsigned par = target & 1u;
      for (int q = 0; q < 4; ++q) st_u64(&slots[blockIdx.x].ph[par][q], wg_ph[q]);
      st_u64(&slots[blockIdx.x].t_seen[par], sweep_t0);
      st_u64(&slots[blockIdx.x].t_arrive[par], t_arrive);
      wait_stores();
    }
    const uint64_t w31 = work < 0x7fffffffull ? work : 0x7fffffffull;
    st_u64(&arrive[blockIdx.x],
           (static_cast<unsigned long long>(target) << 32) | (wg_changed ? (1ull << 31) : 0ull) | w31);
  }
  if (threadIdx.x < kWave) {
    const unsigned lane = threadIdx.x;
    const unsigned G = gridDim.x;
    const unsigned nq = (G + kWave - 1) / kWave;  // words per lane (at most 64)
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
          v[j] = (q < nq && ((pending >> q) & 1ull)) ? ld_u64(&arrive[lane + q * kWave]) : 0ull;
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
    // Cacheable staging: lines of the state and stamps that this CU's L1 or
    // this XCD's L2 kept from earlier sweeps are stale now that every
    // workgroup's (device-coherent, written-through) stores of this sweep
    // are complete; the acquire drops them before the next sweep stages.
    if (a.cached_stage) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    changed = __ballot(changed) != 0ull;
    timeout = __ballot(timeout) != 0ull;
    const bool done = !changed || k + 1 >= a.k_end || timeout;
    if (blockIdx.x == 0) {
      for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(maxwork, off);
        maxwork = o > maxwork ? o : maxwork;
      }
      if (a.prof && row < a.prof_sweeps) {
        const unsigned long long t_all = wall_clock64();
        // first sweep: the kernel start of workgroup 0 stands in for the books
        const unsigned long long rel = row == 0 ? __shfl(sweep_t0, 0) : s_release;
        unsigned long long m[7] = {0, 0, 0, 0, 0, 0, 0};  // ph[4], start, span, arrival
        const unsigned par = target & 1u;
        for (unsigned i = lane; i < G; i += kWave) {
          for (int q = 0; q < 4; ++q) m[q] = max(m[q], ld_u64(&slots[i].ph[par][q]));
          const unsigned long long ts = ld_u64(&slots[i].t_seen[par]), ta = ld_u64(&slots[i].t_arrive[par]);
          m[4] = max(m[4], ts > rel ? ts - rel : 0ull);
          m[5] = max(m[5], ta > ts ? ta - ts : 0ull);
          m[6] = max(m[6], ta);
        }
        for (int q = 0; q < 7; ++q)
          for (int off = 32; off > 0; off >>= 1) m[q] = max(m[q], __shfl_xor(m[q], off));
        if (lane == 0) {
          unsigned long long *r = a.prof + static_cast<size_t>(row) * kProfCols;
          for (int q = 0; q < 6; ++q) r[q] = m[q];
          r[6] = t_all > m[6] ? t_all - m[6] : 0ull;
          r[7] = t_all;  // completed below with the end of the books
        }
      }
      cnt.crit_ticks += maxwork;  // (reduced over the wave) written out once, at the end of the launch
      if (lane == 0) {
        if (done) {
          st_relaxed(&c->book.v[4], static_cast<unsigned>(k));
          st_relaxed(&c->book.v[1], changed ? 0u : 1u);
        }
        // Sweep k + 2 appends to list (k + 3) & 3, last read in sweep k - 1;
        // these stores complete before this wave's next barrier arrival.
        for (uint32_t x = 0; x < a.q.nsub; ++x) st_relaxed(count_of(a.q, k + 3, x), 0u);
        if (a.prof) {
          s_release = wall_clock64();
          if (row < a.prof_sweeps) {
            unsigned long long *r = a.prof + static_cast<size_t>(row) * kProfCols;
            r[7] = s_release - r[7];
          }
        }
      }
    }
    if (lane == 0) {
      if (timeout) st_relaxed(&c->book.v[3], 1u);
      s_done = done ? 1 : 0;
    }
  }
  __syncthreads();
  if (threadIdx.x < kWave) cnt.wait_ticks += wall_clock64() - t_arrive;
  return s_done;
}

__device__ inline void flush_counters(const Args &a, const Counters &cnt) {
  // The counters are wave-uniform (kept in scalar registers), except the
  // walk steps, which each lane counts for its own walks.  One 64-bit add
  // per counter per wave.
  unsigned long long steps = cnt.steps;
  for (int off = 32; off > 0; off >>= 1) steps += __shfl_xor(steps, off);
  const unsigned long long v[16] = {cnt.relax,      cnt.improve,      steps,            cnt.tiles,
                                    cnt.work_ticks, cnt.wait_ticks,   cnt.crit_ticks,   cnt.marks,
                                    cnt.ph_total[0], cnt.ph_total[1], cnt.ph_total[2], cnt.ph_total[3],
                                    cnt.idle,       cnt.idle_far,     0ull,             0ull};
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (v[q]) atomicAdd(&a.ctl->stats[q], v[q]);
  }
}

// Workgroup size bound of the persistent kernel: with the bitset in LDS at
// most 8 waves fit beside it, so that form is bounded at 512 threads, which
// lets it use up to 256 VGPRs (two waves per SIMD) instead of 128.
template <bool kLdsBits>
constexpr int persistent_threads() { return kLdsBits ? 512 : 1024; }

// Persistent form: one launch runs sweeps k_begin.. until a sweep commits
// nothing or k_end; workgroups meet at the grid barrier after each sweep.
// List entry e belongs to workgroup e mod gridDim.x, so tiles queued
// together (neighbours on the fire front) land on different CUs, and each
// workgroup's waves draw its entries dynamically (sweep_wave_dyn).
template <bool kCache, bool kLdsBits, int TW>
__global__ __launch_bounds__(persistent_threads<kLdsBits>()) void propagate_persistent(Args a) {
  // Per-sweep workgroup flags, double-buffered by sweep parity: sweep k
  // writes [k & 1], thread 0 clears it after barrier k, and it is written
  // again only in sweep k + 2, after barrier k + 1.
  __shared__ unsigned s_changed[2];
  __shared__ unsigned s_next[2][kXcds];  // entries drawn from each sub-list in sweep k (by parity, like s_changed)
  __shared__ unsigned long long s_ph[2][4];
  constexpr int kLastChange = last_change_entries<kLdsBits>();
  __shared__ unsigned long long s_lastc[kLastChange];  // (scenario << 32 | sweep) of the last last_change store
  const int nw = blockDim.x / kWave, wave = threadIdx.x / kWave;
  const unsigned lane = threadIdx.x & (kWave - 1);
  // Dynamic LDS: (LDS bitset) the boundary bitset first; then nw tile
  // slices and (deferred walks) nw
  // blocks of deferred work (walk queue and held list entries).
  const unsigned bits_words = kLdsBits ? (a.lds_words + 3u) & ~3u : 0u;  // 16-byte aligned slices
  TileLds<TW> *slices = reinterpret_cast<TileLds<TW> *>(dyn_lds + bits_words);
  TileLds<TW> &s = slices[wave];
  DeferLds *dq = reinterpret_cast<DeferLds *>(slices + nw) + wave;
  const unsigned lds_off = 0;
  if (threadIdx.x < 2) s_changed[threadIdx.x] = 0;
  if (threadIdx.x < 2 * kXcds) s_next[threadIdx.x / kXcds][threadIdx.x % kXcds] = 0;
  if (threadIdx.x < 8) s_ph[threadIdx.x / 4][threadIdx.x % 4] = 0;
  for (unsigned i = threadIdx.x; i < static_cast<unsigned>(kLastChange); i += blockDim.x) s_lastc[i] = ~0ull;
  if (kLdsBits) {
    // 16-byte chunks, eight loads in flight per thread before their LDS
    // writes (one dependent round trip per word otherwise), then the tail.
    unsigned done = 0;
    if ((reinterpret_cast<uintptr_t>(a.lds_src) & 15u) == 0u) {
      const unsigned n4 = a.lds_words / 4u;
      const uint4 *src4 = reinterpret_cast<const uint4 *>(a.lds_src);
      uint4 *dst4 = reinterpret_cast<uint4 *>(dyn_lds);
      for (unsigned i0 = threadIdx.x; i0 < n4; i0 += 8u * blockDim.x) {
        uint4 v[8];
#pragma unroll
        for (unsigned j = 0; j < 8u; ++j) {
          const unsigned i = i0 + j * blockDim.x;
          if (i < n4) v[j] = src4[i];
        }
#pragma unroll
        for (unsigned j = 0; j < 8u; ++j) {
          const unsigned i = i0 + j * blockDim.x;
          if (i < n4) dst4[i] = v[j];
        }
      }
      done = n4 * 4u;
    }
    for (unsigned i = done + threadIdx.x; i < a.lds_words; i += blockDim.x) dyn_lds[i] = a.lds_src[i];
  }
  __syncthreads();
  Counters cnt{};
  uint64_t t0 = wall_clock64();
  for (int k = a.k_begin;; ++k) {
    const bool changed = sweep_wave_dyn<kCache, kLdsBits, TW>(a, s_next[k & 1], lds_off, k, s, dq, cnt, s_lastc);
    if (lane == 0 && changed) s_changed[k & 1] = 1u;
    if (a.prof) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (lane == 0) atomicMax(&s_ph[k & 1][q], static_cast<unsigned long long>(cnt.ph[q]));
        cnt.ph_total[q] += cnt.ph[q];
        cnt.ph[q] = 0;
      }
    }
    wait_stores();
    __syncthreads();
    const int done = grid_barrier(a, k, s_changed[k & 1] != 0u, t0, cnt, s_ph[k & 1]);
    if (threadIdx.x == 0) s_changed[k & 1] = 0;
    t0 = wall_clock64();
    if (threadIdx.x < kXcds) s_next[k & 1][threadIdx.x] = 0;
    if (threadIdx.x < 4) s_ph[k & 1][threadIdx.x] = 0;
    if (done) break;
  }
  flush_counters(a, cnt);
}

// Plain form: one launch per sweep (the reference's relaunch structure).
template <bool kCache, int TW>
__global__ __launch_bounds__(256) void propagate_sweep(Args a, int k) {
  __shared__ TileLds<TW> slices[256 / kWave];
  __shared__ DeferLds dq[256 / kWave];
  __shared__ unsigned s_changed;
  const int wave = threadIdx.x / kWave;
  if (threadIdx.x == 0) s_changed = 0;
  __syncthreads();
  Counters cnt{};
  const unsigned nw = blockDim.x / kWave;
  const unsigned gwave = wave * gridDim.x + blockIdx.x, nwaves = gridDim.x * nw;
  const bool changed = sweep_wave<kCache, false, TW>(a, gwave, nwaves, 0, k, slices[wave], &dq[wave], cnt);
  if ((threadIdx.x & (kWave - 1)) == 0 && changed) s_changed = 1u;
  __syncthreads();
  if (threadIdx.x == 0 && s_changed) add_relaxed(&a.ctl->changes.v[k & 1], 1u);
  flush_counters(a, cnt);
}

// Between plain sweeps: the bookkeeping the persistent barrier does.
__global__ void sweep_book(Args a, int k) {
  Control *c = a.ctl;
  const unsigned ch = c->changes.v[k & 1];
  c->changes.v[(k + 1) & 1] = 0u;
  for (uint32_t x = 0; x < a.q.nsub; ++x) *count_of(a.q, k + 3, x) = 0u;
  const bool done = ch == 0u || k + 1 >= a.k_end;
  if (done) {
    c->book.v[4] = static_cast<unsigned>(k);
    c->book.v[1] = ch == 0u ? 1u : 0u;
  }
  c->book.v[0] = done ? 1u : 0u;
}

// ---------------------------------------------------------------- set-up

__global__ void fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// time_from's per-cell coefficients of a fire (speed, azimuth, eccentricity).
__device__ inline Coef make_coef(float s, float az, float e) {
  double sn, cs;
  sincos_det(static_cast<double>(az), sn, cs);
  Coef k;
  k.a = static_cast<float>(static_cast<double>(e) * cs);
  k.b = static_cast<float>(static_cast<double>(e) * sn);
  k.r = static_cast<float>(1.0 / (static_cast<double>(s) * (1.0 - static_cast<double>(e))));
  k.null_fire = s == 0.0f ? 1.0f : 0.0f;
  return k;
}

// Fire coefficients, packed double-buffered state, stamps, reference
// validation, the line-of-sight flags of the initial references (needs the
// boundary bitset, so it runs after boundaries_kernel), the dependents boxes
// of the initial far references and the marks of the first sweep (tiles
// holding or touching a burned cell).
__global__ void prep(Scen sc, uint32_t si, TileQueue tq, unsigned *err, const float *speed, const float *az,
                     const float *ecc, const float *time, const uint16_t *rx, const uint16_t *ry) {
  const Geom &g = sc.g;
  const size_t n = static_cast<size_t>(g.W) * g.H;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const_cast<Coef *>(sc.coef)[i] = make_coef(speed[i], az[i], ecc[i]);
    const float t = time[i];
    uint32_t ref = static_cast<uint32_t>(rx[i]) | (static_cast<uint32_t>(ry[i]) << 16);
    const bool burned = t < FLT_MAX;
    const int x = static_cast<int>(i % g.W), y = static_cast<int>(i / g.W);
    if (burned && (rx[i] >= g.W || ry[i] >= g.H)) {
      atomicOr(err, 2u);  // invalid reference
      sc.st[0][i] = pack_state(FLT_MAX, kNoRef);
      sc.st[1][i] = pack_state(FLT_MAX, kNoRef);
      sc.stamp[i] = kStampNever;
      continue;
    }
    if (burned && sc.cache) {
      uint32_t steps = 0;
      if (line_clear<false>(x, y, rx[i], ry[i], scen_bits(sc, 0), steps)) ref |= kClearBit;
    }
    sc.st[0][i] = pack_state(t, ref);
    sc.st[1][i] = pack_state(t, ref);
    sc.stamp[i] = burned ? sc.k0 - 1 : kStampNever;
    if (burned) {
      const int tx = x / sc.tw, ty = y / sc.th;
      for (int q = 0; q < 9; ++q) {
        const int X = tx + q % 3 - 1, Y = ty + q / 3 - 1;
        if (X >= 0 && Y >= 0 && X < sc.ntx && Y < sc.nty) queue_tile(tq, sc, si, sc.k0, X, Y);
      }
      if (abs(static_cast<int>(rx[i]) - x) > 1 || abs(static_cast<int>(ry[i]) - y) > 1) {
        const uint32_t rb = rx[i] % sc.tw + (ry[i] % sc.th) * sc.tw;
        const size_t rt = static_cast<size_t>(rx[i] / sc.tw) + static_cast<size_t>(ry[i] / sc.th) * sc.ntx;
        or_relaxed(reinterpret_cast<uint32_t *>(&sc.isref[rt]) + (rb >> 5), 1u << (rb & 31u));
        uint32_t *b = reinterpret_cast<uint32_t *>(&sc.box[rt]);
        min_relaxed(b + 0, static_cast<uint32_t>(x));
        min_relaxed(b + 1, static_cast<uint32_t>(y));
        max_relaxed(b + 2, static_cast<uint32_t>(x));
        max_relaxed(b + 3, static_cast<uint32_t>(y));
      }
    }
  }
}

// pre_propagate's find_boundaries (propag.cu:345-375): a cell is a boundary
// if an in-bounds neighbour has a dissimilar fire that is faster.  One wave
// covers 64 consecutive cells = two bitset words, so no atomics are needed.
// The boundary bitset in 8x8 blocks (the walks' layout, see blk_at).
__global__ void block_bits(const uint32_t *bits, unsigned long long *blk, uint32_t W, uint32_t H, uint32_t bw,
                           uint32_t bh) {
  const size_t nb = static_cast<size_t>(bw) * bh;
  for (size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < nb;
       b += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const uint32_t x0 = static_cast<uint32_t>(b % bw) * 8u, y0 = static_cast<uint32_t>(b / bw) * 8u;
    unsigned long long v = 0;
    for (uint32_t r = 0; r < 8u && y0 + r < H; ++r)
      for (uint32_t c = 0; c < 8u && x0 + c < W; ++c) {
        const size_t i = static_cast<size_t>(x0 + c) + static_cast<size_t>(y0 + r) * W;
        if ((bits[i >> 5] >> (i & 31u)) & 1u) v |= 1ull << (c | (r << 3));
      }
    blk[b] = v;
  }
}

// find_boundaries' test of one neighbour (propag.cu:345-375,296-302): its
// fire is dissimilar (similar_fires compares in double, as the reference's
// f64 thresholds are, azimuths without wrap-around) and faster.
__device__ inline bool boundary_pair(float s0, float a0, float e0, float s1, float a1, float e1) {
  const bool similar = static_cast<double>(fabsf(s1 - s0)) < 0.1 &&
                       static_cast<double>(fabsf(a1 - a0)) < (5.0 * (2 * 3.14159265358979323846) / 360.0) &&
                       static_cast<double>(fabsf(e1 - e0)) < 0.05;
  return !similar && s0 < s1;
}

// The predicate for cell (x, y): some in-bounds neighbour passes the test.
__device__ inline bool boundary_at(const Geom &g, const float *__restrict__ speed, const float *__restrict__ az,
                                   const float *__restrict__ ecc, int x, int y) {
  const size_t i = static_cast<size_t>(x) + static_cast<size_t>(y) * g.W;
  const float s0 = speed[i], a0 = az[i], e0 = ecc[i];
  bool changed = false;
  for (int di = -1; di < 2; ++di) {
    for (int dj = -1; dj < 2; ++dj) {
      if (di == 0 && dj == 0) continue;
      const int nx = x + di, ny = y + dj;
      if (nx < 0 || ny < 0 || nx >= g.W || ny >= g.H) continue;
      const size_t j = static_cast<size_t>(nx) + static_cast<size_t>(ny) * g.W;
      changed = changed || boundary_pair(s0, a0, e0, speed[j], az[j], ecc[j]);
    }
  }
  return changed;
}

__global__ __launch_bounds__(256) void boundaries_kernel(Geom g, const float *speed, const float *az,
                                                        const float *ecc, uint32_t *bits) {
  const size_t n = static_cast<size_t>(g.W) * g.H;
  const size_t nwords = (n + 31) / 32;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t base = static_cast<size_t>(blockIdx.x) * blockDim.x; base < nwords * 32; base += stride) {
    const size_t i = base + threadIdx.x;
    const bool changed =
        i < n && boundary_at(g, speed, az, ecc, static_cast<int>(i % g.W), static_cast<int>(i / g.W));
    const unsigned long long m = __ballot(changed);
    const int lane = threadIdx.x & 63;
    const size_t w0 = (base + (threadIdx.x & ~63u)) / 32;
    if (lane == 0 && w0 < nwords) bits[w0] = static_cast<uint32_t>(m);
    if (lane == 32 && w0 + 1 < nwords) bits[w0 + 1] = static_cast<uint32_t>(m >> 32);
  }
}

// Point ignitions of an admitted scenario, passed by value (at most
// kInlineIgn distinct cells; more go through init_state / set_ignitions).
constexpr int kInlineIgn = static_cast<int>(kAdmitIgnitions);
struct IgnList {
  uint32_t n;
  uint32_t xy[kInlineIgn];  // x | y << 16
  float t[kInlineIgn];
};

// The initial state of the ignited cells of a scenario whose other cells
// admit_cells has set unburned (propag.rs:433-449 for points): each takes
// its time and itself as reference (never walked: an ignition's own
// reference), its stamp marks it changed before the first sweep, and the
// 3x3 tile neighbourhood of its tile is queued for that sweep.  Thread t of
// one workgroup handles ignition t / 9, neighbour tile t % 9.
__device__ inline void seed_ignitions(const Scen &sc, uint32_t si, const TileQueue &q, const IgnList &ign,
                                      float *time, uint16_t *rx, uint16_t *ry) {
  for (unsigned t = threadIdx.x; t < 9u * ign.n; t += blockDim.x) {
    const unsigned e = t / 9u, d = t % 9u;
    const int x = static_cast<int>(ign.xy[e] & 0xffffu), y = static_cast<int>(ign.xy[e] >> 16);
    if (d == 4u) {
      const size_t i = static_cast<size_t>(x) + static_cast<size_t>(y) * sc.g.W;
      const uint2 w = pack_state(ign.t[e], make_ref(x, y));
      sc.st[0][i] = w;
      sc.st[1][i] = w;
      sc.stamp[i] = sc.k0 - 1;
      time[i] = ign.t[e];
      rx[i] = static_cast<uint16_t>(x);
      ry[i] = static_cast<uint16_t>(y);
    }
    const int X = x / sc.tw + static_cast<int>(d % 3u) - 1, Y = y / sc.th + static_cast<int>(d / 3u) - 1;
    if (X >= 0 && Y >= 0 && X < sc.ntx && Y < sc.nty) queue_tile(q, sc, si, sc.k0, X, Y);
  }
}

// Seeding of the ignitions by itself (the bitset pass below does it on
// rasters whose width is a multiple of 32).
struct Seed {
  Scen sc;
  TileQueue q;
  IgnList ign;
  uint32_t si;
  float *time;
  uint16_t *rx, *ry;
};
__global__ void seed_kernel(Seed a) { seed_ignitions(a.sc, a.si, a.q, a.ign, a.time, a.rx, a.ry); }

// boundaries_kernel and block_bits in one pass, for rasters whose width is a
// multiple of 32: a workgroup covers 32 x 8 cells (four 8x8 blocks), wave w
// its rows 2w and 2w + 1, so each row of 32 cells is exactly one word of the
// row-major bitset and one byte of each of the four blocked words.  With
// `seed`, workgroup (0, 0) also seeds the admitted scenario's ignitions.
__global__ __launch_bounds__(256) void bits_fused(Geom g, const float *speed, const float *az, const float *ecc,
                                                 uint32_t *bits, unsigned long long *blk, uint32_t bw, int seed,
                                                 Seed sd) {
  // The 34 x 10 fire window (tile and one-cell ring) in LDS; cells outside
  // the raster hold a NaN speed, which makes every comparison with them
  // false, as skipping them does.
  constexpr int kHW = 34, kHH = 10;
  __shared__ float hs[kHH][kHW], ha[kHH][kHW], he[kHH][kHW];
  __shared__ uint32_t rows[8];
  const int x0 = static_cast<int>(blockIdx.x) * 32, y0 = static_cast<int>(blockIdx.y) * 8;
  for (int k = static_cast<int>(threadIdx.x); k < kHW * kHH; k += 256) {
    const int hy = k / kHW, hx = k % kHW;
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
    const bool inb = gx >= 0 && gy >= 0 && gx < g.W && gy < g.H;
    const size_t gi = inb ? static_cast<size_t>(gx) + static_cast<size_t>(gy) * g.W : 0;
    hs[hy][hx] = inb ? speed[gi] : __builtin_nanf("");
    ha[hy][hx] = inb ? az[gi] : 0.0f;
    he[hy][hx] = inb ? ecc[gi] : 0.0f;
  }
  __syncthreads();
  const int c = static_cast<int>(threadIdx.x & 31u), r = static_cast<int>(threadIdx.x >> 5);
  const int y = y0 + r;
  bool b = false;
  if (y < g.H) {
    const float s0 = hs[r + 1][c + 1], a0 = ha[r + 1][c + 1], e0 = he[r + 1][c + 1];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      if (q == 4) continue;
      const int hx = c + q % 3, hy = r + q / 3;
      b = b || boundary_pair(s0, a0, e0, hs[hy][hx], ha[hy][hx], he[hy][hx]);
    }
  }
  const unsigned long long m = __ballot(b);
  const unsigned lane = threadIdx.x & 63u;
  if (lane == 0u || lane == 32u) {
    const uint32_t word = static_cast<uint32_t>(lane == 0u ? m : m >> 32);
    rows[r] = word;
    if (y < g.H) bits[(static_cast<size_t>(x0) + static_cast<size_t>(y) * g.W) >> 5] = word;
  }
  __syncthreads();
  if (threadIdx.x < 4u) {
    const unsigned j = threadIdx.x;
    unsigned long long v = 0;
#pragma unroll
    for (int rr = 0; rr < 8; ++rr) v |= static_cast<unsigned long long>((rows[rr] >> (8u * j)) & 0xffu) << (8 * rr);
    blk[static_cast<size_t>(x0 / 8 + static_cast<int>(j)) + static_cast<size_t>(blockIdx.y) * bw] = v;
  }
  if (seed && blockIdx.x == 0 && blockIdx.y == 0) seed_ignitions(sd.sc, sd.si, sd.q, sd.ign, sd.time, sd.rx, sd.ry);
}

// Fills of a call's set-up in one launch: word i of job j is (i & 3) < 2 ?
// v01 : v23 (a plain fill has v01 == v23; the dependents boxes, x0 y0 x1 y1
// per tile, are ~0 ~0 0 0), and the scenario's outcome record is reset.
struct FillJob {
  uint32_t *p;
  size_t n;
  uint32_t v01, v23;
};
constexpr int kMaxFillJobs = 8;
struct FillSet {
  FillJob job[kMaxFillJobs];
  int njobs;
  int k0;
  ScenOut *out;
};

__global__ __launch_bounds__(256) void fill_set(FillSet f) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int j = 0; j < f.njobs; ++j) {
    uint32_t *p = f.job[j].p;
    const uint32_t v01 = f.job[j].v01, v23 = f.job[j].v23;
    for (size_t i = t0; i < f.job[j].n; i += stride) p[i] = (i & 2u) ? v23 : v01;
  }
  if (t0 == 0 && f.out) {
    f.out->last_change = f.k0 - 1;
    f.out->capped = 0u;
    f.out->err = 0u;
  }
}

// Admission of a scenario on a resident terrain with spatially constant
// weather, in one pass over its cells: the burn (burn_weather's rule: cells
// without a catalog fuel get a null fire), time_from's coefficients, every
// cell unburned (state, stamp, result rasters; seed_ignitions then sets the
// ignited cells), and the scenario's mark bitmaps, dependents boxes,
// far-reference bitmaps and outcome record cleared.
struct AdmitArgs {
  Scen sc;
  const uint8_t *fuel;
  const float *slope, *aspect;
  float d1, d10, d100, herb, wood, ws, waz;
  float *speed, *az, *ecc;
  float *time;
  uint16_t *rx, *ry;
  ScenOut *out;
  const DevCatalog *cat;
  size_t mark_words, tiles;
};

__global__ __launch_bounds__(256) void admit_cells(AdmitArgs a) {
  __shared__ DevFuel s_cat[kNumFuels];
  firelib::stage_catalog(s_cat, a.cat);
  const Scen &sc = a.sc;
  const size_t n = static_cast<size_t>(sc.g.W) * sc.g.H;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t t0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint2 unburned = pack_state(FLT_MAX, kNoRef);
  for (size_t i = t0; i < n; i += stride) {
    firelib::Fire f{0.0f, 0.0f, 0.0f};
    (void)firelib::burn_cell(s_cat, a.fuel[i], a.d1, a.d10, a.d100, a.herb, a.wood, a.ws, a.waz, a.slope[i],
                             a.aspect[i], f);
    a.speed[i] = f.speed;
    a.az[i] = f.az;
    a.ecc[i] = f.ecc;
    const_cast<Coef *>(sc.coef)[i] = make_coef(f.speed, f.az, f.ecc);
    sc.st[0][i] = unburned;
    sc.st[1][i] = unburned;
    sc.stamp[i] = kStampNever;
    a.time[i] = FLT_MAX;
    a.rx[i] = 0xFFFFu;
    a.ry[i] = 0xFFFFu;
  }
  for (size_t i = t0; i < a.mark_words; i += stride) sc.mark[0][i] = 0u;  // both parities: contiguous
  for (size_t i = t0; i < a.tiles; i += stride) {
    sc.box[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
    sc.isref[i] = 0ull;
  }
  if (t0 == 0) {
    a.out->last_change = sc.k0 - 1;
    a.out->capped = 0u;
    a.out->err = 0u;
  }
}

// post_propagate (propag.cu:110-117) plus unpacking into the caller's rasters.
// post_propagate's cell rule plus unpacking into the caller's rasters, for
// cells [i0, n) with stride `step`.
__device__ inline void post_cells(const Scen &sc, int final_buf, const float *speed, float *time, uint16_t *rx,
                                  uint16_t *ry, size_t i0, size_t step) {
  const size_t n = static_cast<size_t>(sc.g.W) * sc.g.H;
  const uint2 *st = final_buf ? sc.st[1] : sc.st[0];
  for (size_t i = i0; i < n; i += step) {
    const uint2 w = st[i];
    float t = __uint_as_float(w.x);
    // A cell never reached keeps the caller's references untouched.
    if (t < FLT_MAX) {
      rx[i] = static_cast<uint16_t>(sc.cache ? ref_x<true>(w.y) : ref_x<false>(w.y));
      ry[i] = static_cast<uint16_t>(sc.cache ? ref_y<true>(w.y) : ref_y<false>(w.y));
    }
    if (speed[i] == 0.0f || t > sc.g.max_time) t = FLT_MAX;
    time[i] = t;
  }
}

__global__ void post(Scen sc, int final_buf, const ScenOut *out, const float *speed, float *time, uint16_t *rx,
                     uint16_t *ry) {
  if (out) {
    const ScenOut o = *out;
    final_buf = scen_final_buf(sc, o);
  }
  post_cells(sc, final_buf, speed, time, rx, ry, static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x,
             static_cast<size_t>(gridDim.x) * blockDim.x);
}

// The post-pass of several finished scenarios of a batching context in one
// launch: workgroup row y serves items[y], whose descriptor is scen[slot].
struct PostBatch {
  PostItem item[kPostBatch];
};
__global__ __launch_bounds__(256) void post_batch(const Scen *scen, PostBatch b) {
  const PostItem &it = b.item[blockIdx.y];
  const Scen sc = scen[it.slot];
  post_cells(sc, it.final_buf, it.speed, it.time, it.rx, it.ry,
             static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x, static_cast<size_t>(gridDim.x) * blockDim.x);
}

// Initial state of a scenario from point ignitions (propag.rs:433-449 for
// points): FLT_MAX / 0xFFFF everywhere, then each ignited cell its time and
// itself as reference.  The host passes at most one ignition per cell.
__global__ void init_state(float *time, uint16_t *rx, uint16_t *ry, size_t n) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    time[i] = FLT_MAX;
    rx[i] = 0xFFFFu;
    ry[i] = 0xFFFFu;
  }
}

__global__ void set_ignitions(float *time, uint16_t *rx, uint16_t *ry, const pg_ignition *ign, size_t nign,
                              uint32_t W) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nign) return;
  const pg_ignition g = ign[i];
  const size_t c = static_cast<size_t>(g.x) + static_cast<size_t>(g.y) * W;
  time[c] = g.time;
  rx[c] = static_cast<uint16_t>(g.x);
  ry[c] = static_cast<uint16_t>(g.y);
}

// ------------------------------------------------------------- workspaces

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct DeviceBuffer {
  void *base = nullptr;
  size_t bytes = 0;
};

// Grows *buf to at least `bytes` (contents are not kept).
pg_status ensure(DeviceBuffer *buf, size_t bytes) {
  if (buf->bytes >= bytes) return PG_OK;
  if (buf->base) PG_HIP(hipFree(buf->base));
  buf->base = nullptr;
  buf->bytes = 0;
  hipError_t e = hipMalloc(&buf->base, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PG_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  buf->bytes = bytes;
  return PG_OK;
}

// An instantiated launch graph of the persistent kernel: clear the control
// block and barrier slots, record the start event, run the kernel
// (cooperative node), record the end event, copy the control block to
// pinned host memory.  Built once per kernel variant and launch shape and
// re-parameterised per call.
struct LaunchGraph {
  void (*kern)(Args) = nullptr;
  unsigned groups = 0, block = 0;
  size_t shmem = 0;
  void *base = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t knode = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Control *host = nullptr;
};

// Per device: the buffers of the one-scenario entry point and of a launch
// (control block, barrier slots, profile rows), each behind its own lock,
// and the launch graphs (under the launch lock).
// Pinned host memory for a call's small device-to-host results: copies into
// pageable memory go through the runtime's staging (the stream is drained
// first), pinned ones are ordinary asynchronous copies.  The launch part is
// used under launch_mu, the call part under call_mu.
struct HostReturn {
  alignas(128) Control ctl;
  alignas(128) ScenOut out;
  alignas(128) Scen desc;  // the call's descriptor, uploaded from here
};

// A persistent kernel's resident workgroups per CU for a launch shape.
struct Occupancy {
  const void *kern;
  int block;
  size_t shmem;
  int per_cu;
};

struct PerDevice {
  std::mutex call_mu, launch_mu;
  DeviceBuffer call, launch;
  std::vector<LaunchGraph *> graphs;
  HostReturn *ret = nullptr;  // pinned, or null (then pageable copies)
  // Timing events kept across calls: [0, 1] the call's (under call_mu),
  // [2, 3] the relaxation launch's (under launch_mu).
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<Occupancy> occ;  // under launch_mu
};

// The device's timing event i, created on first use.
pg_status device_event(PerDevice *pd, int i, hipEvent_t *out) {
  if (!pd->ev[i]) PG_HIP(hipEventCreate(&pd->ev[i]));
  *out = pd->ev[i];
  return PG_OK;
}

std::mutex g_dev_mu;
std::vector<PerDevice *> g_dev;

PerDevice *per_device(int device) {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  if (g_dev.size() <= static_cast<size_t>(device)) g_dev.resize(device + 1, nullptr);
  if (!g_dev[device]) {
    g_dev[device] = new PerDevice();
    void *p = nullptr;
    if (hipHostMalloc(&p, sizeof(HostReturn)) == hipSuccess)
      g_dev[device]->ret = static_cast<HostReturn *>(p);
    else
      (void)hipGetLastError();
  }
  return g_dev[device];
}

constexpr unsigned kProfRows = 8192;

pg_status build_graph(LaunchGraph *lg, const Args &a, size_t max_groups) {
  PG_HIP(hipGraphCreate(&lg->graph, 0));
  PG_HIP(hipEventCreate(&lg->ev0));
  PG_HIP(hipEventCreate(&lg->ev1));
  PG_HIP(hipHostMalloc(reinterpret_cast<void **>(&lg->host), sizeof(Control), hipHostMallocDefault));
  uint32_t *ctl = reinterpret_cast<uint32_t *>(a.ctl);
  uint32_t *slots = reinterpret_cast<uint32_t *>(a.slots);
  size_t ctl_words = sizeof(Control) / 4, slot_words = slot_bytes(max_groups) / 4;
  uint32_t zero = 0;
  void *args_ctl[] = {&ctl, &ctl_words, &zero};
  void *args_slots[] = {&slots, &slot_words, &zero};
  hipKernelNodeParams p{};
  p.func = reinterpret_cast<void *>(fill_u32);
  p.blockDim = dim3(256);
  p.gridDim = dim3(1);
  p.kernelParams = args_ctl;
  hipGraphNode_t n_ctl, n_slots, n_ev0, n_ev1, n_copy;
  PG_HIP(hipGraphAddKernelNode(&n_ctl, lg->graph, nullptr, 0, &p));
  p.gridDim = dim3(static_cast<unsigned>((slot_words + 255) / 256));
  p.kernelParams = args_slots;
  PG_HIP(hipGraphAddKernelNode(&n_slots, lg->graph, nullptr, 0, &p));
  hipGraphNode_t fills[2] = {n_ctl, n_slots};
  PG_HIP(hipGraphAddEventRecordNode(&n_ev0, lg->graph, fills, 2, lg->ev0));
  Args ak = a;
  void *args_k[] = {&ak};
  hipKernelNodeParams pk{};
  pk.func = reinterpret_cast<void *>(lg->kern);
  pk.gridDim = dim3(lg->groups);
  pk.blockDim = dim3(lg->block);
  pk.sharedMemBytes = static_cast<unsigned>(lg->shmem);
  pk.kernelParams = args_k;
  PG_HIP(hipGraphAddKernelNode(&lg->knode, lg->graph, &n_ev0, 1, &pk));
  hipKernelNodeAttrValue coop{};
  coop.cooperative = 1;
  PG_HIP(hipGraphKernelNodeSetAttribute(lg->knode, hipKernelNodeAttributeCooperative, &coop));
  PG_HIP(hipGraphAddEventRecordNode(&n_ev1, lg->graph, &lg->knode, 1, lg->ev1));
  PG_HIP(hipGraphAddMemcpyNode1D(&n_copy, lg->graph, &n_ev1, 1, lg->host, a.ctl, sizeof(Control),
                                 hipMemcpyDeviceToHost));
  PG_HIP(hipGraphInstantiate(&lg->exec, lg->graph, nullptr, nullptr, 0));
  return PG_OK;
}

void free_graph(LaunchGraph *lg) {
  if (lg->exec) (void)hipGraphExecDestroy(lg->exec);
  if (lg->graph) (void)hipGraphDestroy(lg->graph);
  if (lg->ev0) (void)hipEventDestroy(lg->ev0);
  if (lg->ev1) (void)hipEventDestroy(lg->ev1);
  if (lg->host) (void)hipHostFree(lg->host);
  delete lg;
}

}  // namespace

// Frees every per-device workspace and cached launch graph (the next call
// recreates what it needs).  Runs with no propagation in flight.
void release_launch_resources() {
  std::lock_guard<std::mutex> lk(g_dev_mu);
  int current = 0;
  const bool have = hipGetDevice(&current) == hipSuccess;
  for (size_t d = 0; d < g_dev.size(); ++d) {
    PerDevice *pd = g_dev[d];
    if (!pd) continue;
    if (hipSetDevice(static_cast<int>(d)) != hipSuccess) continue;
    (void)hipDeviceSynchronize();
    for (LaunchGraph *lg : pd->graphs) free_graph(lg);
    pd->graphs.clear();
    for (DeviceBuffer *b : {&pd->call, &pd->launch}) {
      if (b->base) (void)hipFree(b->base);
      b->base = nullptr;
      b->bytes = 0;
    }
    if (pd->ret) (void)hipHostFree(pd->ret);
    pd->ret = nullptr;
    for (hipEvent_t &e : pd->ev) {
      if (e) (void)hipEventDestroy(e);
      e = nullptr;
    }
    pd->occ.clear();
  }
  if (have) (void)hipSetDevice(current);
  (void)hipGetLastError();
}

namespace {

// The cached graph for this launch shape (built on first use).
pg_status launch_graph_for(PerDevice *pd, void (*kern)(Args), unsigned groups, unsigned block, size_t shmem,
                           const Args &a, size_t max_groups, LaunchGraph **out) {
  for (LaunchGraph *lg : pd->graphs)
    if (lg->kern == kern && lg->groups == groups && lg->block == block && lg->shmem == shmem &&
        lg->base == pd->launch.base) {
      *out = lg;
      return PG_OK;
    }
  auto *lg = new LaunchGraph();
  lg->kern = kern;
  lg->groups = groups;
  lg->block = block;
  lg->shmem = shmem;
  lg->base = pd->launch.base;
  pg_status st = build_graph(lg, a, max_groups);
  if (st != PG_OK) {
    free_graph(lg);
    return st;
  }
  pd->graphs.push_back(lg);
  *out = lg;
  return PG_OK;
}

}  // namespace

// ------------------------------------------------------------- scenarios

// Buffers of one scenario, for any tile shape: boxes and far-reference
// bits for the most tiles a shape cuts the raster into, the blocked bitset
// in 8x8 blocks.
size_t scen_buffer_bytes(size_t W, size_t H) {
  const size_t n = W * H;
  const size_t ntiles = max_tiles(W, H);
  const size_t nblocks = ((W + 7) / 8) * ((H + 7) / 8);
  size_t off = 0;
  off = align_up(off + n * sizeof(Coef), 256);
  off = align_up(off + n * sizeof(uint2), 256);
  off = align_up(off + n * sizeof(uint2), 256);
  off = align_up(off + n * sizeof(int), 256);
  off = align_up(off + ntiles * sizeof(uint4), 256);
  off = align_up(off + nblocks * sizeof(unsigned long long), 256);
  off = align_up(off + ntiles * sizeof(unsigned long long), 256);
  return off;
}

ScenBuffers scen_carve(void *base, size_t W, size_t H) {
  const size_t n = W * H;
  char *p = static_cast<char *>(base);
  size_t off = 0;
  ScenBuffers b;
  b.coef = reinterpret_cast<Coef *>(p + off); off = align_up(off + n * sizeof(Coef), 256);
  b.st0 = reinterpret_cast<uint2 *>(p + off); off = align_up(off + n * sizeof(uint2), 256);
  b.st1 = reinterpret_cast<uint2 *>(p + off); off = align_up(off + n * sizeof(uint2), 256);
  b.stamp = reinterpret_cast<int *>(p + off); off = align_up(off + n * sizeof(int), 256);
  const size_t ntiles = max_tiles(W, H);
  const size_t nblocks = ((W + 7) / 8) * ((H + 7) / 8);
  b.box = reinterpret_cast<uint4 *>(p + off); off = align_up(off + ntiles * sizeof(uint4), 256);
  b.blk = reinterpret_cast<unsigned long long *>(p + off); off = align_up(off + nblocks * 8, 256);
  b.isref = reinterpret_cast<unsigned long long *>(p + off);
  return b;
}

pg_status scen_geometry(const Settings *settings, Scen *sc, int tw) {
  if (!settings) return set_error(PG_ERR_INVALID, "settings is NULL");
  if (!tile_shape_ok(static_cast<uint32_t>(tw), static_cast<uint32_t>(kWave / tw)))
    return set_error(PG_ERR_INVALID, "tile shape must be 8x8, 16x4 or 32x2");
  const uint32_t W = settings->geo_ref.width, H = settings->geo_ref.height;
  if (W == 0 || H == 0) return set_error(PG_ERR_INVALID, "raster has no cells");
  if (W > 0xFFFFu || H > 0xFFFFu)
    return set_error(PG_ERR_INVALID, "raster sides must be below 65536 (16-bit references)");
  // A non-finite max_time would let unburned cells (FLT_MAX, no reference)
  // count as burning neighbours.
  if (!std::isfinite(settings->max_time)) return set_error(PG_ERR_INVALID, "max_time must be finite");
  const float gdx = settings->geo_ref.transform.gt.dx, gdy = settings->geo_ref.transform.gt.dy;
  *sc = Scen{};
  sc->g.W = static_cast<int>(W);
  sc->g.H = static_cast<int>(H);
  sc->g.dx = gdx;
  sc->g.dy = gdy;
  sc->g.cell = fabsf(gdx);
  sc->g.max_time = settings->max_time;
  sc->g.north_up = gdy < 0.0f;
  sc->g.square = fabsf(gdx) == fabsf(gdy);
  sc->cache = (W <= 0x7fffu && H <= 0x7fffu) ? 1 : 0;
  sc->tw = tw;
  sc->th = kWave / tw;
  sc->ntx = static_cast<int>((W + sc->tw - 1) / sc->tw);
  sc->nty = static_cast<int>((H + sc->th - 1) / sc->th);
  sc->bw = static_cast<int>((W + 7) / 8);
  sc->bh = static_cast<int>((H + 7) / 8);
  sc->nwords_log2 = need_words_log2(static_cast<size_t>(sc->ntx) * sc->nty);
  sc->max_iters = 64u * (W + H) + 1024u;
  return PG_OK;
}

namespace {

dim3 setup_grid(const DeviceInfo *dev, size_t items, unsigned per_cu) {
  size_t b = (items + 255) / 256;
  const size_t cap = static_cast<size_t>(dev->cu_count) * per_cu;
  if (b > cap) b = cap;
  return dim3(static_cast<unsigned>(b ? b : 1));
}

// The boundary bitset of a scenario in both layouts (row-major `bits`,
// blocked sc.blk), seeding the ignitions of `seed` if given.
pg_status launch_bits(const DeviceInfo *dev, const Scen &sc, const float *speed, const float *az, const float *ecc,
                      uint32_t *bits, const Seed *seed, hipStream_t stream) {
  unsigned long long *blk = const_cast<unsigned long long *>(sc.blk);
  if (sc.g.W % 32 == 0) {
    const dim3 grid(static_cast<unsigned>(sc.g.W / 32), static_cast<unsigned>((sc.g.H + 7) / 8));
    hipLaunchKernelGGL(bits_fused, grid, dim3(256), 0, stream, sc.g, speed, az, ecc, bits, blk,
                       static_cast<uint32_t>(sc.bw), seed ? 1 : 0, seed ? *seed : Seed{});
  } else {
    const size_t n = static_cast<size_t>(sc.g.W) * sc.g.H;
    hipLaunchKernelGGL(boundaries_kernel, setup_grid(dev, (n + 31) / 32 * 32, 8), dim3(256), 0, stream, sc.g, speed,
                       az, ecc, bits);
    hipLaunchKernelGGL(block_bits, setup_grid(dev, static_cast<size_t>(sc.bw) * sc.bh, 8), dim3(256), 0, stream,
                       bits, blk, sc.g.W, sc.g.H, static_cast<uint32_t>(sc.bw), static_cast<uint32_t>(sc.bh));
    if (seed) hipLaunchKernelGGL(seed_kernel, dim3(1), dim3(256), 0, stream, *seed);
  }
  PG_HIP(hipGetLastError());
  return PG_OK;
}

}  // namespace

pg_status scen_prepare(const DeviceInfo *dev, const Scen &sc, uint32_t si, ScenOut *out, const TileQueue &q,
                       const float *speed, const float *az, const float *ecc, const float *time,
                       const uint16_t *refs_x, const uint16_t *refs_y, uint32_t *boundaries, bool validate,
                       hipStream_t stream, bool clear_queue) {
  const size_t n = static_cast<size_t>(sc.g.W) * sc.g.H;
  const size_t ntiles = static_cast<size_t>(sc.ntx) * sc.nty;
  FillSet f{};
  auto job = [&](void *p, size_t words, uint32_t v01, uint32_t v23) {
    f.job[f.njobs++] = FillJob{static_cast<uint32_t *>(p), words, v01, v23};
  };
  job(sc.mark[0], 2 * (size_t{1} << sc.nwords_log2), 0u, 0u);  // both parities: contiguous
  job(sc.box, 4 * ntiles, 0xFFFFFFFFu, 0u);
  job(sc.isref, 2 * ntiles, 0u, 0u);
  if (clear_queue) {
    job(q.list, 8 * static_cast<size_t>(q.nsub) * q.cap, 0xFFFFFFFFu, 0xFFFFFFFFu);
    job(q.count, 128 * static_cast<size_t>(q.nsub), 0u, 0u);
  }
  f.out = out;
  f.k0 = sc.k0;
  size_t most = 0;
  for (int j = 0; j < f.njobs; ++j) most = std::max(most, f.job[j].n);
  hipLaunchKernelGGL(fill_set, setup_grid(dev, most, 8), dim3(256), 0, stream, f);
  PG_HIP(hipGetLastError());
  pg_status st = launch_bits(dev, sc, speed, az, ecc, boundaries, nullptr, stream);
  if (st != PG_OK) return st;
  hipLaunchKernelGGL(prep, setup_grid(dev, n, 8), dim3(256), 0, stream, sc, si, q, &out->err, speed, az, ecc, time,
                     refs_x, refs_y);
  PG_HIP(hipGetLastError());
  unsigned e = 0;
  if (validate) {
    PG_HIP(hipMemcpyAsync(&e, &out->err, sizeof(e), hipMemcpyDeviceToHost, stream));
    PG_HIP(hipStreamSynchronize(stream));
  }
  if (e & 2u) return set_error(PG_ERR_INVALID, "a burned cell (time < FLT_MAX) has a reference outside the raster");
  return PG_OK;
}

pg_status scen_admit(const DeviceInfo *dev, const Scen &sc, uint32_t si, ScenOut *out, const TileQueue &q,
                     const uint8_t *fuel, const float *slope, const float *aspect, const float moisture[5],
                     float wind_speed, float wind_azimuth, float *speed, float *az, float *ecc, float *time,
                     uint16_t *refs_x, uint16_t *refs_y, uint32_t *boundaries, const pg_ignition *ign, size_t n_ign,
                     hipStream_t stream) {
  if (n_ign > static_cast<size_t>(kInlineIgn)) return set_error(PG_ERR_INVALID, "scen_admit: too many ignitions");
  const size_t n = static_cast<size_t>(sc.g.W) * sc.g.H;
  AdmitArgs a{};
  a.sc = sc;
  a.fuel = fuel;
  a.slope = slope;
  a.aspect = aspect;
  a.d1 = moisture[0];
  a.d10 = moisture[1];
  a.d100 = moisture[2];
  a.herb = moisture[3];
  a.wood = moisture[4];
  a.ws = wind_speed;
  a.waz = wind_azimuth;
  a.speed = speed;
  a.az = az;
  a.ecc = ecc;
  a.time = time;
  a.rx = refs_x;
  a.ry = refs_y;
  a.out = out;
  a.cat = dev->catalog;
  a.mark_words = 2 * (size_t{1} << sc.nwords_log2);
  a.tiles = static_cast<size_t>(sc.ntx) * sc.nty;
  hipLaunchKernelGGL(admit_cells, setup_grid(dev, n, 16), dim3(256), 0, stream, a);
  PG_HIP(hipGetLastError());
  Seed sd{};
  sd.sc = sc;
  sd.q = q;
  sd.si = si;
  sd.time = time;
  sd.rx = refs_x;
  sd.ry = refs_y;
  sd.ign.n = static_cast<uint32_t>(n_ign);
  for (size_t i = 0; i < n_ign; ++i) {
    sd.ign.xy[i] = ign[i].x | (ign[i].y << 16);
    sd.ign.t[i] = ign[i].time;
  }
  return launch_bits(dev, sc, speed, az, ecc, boundaries, &sd, stream);
}

pg_status scen_init_state(float *time, uint16_t *refs_x, uint16_t *refs_y, size_t n, uint32_t width,
                          const pg_ignition *ignitions, size_t n_ignitions, hipStream_t stream) {
  size_t b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  hipLaunchKernelGGL(init_state, dim3(static_cast<unsigned>(b ? b : 1)), dim3(256), 0, stream, time, refs_x, refs_y,
                     n);
  if (n_ignitions)
    hipLaunchKernelGGL(set_ignitions, dim3(static_cast<unsigned>((n_ignitions + 255) / 256)), dim3(256), 0, stream,
                       time, refs_x, refs_y, ignitions, n_ignitions, width);
  PG_HIP(hipGetLastError());
  return PG_OK;
}

pg_status scen_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t stream) {
  size_t b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  hipLaunchKernelGGL(fill_u32, dim3(static_cast<unsigned>(b ? b : 1)), dim3(256), 0, stream, p, n, v);
  PG_HIP(hipGetLastError());
  return PG_OK;
}

pg_status scen_finish_batch(const DeviceInfo *dev, const Scen *d_scen, const PostItem *items, size_t n_items,
                            size_t max_cells, hipStream_t stream) {
  for (size_t i0 = 0; i0 < n_items; i0 += kPostBatch) {
    const size_t m = std::min<size_t>(kPostBatch, n_items - i0);
    PostBatch b{};
    for (size_t j = 0; j < m; ++j) b.item[j] = items[i0 + j];
    // enough workgroups per scenario to fill the device with all m of them
    size_t bx = (max_cells + 255) / 256;
    const size_t cap = std::max<size_t>(1, static_cast<size_t>(dev->cu_count) * 16 / m);
    if (bx > cap) bx = cap;
    hipLaunchKernelGGL(post_batch, dim3(static_cast<unsigned>(bx), static_cast<unsigned>(m)), dim3(256), 0, stream,
                       d_scen, b);
    PG_HIP(hipGetLastError());
  }
  return PG_OK;
}

pg_status scen_finish(const Scen &sc, int final_buf, const float *speed, float *time, uint16_t *refs_x,
                      uint16_t *refs_y, hipStream_t stream, const ScenOut *out) {
  const size_t n = static_cast<size_t>(sc.g.W) * sc.g.H;
  size_t b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  hipLaunchKernelGGL(post, dim3(static_cast<unsigned>(b)), dim3(256), 0, stream, sc, final_buf, out, speed, time,
                     refs_x, refs_y);
  PG_HIP(hipGetLastError());
  return PG_OK;
}

// ---------------------------------------------------------------- launches

static pg_status finish_stats(const DeviceInfo *dev, const unsigned *book, const unsigned long long *counters,
                              float rms, unsigned launched_groups, const Args &a, unsigned prof_rows,
                              pg_stats *stats, LaunchResult *res);

size_t launch_ws_bytes(unsigned cu, bool profile) {
  const size_t max_groups = static_cast<size_t>(cu) * 32;
  size_t off = align_up(sizeof(Control), 256);
  off = align_up(off + slot_bytes(max_groups), 256);
  if (profile) off = align_up(off + static_cast<size_t>(kProfRows) * kProfCols * 8, 256);
  return off;
}

// The kernels and LDS slice of a tile width.
static size_t tile_lds_bytes(int tw) {
  return tw == 16 ? sizeof(TileLds<16>) : tw == 32 ? sizeof(TileLds<32>) : sizeof(TileLds<8>);
}
template <int TW>
static void (*persistent_for(bool cache, bool lds))(Args) {
  return cache ? (lds ? propagate_persistent<true, true, TW> : propagate_persistent<true, false, TW>)
               : (lds ? propagate_persistent<false, true, TW> : propagate_persistent<false, false, TW>);
}
static void (*persistent_kernel(int tw, bool cache, bool lds))(Args) {
  return tw == 16 ? persistent_for<16>(cache, lds) : tw == 32 ? persistent_for<32>(cache, lds)
                                                             : persistent_for<8>(cache, lds);
}
template <int TW>
static void (*sweep_for(bool cache))(Args, int) {
  return cache ? propagate_sweep<true, TW> : propagate_sweep<false, TW>;
}
static void (*sweep_kernel(int tw, bool cache))(Args, int) {
  return tw == 16 ? sweep_for<16>(cache) : tw == 32 ? sweep_for<32>(cache) : sweep_for<8>(cache);
}

pg_status run_sweeps(const DeviceInfo *dev, const Launch &L, const pg_tuning *tuning, pg_stats *stats,
                     LaunchResult *res, hipStream_t stream) {
  pg_tuning tn{};
  if (tuning) tn = *tuning;
  const int tw = L.tw ? L.tw : kTile;
  if ((tn.tile_w || tn.tile_h) && (static_cast<int>(tn.tile_w) != tw || static_cast<int>(tn.tile_h) != kWave / tw))
    return set_error(PG_ERR_INVALID, "the tuning's tile shape differs from the scenarios' tiling");
  if (L.k_end <= L.k_begin) return set_error(PG_ERR_INVALID, "empty sweep range");
  // A caller's workspace has no room for profile rows and no cached graphs.
  if (L.ws && (tn.profile || tn.mode == 2))
    return set_error(PG_ERR_INVALID, "profiling and the graph launch form need the library's own workspace");
  PerDevice *pd = L.ws ? nullptr : per_device(dev->device);
  std::unique_lock<std::mutex> lk;
  if (pd) lk = std::unique_lock<std::mutex>(pd->launch_mu);

  const unsigned cu = static_cast<unsigned>(dev->cu_count);
  const size_t max_groups = static_cast<size_t>(cu) * 32;
  const unsigned prof_rows = tn.profile ? kProfRows : 0u;
  const size_t o_ctl = 0;
  const size_t o_slots = align_up(sizeof(Control), 256);
  const size_t o_prof = align_up(o_slots + slot_bytes(max_groups), 256);
  pg_status st = PG_OK;
  char *base = L.ws;
  if (!base) {
    st = ensure(&pd->launch, launch_ws_bytes(cu, true));
    if (st != PG_OK) return st;
    base = static_cast<char *>(pd->launch.base);
  }

  Args a{};
  a.scen = L.scen;
  a.out = L.out;
  a.q = L.q;
  a.nscen = L.nscen;
  a.k_begin = L.k_begin;
  a.k_end = L.k_end;
  a.ctl = reinterpret_cast<Control *>(base + o_ctl);
  a.slots = reinterpret_cast<Slot *>(base + o_slots);
  a.arrive = reinterpret_cast<unsigned long long *>(a.slots + max_groups);
  a.arrive_stride = static_cast<unsigned>(max_groups);
  // 0: the default 2^22 polls; 32 and above: none (a barrier that finds a
  // workgroup not yet arrived gives up at once: tests of the failure path)
  a.spin_limit = tn.spin_limit_log2 >= 32 ? 0u
                 : 1u << (tn.spin_limit_log2 ? (tn.spin_limit_log2 > 30 ? 30 : tn.spin_limit_log2) : 22);
  a.prof = prof_rows ? reinterpret_cast<unsigned long long *>(base + o_prof) : nullptr;
  a.prof_sweeps = prof_rows;

  auto grid_for = [&](size_t items) {
    size_t b = (items + 255) / 256;
    if (b > cu * 8) b = cu * 8;
    return dim3(static_cast<unsigned>(b ? b : 1));
  };
  // mode 2: the persistent launch through a cached hipGraph (no profiling)
  const bool use_graph = tn.mode == 2 && !prof_rows;
  if (tn.mode > 2) return set_error(PG_ERR_INVALID, "mode must be 0, 1 or 2");
  if (!use_graph) {
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, stream, reinterpret_cast<uint32_t *>(a.ctl),
                       sizeof(Control) / 4, 0u);
    hipLaunchKernelGGL(fill_u32, grid_for(slot_bytes(max_groups) / 4), dim3(256), 0, stream,
                       reinterpret_cast<uint32_t *>(a.slots), slot_bytes(max_groups) / 4, 0u);
  }
  if (prof_rows)
    hipLaunchKernelGGL(fill_u32, grid_for(static_cast<size_t>(prof_rows) * kProfCols * 2), dim3(256), 0, stream,
                       reinterpret_cast<uint32_t *>(a.prof), static_cast<size_t>(prof_rows) * kProfCols * 2, 0u);
  PG_HIP(hipGetLastError());

  hipEvent_t evr0 = nullptr, evr1 = nullptr;
  hipEvent_t own0 = nullptr, own1 = nullptr;  // created for this launch only (caller's workspace)
  if (!L.untimed) {
    if (pd) {
      st = device_event(pd, 2, &evr0);
      if (st == PG_OK) st = device_event(pd, 3, &evr1);
      if (st != PG_OK) return st;
    } else {
      PG_HIP(hipEventCreate(&own0));
      PG_HIP(hipEventCreate(&own1));
      evr0 = own0;
      evr1 = own1;
    }
  }
  struct EventGuard {
    hipEvent_t *e[2];
    ~EventGuard() {
      for (hipEvent_t *p : e)
        if (*p) (void)hipEventDestroy(*p);
    }
  } guard{{&own0, &own1}};

  unsigned book[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned launched_groups = 0;
  if (evr0) PG_HIP(hipEventRecord(evr0, stream));
  // One launch per sweep, the host reading the bookkeeping in between.
  auto run_per_sweep = [&]() -> pg_status {
    const dim3 grid(cu * 4);
    launched_groups = grid.x;
    a.lds_bits = 0;
    a.cached_stage = 0;
    void (*sweep)(Args, int) = sweep_kernel(tw, L.cache != 0);
    for (int k = L.k_begin;; ++k) {
      hipLaunchKernelGGL(sweep, grid, dim3(256), 0, stream, a, k);
      hipLaunchKernelGGL(sweep_book, dim3(1), dim3(1), 0, stream, a, k);
      PG_HIP(hipGetLastError());
      PG_HIP(hipMemcpyAsync(book, &a.ctl->book, sizeof(book), hipMemcpyDeviceToHost, stream));
      PG_HIP(hipStreamSynchronize(stream));
      if (book[0]) break;
    }
    if (evr1) PG_HIP(hipEventRecord(evr1, stream));
    return PG_OK;
  };
  // After the relaxation: the caller's follow-up work, the control block
  // (books and counters) to the host, one synchronisation.
  Control hc{};
  auto finish_launch = [&](bool copy_ctl) -> pg_status {
    if (L.after) {
      const pg_status s2 = L.after(L.after_ctx, stream);
      if (s2 != PG_OK) return s2;
    }
    Control *dst = (pd && pd->ret) ? &pd->ret->ctl : &hc;
    if (copy_ctl) PG_HIP(hipMemcpyAsync(dst, a.ctl, sizeof(Control), hipMemcpyDeviceToHost, stream));
    PG_HIP(hipStreamSynchronize(stream));
    if (copy_ctl && dst != &hc) std::memcpy(&hc, dst, sizeof(Control));
    return PG_OK;
  };
  if (tn.mode == 1) {
    st = run_per_sweep();
    if (st != PG_OK) return st;
  } else {
    // Waves per workgroup: as many tile slices as fit beside the bitset when
    // the bitset can live in LDS (one workgroup per CU), else 16 waves.
    const size_t bits_bytes = ((L.lds_bits_words + 3) & ~size_t{3}) * sizeof(uint32_t);  // padded as the kernel lays it out
    const size_t static_lds = 320;  // the persistent kernel's own (static) LDS, rounded up
    int waves = 16;
    a.lds_bits = 0;
    const size_t slice = tile_lds_bytes(tw);
    if (L.nscen == 1 && L.lds_bits_src && bits_bytes + static_lds + slice * 4 <= dev->lds_per_block) {
      a.lds_bits = 1;
      a.lds_src = L.lds_bits_src;
      a.lds_words = static_cast<unsigned>(L.lds_bits_words);
      waves = static_cast<int>((dev->lds_per_block - bits_bytes - static_lds) / slice);
    }
    // per-wave LDS: the tile slice, plus the deferred-work block when walks
    // are deferred (line-of-sight flags cached and the bitset read from L2)
    // With the tiles split by XCD, a scenario's halos are staged through one
    // L2: in batched launches stage them with cacheable loads (the barrier
    // then acquires, which also drops the read-only lines).  A launch of one
    // scenario stages device-coherently and needs no acquire (4096²: 72.9 ->
    // 71.8 ms; the 128-scenario ensemble the other way, 98.7 -> 101.7 ms).
    a.cached_stage = (L.q.nsub > 1 && !a.lds_bits && L.nscen > 1) ? 1 : 0;
    const size_t wave_lds = slice + ((L.cache && !a.lds_bits) ? sizeof(DeferLds) : 0);
    const int max_waves = (a.lds_bits ? persistent_threads<true>() : persistent_threads<false>()) / kWave;
    if (tn.waves_per_group) {
      // with the bitset in LDS, at most the slices that fit beside it
      waves = a.lds_bits ? std::min(waves, static_cast<int>(tn.waves_per_group)) : static_cast<int>(tn.waves_per_group);
    }
    if (waves > max_waves) waves = max_waves;
    const size_t shmem = waves * wave_lds + (a.lds_bits ? bits_bytes : 0);
    const int block = waves * kWave;
    void (*kern)(Args) = persistent_kernel(tw, L.cache != 0, a.lds_bits != 0);
    int per_cu = 0;
    const Occupancy *known = nullptr;
    if (pd)
      for (const Occupancy &o : pd->occ)
        if (o.kern == reinterpret_cast<const void *>(kern) && o.block == block && o.shmem == shmem) known = &o;
    if (known) {
      per_cu = known->per_cu;
    } else {
      PG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, shmem));
      if (pd) pd->occ.push_back(Occupancy{reinterpret_cast<const void *>(kern), block, shmem, per_cu});
    }
    if (per_cu < 1) return set_error(PG_ERR_INVALID, "persistent kernel does not fit on a CU");
    unsigned groups = cu * static_cast<unsigned>(per_cu);
    if (groups > max_groups) groups = static_cast<unsigned>(max_groups);
    if (groups > 64u * kWave) groups = 64u * kWave;  // a barrier poller reads <= 64 words per lane
    if (tn.workgroups && tn.workgroups < groups) groups = tn.workgroups;
    // workgroup w draws from sub-list w % 8: a whole number per XCD
    if (L.q.nsub > 1 && groups >= kXcds) groups -= groups % kXcds;
    launched_groups = groups;
    if (use_graph) {
      LaunchGraph *lg = nullptr;
      st = launch_graph_for(pd, kern, groups, static_cast<unsigned>(block), shmem, a, max_groups, &lg);
      if (st != PG_OK) return st;
      hipKernelNodeParams pk{};
      void *kargs[] = {&a};
      pk.func = reinterpret_cast<void *>(kern);
      pk.gridDim = dim3(groups);
      pk.blockDim = dim3(static_cast<unsigned>(block));
      pk.sharedMemBytes = static_cast<unsigned>(shmem);
      pk.kernelParams = kargs;
      PG_HIP(hipGraphExecKernelNodeSetParams(lg->exec, lg->knode, &pk));
      const hipError_t ge = hipGraphLaunch(lg->exec, stream);
      if (ge != hipErrorIllegalState && ge != hipErrorCooperativeLaunchTooLarge) {
        PG_HIP(ge);
        st = finish_launch(false);  // the graph copies the control block itself
        if (st != PG_OK) return st;
        std::memcpy(book, &lg->host->book, sizeof(book));
        unsigned long long counters[16];
        std::memcpy(counters, lg->host->stats, sizeof(counters));
        float gms = 0.0f;
        PG_HIP(hipEventElapsedTime(&gms, lg->ev0, lg->ev1));
        return finish_stats(dev, book, counters, gms, launched_groups, a, prof_rows, stats, res);
      }
      // Cooperative launch refused (as below): per-sweep launches instead,
      // after the clear the graph would have made.
      (void)hipGetLastError();
      hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, stream, reinterpret_cast<uint32_t *>(a.ctl),
                         sizeof(Control) / 4, 0u);
      PG_HIP(hipGetLastError());
      st = run_per_sweep();
      if (st != PG_OK) return st;
    } else {
      void *kargs[] = {&a};
      const hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kern), dim3(groups), dim3(block),
                                                       kargs, static_cast<unsigned>(shmem), stream);
      if (le == hipErrorIllegalState || le == hipErrorCooperativeLaunchTooLarge) {
        // The device refuses a cooperative launch to this process while
        // another process holds its cooperative-launch resources (or when
        // the grid cannot be co-resident): the same sweeps, one launch each,
        // give the same results (the schedule does not depend on the launch
        // form).
        (void)hipGetLastError();
        st = run_per_sweep();
        if (st != PG_OK) return st;
      } else {
        PG_HIP(le);
        if (evr1) PG_HIP(hipEventRecord(evr1, stream));
      }
    }
  }
  st = finish_launch(true);
  if (st != PG_OK) return st;
  std::memcpy(book, hc.book.v, sizeof(book));
  float rms = 0.0f;
  if (evr0) PG_HIP(hipEventElapsedTime(&rms, evr0, evr1));
  return finish_stats(dev, book, hc.stats, rms, launched_groups, a, prof_rows, stats, res);
}

// Launch outcome and counters into res / stats.
static pg_status finish_stats(const DeviceInfo *dev, const unsigned *book, const unsigned long long *counters, float rms,
                       unsigned launched_groups, const Args &a, unsigned prof_rows, pg_stats *stats,
                       LaunchResult *res) {
  if (book[3] & 1u) return set_error(PG_ERR_TIMEOUT, "a grid barrier wait exceeded its spin limit");
  res->last_sweep = static_cast<int>(book[4]);
  res->finished = book[1] ? 1 : 0;
  if (stats) {
    int clock_khz = 0;
    if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev->device) != hipSuccess ||
        clock_khz <= 0)
      clock_khz = 100000;  // 100 MHz constant clock
    const double groups_d = launched_groups ? static_cast<double>(launched_groups) : 1.0;
    stats->relaxations += counters[0];
    stats->improvements += counters[1];
    stats->dda_steps += counters[2];
    stats->tiles_visited += counters[3];
    stats->tile_marks += counters[7];
    stats->idle_tiles += counters[12];
    stats->idle_far_tiles += counters[13];
    for (int q = 0; q < 4; ++q) stats->phase_total_ms[q] += static_cast<double>(counters[8 + q]) / clock_khz;
    stats->relax_ms += rms;
    stats->work_ms += static_cast<double>(counters[4]) / clock_khz / groups_d;
    stats->wait_ms += static_cast<double>(counters[5]) / clock_khz / groups_d;
    stats->crit_ms += static_cast<double>(counters[6]) / clock_khz;
    stats->workgroups = launched_groups;
    stats->lds_bitset = static_cast<uint32_t>(a.lds_bits);
    if (prof_rows) {
      std::vector<unsigned long long> prof(static_cast<size_t>(prof_rows) * kProfCols);
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
  return PG_OK;
}

// Where one propagation call keeps its buffers: the scenario rasters, tile
// queues and descriptors (`call`, `call_bytes` long) and the launch's control
// block (`launch`, launch_ws_bytes(cu, false) long).  Null: the per-device
// workspaces, grown with hipMalloc and held for the call.
struct CallSpace {
  char *call;
  size_t call_bytes;
  char *launch;
};

// Bytes of the call workspace of one scenario on the raster of sc.
static size_t call_ws_bytes(const Scen &sc, size_t nsub) {
  const size_t W = static_cast<size_t>(sc.g.W), H = static_cast<size_t>(sc.g.H);
  const size_t ntiles = static_cast<size_t>(sc.ntx) * sc.nty;
  const size_t nwords = size_t{1} << sc.nwords_log2;
  size_t off = align_up(scen_buffer_bytes(W, H), 256);
  off = align_up(off + 2 * nwords * sizeof(uint32_t), 256);
  off = align_up(off + 4 * nsub * ntiles * sizeof(unsigned long long), 256);
  off = align_up(off + 512 * nsub, 256);
  off = align_up(off + sizeof(Scen), 256);
  off = align_up(off + sizeof(ScenOut), 256);
  return off;
}

// The one-scenario entry point: prepare, run to convergence (or the cap),
// finish.  Buffers come from `space`, or from a per-device workspace held for
// the call when `space` is null (then the call is also timed with HIP events).
static pg_status propagate_call(const DeviceInfo *dev, const Settings *settings, const float *speed,
                                const float *az, const float *ecc, float *time, uint16_t *refs_x, uint16_t *refs_y,
                                uint32_t *boundaries, const pg_tuning *tuning, pg_stats *stats, hipStream_t stream,
                                const CallSpace *space) {
  pg_tuning tn{};
  if (tuning) tn = *tuning;
  if ((tn.tile_w || tn.tile_h) && !tile_shape_ok(tn.tile_w, tn.tile_h))
    return set_error(PG_ERR_INVALID, "tile shape must be 8x8, 16x4 or 32x2");
  Scen sc;
  pg_status st = scen_geometry(settings, &sc, tn.tile_w ? static_cast<int>(tn.tile_w) : kTile);
  if (st != PG_OK) return st;
  if (!speed || !az || !ecc || !time || !refs_x || !refs_y || !boundaries)
    return set_error(PG_ERR_INVALID, "null raster pointer");
  if (tn.max_iterations) sc.max_iters = tn.max_iterations;
  if (sc.max_iters > 0x3fffffffu) return set_error(PG_ERR_INVALID, "max_iterations too large");

  const size_t W = static_cast<size_t>(sc.g.W), H = static_cast<size_t>(sc.g.H);
  const uint32_t nwords = 1u << sc.nwords_log2;
  const size_t ntiles = static_cast<size_t>(sc.ntx) * sc.nty;
  const size_t o_scen = 0;
  const size_t o_mark = align_up(o_scen + scen_buffer_bytes(W, H), 256);
  const size_t nsub = kXcds;  // tile lists split by XCD (see TileQueue)
  const size_t o_list = align_up(o_mark + 2 * nwords * sizeof(uint32_t), 256);
  const size_t o_count = align_up(o_list + 4 * nsub * ntiles * sizeof(unsigned long long), 256);
  const size_t o_desc = align_up(o_count + 512 * nsub, 256);
  const size_t o_out = align_up(o_desc + sizeof(Scen), 256);
  const size_t need = call_ws_bytes(sc, nsub);
  char *base = nullptr;
  std::unique_lock<std::mutex> lk;
  HostReturn *hr = nullptr;  // pinned descriptor / outcome staging (library workspace only)
  if (space) {
    if (need > space->call_bytes) return set_error(PG_ERR_INVALID, "raster too large for the workspace");
    base = space->call;
  } else {
    PerDevice *pd = per_device(dev->device);
    lk = std::unique_lock<std::mutex>(pd->call_mu);
    st = ensure(&pd->call, need);
    if (st != PG_OK) return st;
    base = static_cast<char *>(pd->call.base);
    hr = pd->ret;
  }
  const ScenBuffers b = scen_carve(base + o_scen, W, H);
  sc.coef = b.coef;
  sc.st[0] = b.st0;
  sc.st[1] = b.st1;
  sc.stamp = b.stamp;
  sc.box = b.box;
  sc.blk = b.blk;
  sc.isref = b.isref;
  sc.bits = boundaries;
  sc.k0 = 0;
  TileQueue q;
  sc.mark[0] = reinterpret_cast<uint32_t *>(base + o_mark);
  sc.mark[1] = sc.mark[0] + nwords;
  q.list = reinterpret_cast<unsigned long long *>(base + o_list);
  q.count = reinterpret_cast<uint32_t *>(base + o_count);
  q.cap = ntiles;
  q.nsub = static_cast<uint32_t>(nsub);
  Scen *d_scen = reinterpret_cast<Scen *>(base + o_desc);
  ScenOut *d_out = reinterpret_cast<ScenOut *>(base + o_out);

  const bool timed = space == nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (timed) {
    PerDevice *pd = per_device(dev->device);  // call_mu is held
    st = device_event(pd, 0, &ev0);
    if (st == PG_OK) st = device_event(pd, 1, &ev1);
    if (st != PG_OK) return st;
  }
  if (ev0) PG_HIP(hipEventRecord(ev0, stream));
  st = scen_prepare(dev, sc, 0, d_out, q, speed, az, ecc, time, refs_x, refs_y, boundaries, false, stream, true);
  if (st != PG_OK) return st;
  if (hr) hr->desc = sc;  // free: the previous call on this device has synchronised
  PG_HIP(hipMemcpyAsync(d_scen, hr ? &hr->desc : &sc, sizeof(Scen), hipMemcpyHostToDevice, stream));
  // Post-pass and outcome right behind the relaxation, so the call waits on
  // its stream once: the post kernel reads the outcome on the device.
  struct After {
    const Scen *sc;
    ScenOut *d_out, *h_out;
    const float *speed;
    float *time;
    uint16_t *rx, *ry;
    hipEvent_t ev1;
  };
  ScenOut o{};
  After aft{&sc, d_out, hr ? &hr->out : &o, speed, time, refs_x, refs_y, ev1};

  Launch L{};
  L.scen = d_scen;
  L.out = d_out;
  L.nscen = 1;
  L.q = q;
  L.k_begin = 0;
  L.k_end = static_cast<int>(sc.max_iters) + 1;  // the last sweep only detects a cap
  L.cache = sc.cache;
  L.tw = sc.tw;
  L.lds_bits_src = boundaries;
  L.lds_bits_words = (W * H + 31) / 32;
  L.ws = space ? space->launch : nullptr;
  L.untimed = timed ? 0 : 1;
  L.after = [](void *p, hipStream_t s) -> pg_status {
    const After &f = *static_cast<const After *>(p);
    const pg_status fs = scen_finish(*f.sc, 0, f.speed, f.time, f.rx, f.ry, s, f.d_out);
    if (fs != PG_OK) return fs;
    PG_HIP(hipMemcpyAsync(f.h_out, f.d_out, sizeof(ScenOut), hipMemcpyDeviceToHost, s));
    if (f.ev1) PG_HIP(hipEventRecord(f.ev1, s));
    return PG_OK;
  };
  L.after_ctx = &aft;
  pg_stats local{};
  LaunchResult res{};
  st = run_sweeps(dev, L, &tn, &local, &res, stream);
  if (st != PG_OK) {
    // Leave no work of this call queued on the workspace.
    (void)hipStreamSynchronize(stream);
    return st;
  }
  if (hr) o = hr->out;  // run_sweeps has synchronised the stream
  if (o.err & 2u) return set_error(PG_ERR_INVALID, "a burned cell (time < FLT_MAX) has a reference outside the raster");
  float ms = 0.0f;
  if (ev0) PG_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  if (stats) {
    *stats = local;
    stats->iterations = o.capped ? sc.max_iters : static_cast<uint64_t>(o.last_change - sc.k0 + 2);
    stats->elapsed_ms = ms;
  }
  if (o.capped) return set_error(PG_ERR_NOT_CONVERGED, "iteration cap reached before convergence");
  return PG_OK;
}

pg_status run_propagate(const DeviceInfo *dev, const Settings *settings, const float *speed, const float *az,
                        const float *ecc, float *time, uint16_t *refs_x, uint16_t *refs_y, uint32_t *boundaries,
                        const pg_tuning *tuning, pg_stats *stats, hipStream_t stream) {
  return propagate_call(dev, settings, speed, az, ecc, time, refs_x, refs_y, boundaries, tuning, stats, stream,
                        nullptr);
}

}  // namespace pg

// ------------------------------------------- allocation-free entry point
//
// pg_propagate_noalloc runs pg_propagate's kernels with every buffer beyond
// the caller's inputs and outputs carved from `scratch`, a device array of
// this code object: each device the library runs on holds its own copy
// (kNoAllocScratchBytes) until the process ends.  Its address on a device is
// taken once, with that device current, by hipGetSymbolAddress; host code
// passes only that address to kernels and copies (the array's name is a
// host-side stand-in, not a device address).  Calls on one device are
// serialised by a lock and each one synchronises its stream before it
// returns; a call clears what must start at zero and leaves its contents
// behind for the next.
static __device__ __align__(256) unsigned char scratch[pg::kNoAllocScratchBytes];

namespace pg {
namespace {

constexpr int kNoAllocDevices = 64;

struct NoAllocDevice {
  std::mutex mu;        // held for a whole call
  bool ready = false;   // set-up done
  char *ptr = nullptr;  // this device's `scratch`
  DeviceInfo info{};    // properties (no catalog: the propagation does not burn)
};

NoAllocDevice g_noalloc[kNoAllocDevices];

// Set-up on the current device (under d.mu): device properties and the
// address of this device's `scratch`.
pg_status noalloc_setup(NoAllocDevice &d, int dev) {
  hipDeviceProp_t prop;
  PG_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(PG_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                           "; this library is built for gfx950 only");
  int lds = 0;
  PG_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    return set_error(PG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  d.info.device = dev;
  d.info.cu_count = prop.multiProcessorCount;
  d.info.lds_per_block = static_cast<size_t>(lds);
  d.info.catalog = nullptr;
  d.ptr = static_cast<char *>(p);
  d.ready = true;
  return PG_OK;
}

}  // namespace
}  // namespace pg

extern "C" pg_status pg_propagate_noalloc(const Settings *settings, const float *speed_max,
                                          const float *azimuth_max, const float *eccentricity, float *time,
                                          uint16_t *refs_x, uint16_t *refs_y, uint32_t *boundaries,
                                          const pg_tuning *tuning, pg_stats *stats, void *stream) {
  using namespace pg;
  if (!settings) return set_error(PG_ERR_INVALID, "settings is NULL");
  const uint64_t cells = static_cast<uint64_t>(settings->geo_ref.width) * settings->geo_ref.height;
  if (cells > PG_NOALLOC_MAX_CELLS)
    return set_error(PG_ERR_INVALID, "pg_propagate_noalloc: raster of " + std::to_string(cells) +
                                         " cells exceeds its maximum of " + std::to_string(PG_NOALLOC_MAX_CELLS));
  if (tuning && (tuning->mode == 2 || tuning->profile))
    return set_error(PG_ERR_INVALID, "pg_propagate_noalloc: mode 2 (graph) and profiling allocate; use mode 0 or 1");
  int dev = 0;
  PG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return set_error(PG_ERR_NO_DEVICE, "device ordinal out of range");
  NoAllocDevice &d = g_noalloc[dev];
  std::lock_guard<std::mutex> lk(d.mu);
  if (!d.ready) {
    const pg_status s = noalloc_setup(d, dev);
    if (s != PG_OK) return s;
  }
  const size_t launch_bytes = launch_ws_bytes(static_cast<unsigned>(d.info.cu_count), false);
  if (launch_bytes >= kNoAllocScratchBytes) return set_error(PG_ERR_INVALID, "device too large for the scratch array");
  const CallSpace space{d.ptr + launch_bytes, kNoAllocScratchBytes - launch_bytes, d.ptr};
  return propagate_call(&d.info, settings, speed_max, azimuth_max, eccentricity, time, refs_x, refs_y, boundaries,
                        tuning, stats, static_cast<hipStream_t>(stream), &space);
}
