/*
 * Oracle (test infrastructure only): serial transcription of propag25's
 * arrival-time propagation.  See oracle.h for who may use this.
 *
 *  - update rule        crates/propag/src/propag.cu:124-190
 *  - commit             crates/propag/src/propag.cu:192-207
 *  - load_point         crates/propag/src/propag.cu:209-229 (a burned cell's
 *                       reference time is re-read from the time raster)
 *  - time_from          crates/propag/src/propag.cu:269-294
 *  - boundaries         crates/propag/src/propag.cu:296-302,345-375
 *  - post_propagate     crates/propag/src/propag.cu:110-117
 *  - DDA                crates/propag/src/propag.h:89-135 (f32 tmax/delta)
 *  - schedule           every sweep reads the previous sweep's state and
 *                       commits afterwards, which is what one cooperative
 *                       launch does between its two grid syncs
 *                       (propag.cu:62-109).  Sweeps repeat until one commits
 *                       nothing (the host loop of propag.rs:502-536 with one
 *                       super-tile).
 *
 * time_from, formula 0: t = t_from + dist * (1 - e*cos(az - az_max)) * R with
 *   e*cos(az - az_max) = (A*vy + B*vx) / r, A = e*cos(az_max),
 *   evaluated for square cells as t_from + ((r - (A*vy + B*vx)) * cell) * R
 *   (dist = cell * r), otherwise as t_from + (dist * (1 - (A*vy+B*vx)/r)) * R,
 *   B = e*sin(az_max), R = 1/(speed_max*(1-e)), (vx, vy) the cell offset in
 *   the azimuth frame, r its length.  This is the reference's
 *   speed_max*(1-e)/(1-e*cos|az-az_max|) with cos(atan2(..) - az_max)
 *   expanded; A, B, R are formed once per cell in double with oracle_sincos
 *   and rounded to float.  All steps are IEEE +,-,*,/,sqrt so the HIP kernel
 *   reproduces them bit for bit.
 * time_from, formula 1: the reference's literal arithmetic (atan2f, cosf,
 *   reciprocal, sqrt) with libm and IEEE division.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NO_REF 0xFFFFu

/* Planted mistakes for the tests that check the known answers can fail
 * (tests/test_oracle_kat.py builds one library per value; the default build
 * has none):
 *   1 the blocked-line fallback is ignored (ct kept despite a boundary hit)
 *   2 the line-of-sight walk starts at the cell instead of the neighbour
 *   3 the boundary predicate takes the faster side (> for < on speed)
 *   4 similar_fires compares azimuths with wrap-around
 *   5 neighbours visited i outer, j inner
 *   6 ties keep the last candidate (<= for <)
 *   7 a neighbour referencing the cell is not skipped
 *   8 a neighbour burning exactly at max_time spreads (<= max_time)
 *   9 the post-pass drops arrivals equal to max_time (>= for >) */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

/* ------------------------------------------------------------- sincos */

void oracle_sincos(double x, double *s, double *c) {
  /* pi/2 split so k*P1 is exact for |k| < 2^20 */
  const double TWO_OVER_PI = 0.63661977236758134308;
  const double P1 = 1.5707963267341256;
  const double P2 = 6.0771005065061922e-11;
  double k = rint(x * TWO_OVER_PI);
  double r = (x - k * P1) - k * P2;
  double r2 = r * r;
  /* Horner over r^2 with 1/n! (n odd up to 19 for sin, even up to 20 for cos) */
  double ps = -1.0 / 121645100408832000.0; /* -1/19! */
  ps = ps * r2 + 1.0 / 355687428096000.0;  /* +1/17! */
  ps = ps * r2 - 1.0 / 1307674368000.0;    /* -1/15! */
  ps = ps * r2 + 1.0 / 6227020800.0;       /* +1/13! */
  ps = ps * r2 - 1.0 / 39916800.0;         /* -1/11! */
  ps = ps * r2 + 1.0 / 362880.0;           /* +1/9!  */
  ps = ps * r2 - 1.0 / 5040.0;             /* -1/7!  */
  ps = ps * r2 + 1.0 / 120.0;              /* +1/5!  */
  ps = ps * r2 - 1.0 / 6.0;                /* -1/3!  */
  ps = ps * r2;
  double sn = r + r * ps;
  double pc = 1.0 / 2432902008176640000.0; /* +1/20! */
  pc = pc * r2 - 1.0 / 6402373705728000.0; /* -1/18! */
  pc = pc * r2 + 1.0 / 20922789888000.0;   /* +1/16! */
  pc = pc * r2 - 1.0 / 87178291200.0;      /* -1/14! */
  pc = pc * r2 + 1.0 / 479001600.0;        /* +1/12! */
  pc = pc * r2 - 1.0 / 3628800.0;          /* -1/10! */
  pc = pc * r2 + 1.0 / 40320.0;            /* +1/8!  */
  pc = pc * r2 - 1.0 / 720.0;              /* -1/6!  */
  pc = pc * r2 + 1.0 / 24.0;               /* +1/4!  */
  pc = pc * r2 - 0.5;                      /* -1/2!  */
  double cs = 1.0 + r2 * pc;
  long q = ((long)k) & 3;
  switch (q) {
  case 0: *s = sn; *c = cs; break;
  case 1: *s = cs; *c = -sn; break;
  case 2: *s = -sn; *c = -cs; break;
  default: *s = -cs; *c = sn; break;
  }
}

/* ------------------------------------------------------------- helpers */

typedef struct {
  float A, B, R; /* e*cos(az_max), e*sin(az_max), 1/(speed_max*(1-e)) */
} coeffs;

static coeffs fire_coeffs(float smax, float az, float ecc) {
  coeffs k;
  double s, c;
  oracle_sincos((double)az, &s, &c);
  k.A = (float)((double)ecc * c);
  k.B = (float)((double)ecc * s);
  k.R = (float)(1.0 / ((double)smax * (1.0 - (double)ecc)));
  return k;
}

static inline float time_from0(float from_t, int fx, int fy, int tx, int ty,
                               coeffs k, float gdx, float gdy, int north_up) {
  int ix = tx - fx;
  int iy = ty - fy;
  float ux = (float)ix;
  float uy = (float)(north_up ? -iy : iy);
  float r = sqrtf(ux * ux + uy * uy);
  float proj = k.A * uy + k.B * ux; /* e * r * cos(az - az_max) */
  if (fabsf(gdx) == fabsf(gdy)) /* square cells: |d| = cell * r */
    return from_t + ((r - proj) * fabsf(gdx)) * k.R;
  float ex = ux * gdx;
  float ey = (float)iy * gdy;
  float dist = sqrtf(ex * ex + ey * ey);
  return from_t + (dist * (1.0f - proj / r)) * k.R;
}

static inline float time_from1(float from_t, int fx, int fy, int tx, int ty,
                               float smax, float azm, float ecc, float gdx,
                               float gdy, int north_up) {
  float az;
  if (north_up)
    az = atan2f((float)(tx - fx), (float)(fy - ty));
  else
    az = atan2f((float)(tx - fx), (float)(ty - fy));
  float angle = fabsf(az - azm);
  float denom = (float)(1.0 - (double)(ecc * cosf(angle)));
  float factor = (float)((1.0 - (double)ecc) * (double)(1.0f / denom));
  float speed = smax * factor;
  float ddx = (float)(fx - tx) * gdx;
  float ddy = (float)(fy - ty) * gdy;
  float distance = sqrtf(ddx * ddx + ddy * ddy);
  return from_t + distance * (1.0f / speed);
}

float oracle_time_from(const oracle_grid *g, float from_time, int fx, int fy,
                       int tx, int ty, float speed_max, float azimuth_max,
                       float eccentricity, int formula) {
  int north_up = g->dy < 0.0f;
  if (formula == 1)
    return time_from1(from_time, fx, fy, tx, ty, speed_max, azimuth_max,
                      eccentricity, g->dx, g->dy, north_up);
  return time_from0(from_time, fx, fy, tx, ty,
                    fire_coeffs(speed_max, azimuth_max, eccentricity), g->dx,
                    g->dy, north_up);
}

/* ---------------------------------------------------------------- DDA */

typedef struct {
  int tx, ty, sx, sy, cx, cy;
  float tmx, tmy, dlx, dly;
} dda_t;

static inline int signum_i(int v) { return (0 < v) - (v < 0); }

static inline void dda_init(dda_t *d, int fx, int fy, int tx, int ty) {
  d->tx = tx;
  d->ty = ty;
  d->sx = signum_i(tx - fx);
  d->sy = signum_i(ty - fy);
  d->cx = fx;
  d->cy = fy;
  /* 1.0 / float(abs(..)) is a double division narrowed to float */
  d->tmx = (float)(1.0 / (double)(float)abs(tx - fx));
  d->tmy = (float)(1.0 / (double)(float)abs(ty - fy));
  d->dlx = d->tmx;
  d->dly = d->tmy;
}

static inline int dda_next(dda_t *d, int *rx, int *ry) {
  *rx = d->cx;
  *ry = d->cy;
  if (d->cx == d->tx && d->cy == d->ty) {
    d->cx += d->sx;
    d->cy += d->sy;
    return 1;
  }
  int vx = d->sx > 0 ? d->cx <= d->tx : d->cx >= d->tx;
  int vy = d->sy > 0 ? d->cy <= d->ty : d->cy >= d->ty;
  if (!(vx && vy)) return 0;
  if ((double)fabsf(d->tmx - d->tmy) < 1e-9) {
    d->cx += d->sx;
    d->cy += d->sy;
    d->tmx += d->dlx;
    d->tmy += d->dly;
  } else if (d->tmx < d->tmy) {
    d->cx += d->sx;
    d->tmx += d->dlx;
  } else {
    d->cy += d->sy;
    d->tmy += d->dly;
  }
  return 1;
}

int oracle_dda(int fx, int fy, int tx, int ty, int32_t *out_xy, int max_out) {
  dda_t d;
  int x, y, n = 0;
  if (fx == tx && fy == ty) { /* the device walk never runs with from == to */
    if (max_out > 0) { out_xy[0] = fx; out_xy[1] = fy; }
    return 1;
  }
  dda_init(&d, fx, fy, tx, ty);
  while (dda_next(&d, &x, &y)) {
    if (n < max_out) {
      out_xy[2 * n] = x;
      out_xy[2 * n + 1] = y;
    }
    n++;
    if (n > 1 << 24) break; /* guard; a walk is at most |dx|+|dy|+1 long */
  }
  return n;
}

/* ---------------------------------------------------------- boundaries */

static inline int similar_fires(float s1, float a1, float e1, float s2,
                                float a2, float e2) {
  double da = (double)fabsf(a1 - a2);
#if ORACLE_MUTANT == 4
  if (da > 3.14159265358979323846) da = 2 * 3.14159265358979323846 - da;
#endif
  return (double)fabsf(s1 - s2) < 0.1 &&
         da < (5.0 * (2 * 3.14159265358979323846) / 360.0) &&
         (double)fabsf(e1 - e2) < 0.05;
}

void oracle_boundaries(const oracle_grid *g, const float *speed_max,
                       const float *azimuth_max, const float *eccentricity,
                       uint32_t *boundaries) {
  const int W = (int)g->width, H = (int)g->height;
  for (int y = 0; y < H; y++) {
    for (int x = 0; x < W; x++) {
      size_t c = (size_t)x + (size_t)y * W;
      int changed = 0;
      for (int i = -1; i < 2; i++) {
        for (int j = -1; j < 2; j++) {
          if (i == 0 && j == 0) continue;
          int nx = x + i, ny = y + j;
          if (!(nx >= 0 && ny >= 0 && nx < W && ny < H)) continue;
          size_t n = (size_t)nx + (size_t)ny * W;
          changed |= !similar_fires(speed_max[n], azimuth_max[n],
                                    eccentricity[n], speed_max[c],
                                    azimuth_max[c], eccentricity[c]) &&
#if ORACLE_MUTANT == 3
                     speed_max[c] > speed_max[n];
#else
                     speed_max[c] < speed_max[n];
#endif
        }
      }
      if (changed) boundaries[c >> 5] |= 1u << (c & 31);
    }
  }
}

static inline int is_boundary(const uint32_t *b, size_t ix) {
  return (b[ix >> 5] >> (ix & 31)) & 1u;
}

/* --------------------------------------------------------- propagation */

typedef struct {
  uint32_t idx;
  float t;
  uint16_t rx, ry;
} change_t;

typedef struct {
  change_t *v;
  size_t n, cap;
} change_list;

static void push_change(change_list *l, change_t c) {
  if (l->n == l->cap) {
    l->cap = l->cap ? l->cap * 2 : 1024;
    l->v = (change_t *)realloc(l->v, l->cap * sizeof(change_t));
  }
  l->v[l->n++] = c;
}

typedef struct {
  const oracle_grid *g;
  const float *smax, *azm, *ecc;
  const coeffs *k;
  const float *time;
  const uint16_t *rx, *ry;
  const uint32_t *bnd;
  int formula, north_up;
} ctx_t;

/* propag.cu:124-190 for the cell (x, y); returns 1 and the new state if the
 * best candidate is strictly earlier than the cell's current time. */
static int relax_cell(const ctx_t *C, int x, int y, float *out_t,
                      uint16_t *out_rx, uint16_t *out_ry, int64_t *dda_steps) {
  const int W = (int)C->g->width, H = (int)C->g->height;
  const float max_time = C->g->max_time;
  size_t me = (size_t)x + (size_t)y * W;
  float best_t = C->time[me];
  uint16_t best_rx = NO_REF, best_ry = NO_REF;
  for (int j = -1; j < 2; j++) {
    for (int i = -1; i < 2; i++) {
#if ORACLE_MUTANT == 5
      { int t_ = i; i = j; j = t_; }
#endif
      int nx = x + i, ny = y + j;
#if ORACLE_MUTANT == 5
      { int t_ = i; i = j; j = t_; }
#endif
      if (nx == x && ny == y) continue;
      if (!(nx >= 0 && ny >= 0 && nx < W && ny < H)) continue;
      size_t n = (size_t)nx + (size_t)ny * W;
      float n_t = C->time[n];
      float n_s = C->smax[n];
#if ORACLE_MUTANT == 8
      if (!(n_t <= max_time && !(n_s == 0.0f))) continue;
#else
      if (!(n_t < max_time && !(n_s == 0.0f))) continue;
#endif
      int rfx = C->rx[n], rfy = C->ry[n];
#if ORACLE_MUTANT != 7
      if (rfx == x && rfy == y) continue;
#endif
      float ref_t = C->time[(size_t)rfx + (size_t)rfy * W];
      float nt, ct;
      if (C->formula == 1) {
        nt = time_from1(n_t, nx, ny, x, y, n_s, C->azm[n], C->ecc[n],
                        C->g->dx, C->g->dy, C->north_up);
        ct = time_from1(ref_t, rfx, rfy, x, y, n_s, C->azm[n], C->ecc[n],
                        C->g->dx, C->g->dy, C->north_up);
      } else {
        nt = time_from0(n_t, nx, ny, x, y, C->k[n], C->g->dx, C->g->dy,
                        C->north_up);
        ct = time_from0(ref_t, rfx, rfy, x, y, C->k[n], C->g->dx, C->g->dy,
                        C->north_up);
      }
#if ORACLE_MUTANT == 6
#define ORACLE_LESS(a, b) ((a) <= (b))
#else
#define ORACLE_LESS(a, b) ((a) < (b))
#endif
      if (ct < nt && ORACLE_LESS(ct, best_t)) {
        int cand_rx = rfx, cand_ry = rfy;
        float cand_t = ct;
        dda_t d;
        int px, py;
#if ORACLE_MUTANT == 2
        dda_init(&d, x, y, rfx, rfy);
#else
        dda_init(&d, nx, ny, rfx, rfy);
#endif
        while (dda_next(&d, &px, &py)) {
          (*dda_steps)++;
          if (is_boundary(C->bnd, (size_t)px + (size_t)py * W)) {
#if ORACLE_MUTANT != 1
            cand_rx = nx;
            cand_ry = ny;
            cand_t = nt;
#endif
            break;
          }
        }
        if (ORACLE_LESS(cand_t, best_t)) {
          best_t = cand_t;
          best_rx = (uint16_t)cand_rx;
          best_ry = (uint16_t)cand_ry;
        }
      } else if (ORACLE_LESS(nt, best_t)) {
        best_t = nt;
        best_rx = (uint16_t)nx;
        best_ry = (uint16_t)ny;
      }
    }
  }
  if (best_t < C->time[me]) {
    *out_t = best_t;
    *out_rx = best_rx;
    *out_ry = best_ry;
    return 1;
  }
  return 0;
}

int oracle_propagate(const oracle_grid *g, const float *speed_max,
                     const float *azimuth_max, const float *eccentricity,
                     float *time, uint16_t *refs_x, uint16_t *refs_y,
                     uint32_t *boundaries, int formula, int64_t max_sweeps,
                     int nthreads, oracle_stats *stats) {
  const int W = (int)g->width, H = (int)g->height;
  const size_t N = (size_t)W * H;
  oracle_stats st = {0, 0, 0, 0};
  coeffs *k = (coeffs *)malloc(N * sizeof(coeffs));
  for (size_t i = 0; i < N; i++)
    k[i] = fire_coeffs(speed_max[i], azimuth_max[i], eccentricity[i]);

  /* pre_propagate */
  memset(boundaries, 0, ((N + 31) / 32) * sizeof(uint32_t));
  oracle_boundaries(g, speed_max, azimuth_max, eccentricity, boundaries);

  ctx_t C = {g, speed_max, azimuth_max, eccentricity, k, time,
             refs_x, refs_y, boundaries, formula, g->dy < 0.0f};

  /* Cells whose eight neighbours are all unburned cannot change in a sweep,
   * so each sweep visits the burned bounding box grown by one cell. */
  int bx0 = W, by0 = H, bx1 = -1, by1 = -1;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++)
      if (time[(size_t)x + (size_t)y * W] < FLT_MAX) {
        if (x < bx0) bx0 = x;
        if (x > bx1) bx1 = x;
        if (y < by0) by0 = y;
        if (y > by1) by1 = y;
      }
  if (nthreads < 1) nthreads = 1;
  change_list *lists = (change_list *)calloc((size_t)nthreads, sizeof(change_list));
  int64_t *dda = (int64_t *)calloc((size_t)nthreads, sizeof(int64_t));
  int hit_cap = 0;
  while (bx1 >= 0) {
    if (max_sweeps > 0 && st.sweeps >= max_sweeps) {
      hit_cap = 1;
      break;
    }
    int x0 = bx0 > 0 ? bx0 - 1 : 0, x1 = bx1 < W - 1 ? bx1 + 1 : W - 1;
    int y0 = by0 > 0 ? by0 - 1 : 0, y1 = by1 < H - 1 ? by1 + 1 : H - 1;
    for (int t = 0; t < nthreads; t++) lists[t].n = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int y = y0; y <= y1; y++) {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      for (int x = x0; x <= x1; x++) {
        change_t c;
        if (relax_cell(&C, x, y, &c.t, &c.rx, &c.ry, &dda[tid])) {
          c.idx = (uint32_t)((size_t)x + (size_t)y * W);
          push_change(&lists[tid], c);
        }
      }
    }
    st.sweeps++;
    st.relaxations += (int64_t)(x1 - x0 + 1) * (y1 - y0 + 1);
    size_t nchg = 0;
    for (int t = 0; t < nthreads; t++) {
      for (size_t i = 0; i < lists[t].n; i++) {
        change_t c = lists[t].v[i];
        time[c.idx] = c.t;
        refs_x[c.idx] = c.rx;
        refs_y[c.idx] = c.ry;
        int x = (int)(c.idx % (uint32_t)W), y = (int)(c.idx / (uint32_t)W);
        if (x < bx0) bx0 = x;
        if (x > bx1) bx1 = x;
        if (y < by0) by0 = y;
        if (y > by1) by1 = y;
      }
      nchg += lists[t].n;
    }
    st.improvements += (int64_t)nchg;
    if (nchg == 0) break;
  }
  for (int t = 0; t < nthreads; t++) {
    st.dda_steps += dda[t];
    free(lists[t].v);
  }
  free(lists);
  free(dda);

  /* post_propagate */
  for (size_t i = 0; i < N; i++)
#if ORACLE_MUTANT == 9
    if (speed_max[i] == 0.0f || time[i] >= g->max_time) time[i] = FLT_MAX;
#else
    if (speed_max[i] == 0.0f || time[i] > g->max_time) time[i] = FLT_MAX;
#endif
  free(k);
  if (stats) *stats = st;
  return hit_cap;
}
