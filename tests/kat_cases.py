"""Hand-derived known answers for Stage P's update rule, boundary predicate
and post-pass, on rasters of at most 6 x 4 cells.

Each case is worked out from the reference's rules, not from the oracle:
find_neighbor_which_reaches_first (crates/propag/src/propag.cu:124-190),
time_from (269-294), similar_fires / find_boundaries (296-302, 345-375),
post_propagate (110-117) and the DDA (crates/propag/src/propag.h:91-135).
Every case uses square 1 m cells, south-up, eccentricity 0 and speed 1 m/s
unless it says otherwise, so time_from(a -> b) is t_a + |b - a| exactly as
f32 arithmetic rounds it: an adjacent step adds 1 or sqrtf(2), a far
reference adds sqrtf(dx^2 + dy^2) (every such value below is an exact
integer or a correctly rounded square root).  Azimuths matter only to the
boundary predicate when the eccentricity is 0.

A case is a dict: raster size, fire rasters, initial state, max_time, the
number of Jacobi sweeps to run (0: to convergence), and the expectations:
`cells` maps (x, y) to (time, (ref_x, ref_y)) after the post-pass (time
FLT_MAX and ref 0xFFFF for a cell never reached), `bits` is the exact set of
boundary cells when given.  `catches` names the planted oracle mistakes
(ORACLE_MUTANT values of oracle/propag_oracle.c) that the case must fail.
"""
import math

import numpy as np

FLT_MAX = np.float32(np.finfo(np.float32).max)
NO_REF = (0xFFFF, 0xFFFF)
F = np.float32


def _fire(W, H, speed=1.0, az=0.0, ecc=0.0):
    return (np.full(W * H, F(speed), np.float32), np.full(W * H, F(az), np.float32),
            np.full(W * H, F(ecc), np.float32))


def _state(W, H, burning):
    """burning: {(x, y): (t, (rx, ry))}; everything else FLT_MAX / 0xFFFF."""
    t = np.full(W * H, FLT_MAX, np.float32)
    rx = np.full(W * H, 0xFFFF, np.uint16)
    ry = np.full(W * H, 0xFFFF, np.uint16)
    for (x, y), (tt, (ax, ay)) in burning.items():
        t[x + y * W] = F(tt)
        rx[x + y * W], ry[x + y * W] = ax, ay
    return t, rx, ry


def _case(name, W, H, fire, burning, max_time, sweeps, cells, bits=None, catches=(), doc=""):
    t, rx, ry = _state(W, H, burning)
    return {"name": name, "W": W, "H": H, "speed": fire[0], "az": fire[1], "ecc": fire[2],
            "time": t, "refs_x": rx, "refs_y": ry, "max_time": F(max_time), "sweeps": sweeps,
            "cells": cells, "bits": bits, "catches": tuple(catches), "doc": doc}


SQRT10 = F(math.sqrt(10.0))  # |(3,1) - (0,0)|, correctly rounded


def fallback_taken():
    """A neighbour's line to its far reference crosses a boundary cell.

    6 x 3 raster; B = (2,1) has speed 0.95 and azimuth 1 rad, every other
    cell speed 1, azimuth 0.  B is the only boundary: its neighbours are
    faster and dissimilar (|d az| = 1 rad > 5 deg), while no other cell has
    a faster dissimilar neighbour.  Initial state: R = (0,0) burning at 0
    (its own reference) and N = (3,1) at sqrtf(10) with reference R.
    One sweep.  The cell c = (4,1) has exactly one burning neighbour, N:
      nt = t_N + 1 = sqrtf(10) + 1 = 4.1622777
      ct = t_R + |c - R| = sqrtf(17) = 4.1231055 < nt,
    so the walk from N to R runs: the DDA from (3,1) to (0,0) visits
    (3,1), (2,1), (1,1), (0,0) (tmax_x = 1/3 < tmax_y = 1 twice, then the
    tie step), which holds B.  The candidate falls back to nt with N as
    reference: c = (sqrtf(10) + 1, (3,1)).  Keeping ct despite the hit
    gives (sqrtf(17), (0,0))."""
    W, H = 6, 3
    s, a, e = _fire(W, H)
    s[2 + 1 * W], a[2 + 1 * W] = F(0.95), F(1.0)
    return _case("fallback_taken", W, H, (s, a, e),
                 {(0, 0): (0.0, (0, 0)), (3, 1): (SQRT10, (0, 0))}, 100.0, 1,
                 {(4, 1): (F(SQRT10 + F(1.0)), (3, 1))}, bits={(2, 1)}, catches=(1,),
                 doc=fallback_taken.__doc__)


def line_clear():
    """The same set-up with the slow cell at B = (2,0): off the line from the
    neighbour to the reference, on the line from the cell to it.

    c = (4,0) has one burning neighbour, N = (3,1) (diagonal):
      nt = sqrtf(10) + sqrtf(2) = 4.5764914
      ct = |(4,0) - (0,0)| = 4 < nt,
    and the DDA from N = (3,1) to R = (0,0) visits (3,1), (2,1), (1,1),
    (0,0): no boundary, so c = (4, (0,0)).  Walking from the cell instead
    ((4,0), (3,0), (2,0), ...) meets B and gives (nt, (3,1))."""
    W, H = 6, 3
    s, a, e = _fire(W, H)
    s[2], a[2] = F(0.95), F(1.0)
    return _case("line_clear", W, H, (s, a, e),
                 {(0, 0): (0.0, (0, 0)), (3, 1): (SQRT10, (0, 0))}, 100.0, 1,
                 {(4, 0): (F(4.0), (0, 0))}, bits={(2, 0)}, catches=(2,), doc=line_clear.__doc__)


def step_field():
    """Two-value step field: speed 1 for x < 3, 2 for x >= 3 (6 x 4).

    A cell is a boundary iff a neighbour is dissimilar (|d speed| = 1 >= 0.1)
    and faster, i.e. exactly the slower side's column next to the step,
    x = 2, every row."""
    W, H = 6, 4
    s, a, e = _fire(W, H)
    for y in range(H):
        for x in range(3, W):
            s[x + y * W] = F(2.0)
    return _case("step_field", W, H, (s, a, e), {}, 100.0, 0, {},
                 bits={(2, y) for y in range(H)}, catches=(3,), doc=step_field.__doc__)


def _pair(name, f0, f1, bit0, catches=(), doc=""):
    """2 x 1 raster, cell 0 fire f0, cell 1 fire f1; cell 1 is never slower,
    so the only possible boundary is cell 0, iff the fires are dissimilar."""
    s = np.array([f0[0], f1[0]], np.float32)
    a = np.array([f0[1], f1[1]], np.float32)
    e = np.array([f0[2], f1[2]], np.float32)
    return _case(name, 2, 1, (s, a, e), {}, 100.0, 0, {}, bits={(0, 0)} if bit0 else set(),
                 catches=catches, doc=doc)


def _float_at_least(v):
    """The least float whose value is >= the double v."""
    f = F(v)
    return f if float(f) >= v else np.nextafter(f, F(np.inf))


def thresholds():
    """similar_fires thresholds (propag.cu:296-302): |d speed| < 0.1,
    |d az| < 5 deg = 5 * 2 pi / 360 rad, |d ecc| < 0.05, each difference
    taken in f32 and compared as a double, strictly.  For each quantity the
    least f32 difference at or above the threshold makes the pair
    dissimilar (cell 0, the slower, is a boundary) and the float below it
    keeps it similar.  Azimuths do not wrap: 359 deg against 1 deg differs
    by 358 deg, dissimilar."""
    out = []
    s_hi = F(1.1)                         # 1.1f - 1.0f = 0.10000002 >= 0.1
    assert float(s_hi - F(1.0)) >= 0.1
    s_lo = np.nextafter(s_hi, F(0.0))     # 1.0999999f - 1.0f = 0.0999999 < 0.1
    assert float(s_lo - F(1.0)) < 0.1
    out.append(_pair("speed_at_threshold", (1.0, 0.0, 0.0), (s_hi, 0.0, 0.0), True))
    out.append(_pair("speed_below_threshold", (1.0, 0.0, 0.0), (s_lo, 0.0, 0.0), False))
    deg5 = 5.0 * (2 * math.pi) / 360.0
    a_hi = _float_at_least(deg5)
    a_lo = np.nextafter(a_hi, F(0.0))
    out.append(_pair("azimuth_at_threshold", (1.0, 0.0, 0.0), (1.01, a_hi, 0.0), True))
    out.append(_pair("azimuth_below_threshold", (1.0, 0.0, 0.0), (1.01, a_lo, 0.0), False))
    e_hi = _float_at_least(0.05)
    e_lo = np.nextafter(e_hi, F(0.0))
    out.append(_pair("ecc_at_threshold", (1.0, 0.0, 0.0), (1.01, 0.0, e_hi), True))
    out.append(_pair("ecc_below_threshold", (1.0, 0.0, 0.0), (1.01, 0.0, e_lo), False))
    out.append(_pair("azimuth_no_wrap", (1.0, F(math.radians(359.0)), 0.0),
                     (1.05, F(math.radians(1.0)), 0.0), True, catches=(4,),
                     doc="359 deg vs 1 deg: |d az| = 358 deg, dissimilar"))
    return out


def tie_order():
    """Equal candidates: the first neighbour in j-outer / i-inner order wins
    (strict <, propag.cu:129-187).

    3 x 3 raster, ignitions (1,0) and (0,1) at 0.  One sweep.  For c = (1,1)
    the neighbour (1,0) (j = -1) comes before (0,1) (j = 0): both give
    nt = ct = 1 (ct < nt fails, nt < best takes the first), so
    c = (1, (1,0)).  i-outer order, or <= in the comparisons, gives
    (1, (0,1))."""
    W, H = 3, 3
    return _case("tie_order", W, H, _fire(W, H), {(1, 0): (0.0, (1, 0)), (0, 1): (0.0, (0, 1))},
                 100.0, 1, {(1, 1): (F(1.0), (1, 0))}, bits=set(), catches=(5, 6),
                 doc=tie_order.__doc__)


def self_reference():
    """A neighbour whose reference is the cell itself offers nothing
    (propag.cu:153-156).

    2 x 1 raster: (0,0) burns at 0 with reference (1,0); (1,0) is unburned.
    One sweep: (1,0)'s only neighbour references it, so it stays unburned
    (FLT_MAX, 0xFFFF).  Without the check, nt = 0 + 1 wins
    (ct = t(1,0) + 0 = FLT_MAX) and (1,0) = (1, (0,0))."""
    return _case("self_reference", 2, 1, _fire(2, 1), {(0, 0): (0.0, (1, 0))}, 100.0, 1,
                 {(1, 0): (FLT_MAX, NO_REF)}, bits=set(), catches=(7,), doc=self_reference.__doc__)


def max_time_gate():
    """A neighbour burns only if its time is below max_time (propag.cu:146-150),
    and the post-pass keeps a time equal to max_time (propag.cu:110-117).

    3 x 1 raster, ignition (0,0) at 10 = max_time: it spreads to nothing, so
    (1,0) stays (FLT_MAX, 0xFFFF), and (0,0) keeps 10 after the post-pass.
    With <= max_time, (1,0) takes (11, (0,0)) and the post-pass clears only
    its time: its reference (0,0) remains."""
    return _case("max_time_gate", 3, 1, _fire(3, 1), {(0, 0): (10.0, (0, 0))}, 10.0, 0,
                 {(0, 0): (F(10.0), (0, 0)), (1, 0): (FLT_MAX, NO_REF)}, bits=set(), catches=(8, 9),
                 doc=max_time_gate.__doc__)


def post_equal_kept():
    """3 x 1 raster, ignition (0,0) at 0, max_time 2, to convergence.
    (1,0) = (1, (0,0)); (2,0): nt = 1 + 1 = 2 via (1,0), ct = 0 + 2 = 2 is
    not below nt, so (2, (1,0)); 2 is not later than max_time and stays
    (propag.cu:110-117: only time > max_time is cleared)."""
    return _case("post_equal_kept", 3, 1, _fire(3, 1), {(0, 0): (0.0, (0, 0))}, 2.0, 0,
                 {(0, 0): (F(0.0), (0, 0)), (1, 0): (F(1.0), (0, 0)), (2, 0): (F(2.0), (1, 0))},
                 bits=set(), catches=(9,), doc=post_equal_kept.__doc__)


def all_cases():
    return ([fallback_taken(), line_clear(), step_field()] + thresholds()
            + [tie_order(), self_reference(), max_time_gate(), post_equal_kept()])


def mismatches(case, time, refs_x, refs_y, boundaries):
    """Expectations of `case` that a result violates (empty: all hold)."""
    W, H = case["W"], case["H"]
    bad = []
    for (x, y), (t, (ax, ay)) in case["cells"].items():
        i = x + y * W
        if np.float32(time[i]).view(np.uint32) != np.float32(t).view(np.uint32):
            bad.append(f"time{(x, y)} = {time[i]!r}, want {t!r}")
        if (int(refs_x[i]), int(refs_y[i])) != (ax, ay):
            bad.append(f"ref{(x, y)} = {(int(refs_x[i]), int(refs_y[i]))}, want {(ax, ay)}")
    if case["bits"] is not None:
        got = {(i % W, i // W) for i in range(W * H) if (int(boundaries[i >> 5]) >> (i & 31)) & 1}
        if got != case["bits"]:
            bad.append(f"boundaries {sorted(got)}, want {sorted(case['bits'])}")
    return bad
