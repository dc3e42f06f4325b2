"""Pins of the oracle's hierarchical integrator (main.go:159-199) at 1e-12.

The reference's own tests check this integrator only to 0.05 or DS
(main_test.go:365-522).  A dropped right-endpoint term or a fine walk that
starts at `left` moves T by rho*DS/10 per transition, which they cannot see.
These tests fix T exactly from the definition of the two FP64 grids the
integrator walks:

* coarse samples R_0 = smin + DS, R_{n+1} = fl(R_n + DS) while R_n <= smax;
* on a transition of (rho == 0) between R_{n-1} and R_n (R_{-1} = smin,
  rho(R_{-1}) taken as 0) the fine samples f_1 = fl(R_{n-1} + DS/10),
  f_{m+1} = fl(f_m + DS/10) while f_m < R_n, each weighted DS/10, plus
  rho(R_n) * DS/10; otherwise rho(R_n) * DS.

Densities come from analytic chords of spheres, slabs and cylinders, not
from the oracle's own membership tests.  The last test compiles the oracle
with planted mistakes (xrp_oracle.c, ORC_MUTANT) and shows that every one of
them fails a pin here.
"""
import math

import numpy as np
import pytest

from helpers import O
from test_oracle_closed_forms import (chord_box, chord_cylinder, chord_sphere,
                                      random_rays)


def coarse_grid(smin, smax, DS):
    out, r = [], smin + DS
    while r <= smax:
        out.append(r)
        r += DS
    return np.array(out)


def fine_grid(left, right, DS):
    ds = DS / 10.0
    out, f = [], left + ds
    while f < right:
        out.append(f)
        f += ds
    return np.array(out)


def simple_grid(smin, smax, ds):
    out, s = [], smin
    while s < smax:
        out.append(s)
        s += ds
    return np.array(out)


def hierarchical_identity(dens, smin, smax, DS, ff=0.0):
    """T from the grid definition above; dens maps an array of s to rho."""
    R = coarse_grid(smin, smax, DS)
    if len(R) == 0:
        return ff
    rho = dens(R)
    prev = np.concatenate([[0.0], rho[:-1]])
    trans = (rho == 0) != (prev == 0)
    lefts = np.concatenate([[smin], R[:-1]])
    ds = DS / 10.0
    T = ff + DS * float(np.sum(rho[~trans]))
    for n in np.nonzero(trans)[0]:
        F = fine_grid(lefts[n], R[n], DS)
        T += ds * (float(np.sum(dens(F))) if len(F) else 0.0) + ds * rho[n]
    return T


def chord_density(chords):
    """rho(s) = clamp(sum of rho_i over chords (t_in, t_out) holding s, 0, 1)
    (an object_collection of convex members, objects.go:422-438)."""
    def dens(s):
        s = np.asarray(s, dtype=np.float64)
        tot = np.zeros_like(s)
        for ch, rho in chords:
            if ch is not None:
                tot += rho * ((s > ch[0]) & (s < ch[1]))
        return np.clip(tot, 0.0, 1.0) if len(chords) > 1 else tot
    return dens


# ---- uniform medium: every coarse sample inside ---------------------------
@pytest.mark.parametrize("smin,smax,DS,rho,ff", [
    (2.26, 5.74, 0.05, 1.0, 0.0), (2.26, 5.74, 0.013, 0.7, 0.0),
    (0.0, 1.0, 0.1, 2.0, 0.3), (4.5, 5.5, 0.0071, 1.3, 0.0), (2.26, 5.74, 0.25, 0.4, 1.0),
    (0.0, 5.0, 0.625, 1.0, 0.0)])
def test_uniform_medium_closed_form(smin, smax, DS, rho, ff):
    """Only the first coarse step is a transition (prev_rho starts at 0):
    T = ff + rho*(DS/10)*(#fine(smin, R_0) + 1) + rho*DS*(N - 1).
    With DS = 0.625 (DS/10 = 1/16) both grids are exact, the fine grid meets
    R_0 itself and the walk stops before it (9 fine samples)."""
    box = O.build_object({"type": "cube", "center": [0.0, 0.0, 0.0], "side": 30.0,
                          "rho": rho}, numbers="json")
    sc = O.Scene(box, flat_field=ff)
    R = coarse_grid(smin, smax, DS)
    nf = len(fine_grid(smin, R[0], DS))
    assert nf in (9, 10)
    want = ff + rho * (DS / 10.0) * (nf + 1) + rho * DS * (len(R) - 1)
    for o, d in ([0, 0, -5.0], [0, 0, 1.0]), ([1.0, -2.0, 0.5], [0.6, 0.0, 0.8]):
        got = -math.log(sc.integrate_hierarchical(o, d, DS, smin, smax))
        assert abs(got - want) <= 1e-12, (got, want)


def test_simple_uniform_medium_closed_form():
    """T = ff + rho*ds*#{s_k < smax}; ds = 1/8 makes s_k exact, so the grid
    reaches smax itself and a `<=` bound would add a sample."""
    box = O.build_object({"type": "cube", "center": [0.0, 0.0, 0.0], "side": 30.0,
                          "rho": 1.5}, numbers="json")
    sc = O.Scene(box, flat_field=0.25)
    got = -math.log(sc.integrate_simple([0, 0, -5.0], [0, 0, 1.0], 0.125, 0.0, 1.0))
    assert abs(got - (0.25 + 1.5 * 0.125 * 8)) <= 1e-12


# ---- one chord along the ray ------------------------------------------------
def single_chord_closed_form(a, b, rho, smin, smax, DS):
    """For one convex member with chord (a, b): entry step i (first R in the
    chord) and exit step e (first R after it outside) give
    T = rho*ds*(#(F_i in chord) + 1) + rho*DS*(e - i - 1) + rho*ds*#(F_e in chord)."""
    R = coarse_grid(smin, smax, DS)
    inside = (R > a) & (R < b)
    if not inside.any():
        return 0.0
    i = int(np.argmax(inside))
    outside_after = np.nonzero(~inside[i:])[0]
    e = i + int(outside_after[0]) if len(outside_after) else len(R)
    ds = DS / 10.0
    lefts = np.concatenate([[smin], R[:-1]])

    def cnt(n):
        F = fine_grid(lefts[n], R[n], DS)
        return int(np.sum((F > a) & (F < b)))
    T = rho * ds * (cnt(i) + 1) + rho * DS * (e - i - 1)
    if e < len(R):
        T += rho * ds * cnt(e)
    return T


@pytest.mark.parametrize("prim", ["sphere", "box", "cylinder"])
def test_single_chord_closed_form(prim):
    DS, smin, smax = 0.037, 4.0 - 1.74, 4.0 + 1.74
    rho = 0.7
    if prim == "sphere":
        obj = {"type": "sphere", "center": [0.05, -0.02, 0.1], "radius": 0.33, "rho": rho}
        chord = lambda o, d: chord_sphere(o, d, np.array(obj["center"]), obj["radius"])
    elif prim == "box":
        obj = {"type": "box", "center": [0.05, -0.02, 0.1], "sides": [0.5, 0.3, 0.7],
               "rho": rho}
        chord = lambda o, d: chord_box(o, d, np.array(obj["center"]),
                                       np.array(obj["sides"]) / 2)
    else:
        obj = {"type": "cylinder", "p0": [-0.3, 0.1, -0.2], "p1": [0.25, -0.1, 0.3],
               "radius": 0.12, "rho": rho}
        chord = lambda o, d: chord_cylinder(o, d, np.array(obj["p0"]), np.array(obj["p1"]),
                                            obj["radius"])
    sc = O.Scene(O.build_object(obj, numbers="json"))
    hits = 0
    for o, d in random_rays(300, 13):
        ch = chord(o, d)
        want = 0.0 if ch is None else single_chord_closed_form(ch[0], ch[1], rho, smin,
                                                               smax, DS)
        hits += want > 0
        got = -math.log(sc.integrate_hierarchical(o, d, DS, smin, smax))
        assert abs(got - want) <= 1e-12, (o, d, got, want)
    assert hits > 50


def test_thin_slab_between_coarse_samples_is_missed():
    """A slab thinner than DS lying strictly between two coarse samples
    causes no transition, so T is the flat field alone; moved onto a coarse
    sample it is counted on the fine grids of both neighbouring steps."""
    DS, smin, smax = 0.05, 4.5, 5.5
    R = coarse_grid(smin, smax, DS)
    mid = 0.5 * (R[7] + R[8]) - 5.0           # z of the gap centre
    slab = lambda zc: O.Scene(O.build_object(
        {"type": "box", "center": [0.0, 0.0, zc], "sides": [4.0, 4.0, 0.02], "rho": 1.0},
        numbers="json"), flat_field=0.125)
    got = -math.log(slab(mid).integrate_hierarchical([0, 0, -5.0], [0, 0, 1.0], DS,
                                                     smin, smax))
    assert abs(got - 0.125) <= 1e-15
    zc = R[8] - 5.0 + 0.0013                  # covers R[8]
    a, b = 5.0 + zc - 0.01, 5.0 + zc + 0.01
    want = 0.125 + single_chord_closed_form(a, b, 1.0, smin, smax, DS)
    got = -math.log(slab(zc).integrate_hierarchical([0, 0, -5.0], [0, 0, 1.0], DS,
                                                    smin, smax))
    assert abs(got - want) <= 1e-12
    assert want > 0.125 + 0.009               # about 0.02 of chord, at DS/10 resolution


# ---- interval-count identity over collections ------------------------------
@pytest.mark.parametrize("DS", [0.02, 0.05, 0.013])
def test_hierarchical_identity_csg(DS):
    """cube minus sphere minus cylinder (clamped sum), random rays."""
    smin, smax = 2.26, 5.74
    obj = {"type": "object_collection", "objects": [
        {"type": "cube", "center": [0.0, 0.0, 0.0], "side": 1.0, "rho": 1.0},
        {"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": -1.0},
        {"type": "cylinder", "p0": [-0.6, 0.0, 0.0], "p1": [0.6, 0.0, 0.0], "radius": 0.15,
         "rho": -1.0}]}
    sc = O.Scene(O.build_object(obj, numbers="json"), flat_field=0.05)
    trans = 0
    for o, d in random_rays(200, 17):
        chords = [(chord_box(o, d, np.zeros(3), np.full(3, 0.5)), 1.0),
                  (chord_sphere(o, d, np.zeros(3), 0.3), -1.0),
                  (chord_cylinder(o, d, np.array([-0.6, 0, 0]), np.array([0.6, 0, 0]),
                                  0.15), -1.0)]
        dens = chord_density(chords)
        want = hierarchical_identity(dens, smin, smax, DS, ff=0.05)
        R = coarse_grid(smin, smax, DS)
        rho = dens(R)
        trans += int(np.sum((rho == 0) != (np.concatenate([[0.0], rho[:-1]]) == 0)))
        got = -math.log(sc.integrate_hierarchical(o, d, DS, smin, smax))
        assert abs(got - want) <= 1e-12, (o, d, got, want)
    assert trans > 200


def test_hierarchical_identity_overlapping_spheres():
    """Two overlapping positive spheres: the clamp to 1 inside the overlap
    and interior transitions where only the density value changes (no
    transition of rho == 0)."""
    smin, smax, DS = 2.26, 5.74, 0.03
    c1, c2 = np.array([0.1, 0.0, 0.0]), np.array([-0.1, 0.05, 0.0])
    obj = {"type": "object_collection", "objects": [
        {"type": "sphere", "center": list(c1), "radius": 0.3, "rho": 0.6},
        {"type": "sphere", "center": list(c2), "radius": 0.25, "rho": 0.7}]}
    sc = O.Scene(O.build_object(obj, numbers="json"))
    for o, d in random_rays(200, 19):
        dens = chord_density([(chord_sphere(o, d, c1, 0.3), 0.6),
                              (chord_sphere(o, d, c2, 0.25), 0.7)])
        want = hierarchical_identity(dens, smin, smax, DS)
        got = -math.log(sc.integrate_hierarchical(o, d, DS, smin, smax))
        assert abs(got - want) <= 1e-12, (o, d, got, want)


# ---- planted mistakes --------------------------------------------------------
def _pins_pass(L) -> bool:
    """Runs a compact version of the pins above through library L."""
    try:
        with O.using_library(L):
            test_uniform_medium_closed_form(2.26, 5.74, 0.05, 1.0, 0.0)
            test_uniform_medium_closed_form(0.0, 1.0, 0.1, 2.0, 0.3)
            test_uniform_medium_closed_form(0.0, 5.0, 0.625, 1.0, 0.0)
            test_simple_uniform_medium_closed_form()
            test_thin_slab_between_coarse_samples_is_missed()
            test_single_chord_closed_form("sphere")
    except AssertionError:
        return False
    return True


def test_faithful_oracle_passes_compact_pins():
    assert _pins_pass(O.lib())


@pytest.mark.parametrize("mutant", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_fail_the_pins(mutant, tmp_path):
    """Each ORC_MUTANT build of xrp_oracle.c (a dropped, shifted, extended
    or doubled term of either integrator) fails at least one pin."""
    L = O.build_mutant(mutant, str(tmp_path))
    assert not _pins_pass(L)
