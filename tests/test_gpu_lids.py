"""Damping lids (hbg_set_lids; the reference's SurfaceKind::Lid, lo_frequency.rs:85-97, 263-268,
1150-1156, low_order.rs:811-819; GPU). No baseline of the reference carries a lid, so the lid is
pinned by what the formulation fixes:
- eps = 0 is the free surface itself (kappa = K): the lid columns and boundary values vanish and the
  body's results equal those of the same solve without the lid (FP64, mixed; infinite and finite
  depth), although the lid unknowns stay in the system;
- with damping the results move, and the Haskind excitation (with the lid term (K - kappa) phi_i
  phi_I) keeps agreeing with the direct one to a small part of that move: a wrong sign or a dropped
  factor of the lid's boundary value, column or Haskind term breaks the agreement;
- a quadrant with its lid and both symmetry planes equals the whole body with the whole lid;
- at the frequency limits the lid unknowns are not used: the limits equal those without the lid;
- refusals: field points with lids, a vertex off z = 0, a negative or non-finite damping factor.
"""

import math
import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu

RHO, G = 1000.0, 9.80665
KEYS = ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_froude_krylov",
        "excitation_scattering", "excitation_haskind")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def ring(r0, r1, n_r, n_t, quarter):
    """Quads of the annulus r0 <= r <= r1 in z = 0 (the quadrant x, y >= 0, or the whole),
    counter-clockwise from above."""
    t = np.linspace(0.0, (0.5 if quarter else 2.0) * np.pi, n_t + 1)
    r = np.linspace(r0, r1, n_r + 1)
    out = []
    for i in range(n_r):
        for j in range(n_t):
            pts = [(r[i], t[j]), (r[i + 1], t[j]), (r[i + 1], t[j + 1]), (r[i], t[j + 1])]
            out.append([[a * math.cos(b), a * math.sin(b), 0.0] for a, b in pts])
    return np.array(out)


def body(quarter):
    q = meshes.cylinder_panels(1.0, 0.5, n_phi=8, n_z=4, n_r=4, quarter=True)
    return q if quarter else meshes.expand_symmetry(q)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def solve(depth=math.inf, lid=None, eps=0.0, omegas=None, heads=(0.0, 30.0), precision="fp64", quarter=False):
    w = np.sqrt(G * np.array([0.8, 1.6, 2.4])) if omegas is None else omegas
    with hb.Session(0) as s:
        s.set_environment(RHO, G, depth)
        if quarter:
            s.set_symmetry(True, True)
        s.set_panels(body(quarter))
        if lid is not None:
            s.set_lids(lid, eps)
        return s.solve(w, heads, precision=precision)


@pytest.mark.parametrize("precision,depth", [("fp64", math.inf), ("mixed", math.inf), ("fp64", 3.0)])
def test_lid_without_damping_leaves_the_body_results(precision, depth):
    lid = ring(1.25, 2.0, 3, 32, False)
    say(f"case: cylinder 256 panels + lid {len(lid)} panels eps 0 vs no lid, {precision}, depth {depth}")
    plain = solve(depth, precision=precision)
    lidded = solve(depth, lid, 0.0, precision=precision)
    for key in KEYS:
        e = rel(lidded[key], plain[key])
        say(f"  {key}: {e:.2e}")
        assert e < 1e-9, (key, e)


@pytest.mark.parametrize("depth", [math.inf, 3.0])
def test_lid_damping_moves_the_results_and_keeps_haskind(depth):
    lid = ring(1.25, 2.0, 3, 32, False)
    plain = solve(depth)
    damped = solve(depth, lid, 1.0)
    move = rel(damped["excitation_diffraction"], plain["excitation_diffraction"])
    move_b = rel(damped["radiation_damping"], plain["radiation_damping"])
    d_plain = rel(plain["excitation_haskind"], plain["excitation_diffraction"])
    d_damped = rel(damped["excitation_haskind"], damped["excitation_diffraction"])
    say(f"case: lid eps 1, depth {depth}: X moves {move:.2e}, B moves {move_b:.2e}; Haskind vs direct "
        f"{d_damped:.2e} (without the lid {d_plain:.2e})")
    assert move > 1e-2 and move_b > 1e-2
    assert d_damped < 0.2 * move and d_damped < 3.0 * d_plain + 2e-3


def test_lids_of_a_quadrant_equal_the_whole():
    say("case: quadrant cylinder + quarter lid with both planes vs whole, eps 0.5")
    got = solve(math.inf, ring(1.25, 2.0, 3, 8, True), 0.5, quarter=True)
    ref = solve(math.inf, ring(1.25, 2.0, 3, 32, False), 0.5)
    for key in KEYS:
        e = rel(got[key], ref[key])
        say(f"  {key}: {e:.2e}")
        assert e < 1e-9, (key, e)


def test_lid_limits_and_refusals():
    say("case: limits with a lid eps 0.5 vs without; refusals")
    lim = [0.0, math.inf]
    plain = solve(math.inf, omegas=lim, heads=())
    lidded = solve(math.inf, ring(1.25, 2.0, 3, 32, False), 0.5, omegas=lim, heads=())
    e = rel(lidded["added_mass"], plain["added_mass"])
    say(f"  added mass at 0 and inf: {e:.2e}")
    assert e < 1e-9
    with hb.Session(0) as s:
        s.set_environment(RHO, G)
        s.set_panels(body(False))
        lid = ring(1.25, 2.0, 2, 16, False)
        for bad, eps in (("off", 0.0), ("neg", -1.0), ("nan", float("nan"))):
            v = lid.copy()
            if bad == "off":
                v[3, 1, 2] = 1e-3
            with pytest.raises(hb.HbgError) as err:
                s.set_lids(v, eps)
            assert err.value.status == 1
        s.set_lids(lid, 0.5)
        s.set_field_points(np.array([[3.0, 0.0, -0.5]]))
        with pytest.raises(hb.HbgError) as err:
            s.solve([2.0], [0.0])
        assert err.value.status == 3 and "field points with lids" in err.value.message
