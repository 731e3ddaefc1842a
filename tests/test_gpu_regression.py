"""The reference's own regression baselines (tests/data/regression/*.npz of the reference, copied
into tests/golden/, loaded with allow_pickle=False) for the outputs of this path beyond A, B, X:
Kochin functions, motions, the frequency limits, field-point pressures, walls in finite depth
(GPU).

The baselines are the reference's production results (complex64 matrices and LU, float32 Rankine
tiers, results stored as float32); this path computes in FP64, as the reference's double mode
does. Two rules hold every array:
- the reference's rule between its production and double modes
  (tests/python/test_double_precision_mode.py:48-53): max |a - b| <= 1e-5 of the array's peak;
- the per-entry rule of its regression test (tests/python/test_regression.py:25-75: each entry
  against its own largest magnitude over the first axis, at least 1e-3 of the peak), whose RTOL
  1e-5 is meant for a rerun of the same production arithmetic; between the two precisions it is
  held at 5e-5 (an entry of a column that is nearly zero, e.g. the heave column of an angular
  derivative, sits at the float32 floor of the baseline: 2e-5 of its 1e-3-of-peak scale is
  2e-8 of the peak). The problems are those of tests/python/regression_problems.py (the reference): rho 1000,
g 9.80665, omega = sqrt(0.5 g), sqrt(1.5 g), headings 0 and 30 degrees.
"""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from conftest import GOLDEN
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu

RHO, G = 1000.0, 9.80665
RTOL, ATOL = 1e-5, 1e-14   # the regression rule's tolerances (test_regression.py:27-28)
PEAK_TOL = 1e-5             # production vs double mode, of the peak (test_double_precision_mode.py:52)
ENTRY_TOL = 5e-5            # the per-entry rule across the two precisions
POINTS = np.array([[3.0, 0.5, 0.0], [2.0, 2.0, -0.4], [-2.5, 1.0, -0.2]])


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def regression_error(a, b, rtol=RTOL):
    """max |a - b| / scale under the reference's regression rule (scale: per-entry largest magnitude
    over the first axis, at least 1e-3 of the peak), and whether every entry is within
    rtol scale + ATOL."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b).astype(np.complex128)
    assert a.shape == b.shape, (a.shape, b.shape)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak) if b.ndim > 1 else peak
    d = np.abs(a - b)
    return float((d / scale).max()), bool(np.all(d <= rtol * scale + ATOL))


def peak_error(a, b):
    """max |a - b| / max |b| (the production-vs-double rule)."""
    b = np.asarray(b).astype(np.complex128)
    return float(np.abs(np.asarray(a, dtype=np.complex128) - b).max() / np.abs(b).max())


def check(name, got, ref):
    err, ok = regression_error(got, ref, ENTRY_TOL)
    pk = peak_error(got, ref)
    say(f"  {name}: {pk:.2e} of the peak, {err:.2e} of the per-entry scale")
    assert pk <= PEAK_TOL and ok, (name, pk, err)


def quadrant_cylinder():
    # regression_problems.py _cylinder / bodies.py lo_free_cylinder: R 1, draft 0.5, quadrant of
    # 8 x 4 side and 8 x 4 bottom panels, symmetric about x = 0 and y = 0
    return meshes.cylinder_panels(1.0, 0.5, n_phi=8, n_z=4, n_r=4, quarter=True)


def rigid_mass_matrix(m, cog, inertia_cog, centre=(0.0, 0.0, 0.0)):
    """6 x 6 mass matrix of a rigid body about `centre`: [[m I, -m [c]x], [m [c]x, J]] with c the
    centre of gravity relative to the centre and J the inertia moved there (parallel axes)."""
    c = np.asarray(cog, float) - np.asarray(centre, float)
    cx = np.array([[0.0, -c[2], c[1]], [c[2], 0.0, -c[0]], [-c[1], c[0], 0.0]])
    J = np.asarray(inertia_cog, float) + m * (np.dot(c, c) * np.eye(3) - np.outer(c, c))
    M = np.zeros((6, 6))
    M[:3, :3] = m * np.eye(3)
    M[:3, 3:] = -m * cx
    M[3:, :3] = m * cx
    M[3:, 3:] = J
    return M


@pytest.mark.parametrize("symmetric", [False, True], ids=["whole", "quadrant"])
def test_kochin_functions_and_motions_of_the_drift_cylinder(symmetric):
    say(f"case: lo_cylinder_drift Kochin functions (104 angles) and motions, {'quadrant' if symmetric else 'whole'}")
    ref = np.load(GOLDEN / "lo_cylinder_drift.npz", allow_pickle=False)
    NK = len(ref["kochin_angles"])
    assert np.allclose(ref["kochin_angles"], 2 * np.pi * np.arange(NK) / NK)
    with hb.Session(0) as s:
        s.set_environment(RHO, G)
        if symmetric:
            s.set_panels(quadrant_cylinder())
            s.set_symmetry(True, True)
        else:
            s.set_panels(meshes.expand_symmetry(quadrant_cylinder()))
        s.set_kochin(NK)
        res = s.solve(ref["omegas"], ref["headings"])
        k = s.kochin_values()
        for key, got in (("kochin_radiation", k["kr"]), ("kochin_radiation_derivative", k["krt"]),
                         ("kochin_diffraction", k["kd"]), ("kochin_diffraction_derivative", k["kdt"])):
            check(key, got, ref[key])
        # bodies.py _free: mass 1000 pi / 2, centre of gravity (0, 0, -0.2), Ixx = Iyy = 0.36 m,
        # Izz = 0.49 m about it; the stiffness of the baseline's hydrostatics; rotation centre 0
        m = float(ref["hydrostatics_mass"][0])
        M = rigid_mass_matrix(m, ref["hydrostatics_center_of_gravity"][0], np.diag([0.36 * m, 0.36 * m, 0.49 * m]))
        C = ref["hydrostatics_stiffness"][0]
        rao = s.motions(res, M, C)
        check("motion", rao, ref["motion"])


@pytest.mark.parametrize("form", ["whole mesh", "quadrant with two symmetry planes"])
def test_frequency_limits_of_the_cylinder(form):
    say(f"case: lo_cylinder_limits: omega 0, sqrt(g/2), infinity ({form})")
    ref = np.load(GOLDEN / "lo_cylinder_limits.npz", allow_pickle=False)
    with hb.Session(0) as s:
        s.set_environment(RHO, G)
        if form == "whole mesh":
            s.set_panels(meshes.expand_symmetry(quadrant_cylinder()))
        else:  # the problem as the reference states it (regression_problems.py)
            s.set_symmetry(True, True)
            s.set_panels(quadrant_cylinder())
        res = s.solve([0.0, float(ref["omegas"][0]), np.inf], [])
        check("added_mass", res["added_mass"][1:2], ref["added_mass"])
        check("radiation_damping", res["radiation_damping"][1:2], ref["radiation_damping"])
        check("added_mass_zero_frequency", res["added_mass"][0], ref["added_mass_zero_frequency"])
        check("added_mass_infinite_frequency", res["added_mass"][2], ref["added_mass_infinite_frequency"])


def test_field_point_pressures_of_the_cylinder():
    say("case: lo_cylinder_field_body_points: field-point pressures, quadrant with two symmetry planes")
    ref = np.load(GOLDEN / "lo_cylinder_field_body_points.npz", allow_pickle=False)
    assert np.array_equal(ref["field_points"], POINTS)
    with hb.Session(0) as s:
        s.set_environment(RHO, G)
        s.set_panels(quadrant_cylinder())
        s.set_symmetry(True, True)
        s.set_field_points(POINTS)
        res = s.solve(ref["omegas"], ref["headings"])
        for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
            check(key, res[key], ref[key])
        f = s.field_values()
        # p / (rho g) below z = 0 and the elevation on it: -(i w / g) phi, per unit displacement of
        # the radiation columns
        for key, got in (("field_points_pressure_radiation", f["frad"]),
                         ("field_points_pressure_incident", f["finc"]),
                         ("field_points_pressure_scattering", f["fsca"]),
                         ("field_points_pressure_diffraction", f["ftot"])):
            check(key, got, ref[key])


def test_box_between_walls_in_finite_depth():
    say("case: lo_box_walls_finite: box at (2, 1.5) by the walls x = 0 and y = 0, depth 3, field points")
    ref = np.load(GOLDEN / "lo_box_walls_finite.npz", allow_pickle=False)
    pts = POINTS + np.array([3.0, 3.0, 0.0])
    assert np.array_equal(ref["field_points"], pts)
    pos = np.array([2.0, 1.5, 0.0])
    # regression_problems.py _box: the quadrant box_panels(1, 0.5, 0.5, (6, 3, 3)) with its two
    # symmetry planes about the body's own axes, placed at (2, 1.5): written out as the whole box
    box = meshes.place(meshes.expand_symmetry(meshes.box_panels(1.0, 0.5, 0.5, n=(6, 3, 3), quarter=True)), pos)
    with hb.Session(0) as s:
        s.set_environment(RHO, G, 3.0)
        s.set_panels(box, rotation_centres=pos[None, :])
        s.set_walls(wall_x=True, wall_y=True)
        s.set_field_points(pts)
        res = s.solve(ref["omegas"], ref["headings"])
        for key in ("excitation_diffraction", "excitation_froude_krylov", "excitation_scattering"):
            check(key, res[key], ref[key])
        f = s.field_values()
        for key, got in (("field_points_pressure_incident", f["finc"]),
                         ("field_points_pressure_scattering", f["fsca"]),
                         ("field_points_pressure_diffraction", f["ftot"])):
            check(key, got, ref[key])
