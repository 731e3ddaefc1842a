"""Symmetry planes (GPU): a sector of a symmetric body solved with hbg_set_symmetry against the whole
mesh solved without it, for the planes x = 0, y = 0 and both, FP64 and mixed precision, infinite and
finite depth, headings that break the symmetry. The character systems and the whole mesh are the
same problem, so the two agree to the solver's rounding; the whole mesh itself is held to the FP64
oracle in the parity tests, and once here. The reference holds its sectors to the whole body within
1e-6 (tests/python/test_low_order.py:143-155)."""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu
KEYS = ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_froude_krylov",
        "excitation_scattering", "excitation_haskind")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_per_entry(a, b):
    b = np.asarray(b)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def reflect(v, sx, sy):
    """The panels reflected (x -> sx x, y -> sy y), the vertex order reversed where a single plane
    reverses the orientation."""
    r = np.asarray(v, dtype=float) * np.array([sx, sy, 1.0])
    return r[:, [0, 3, 2, 1]] if sx * sy < 0 else r


def sector_and_whole(quadrant, sym_x, sym_y):
    """The stored sector for the planes asked and the whole mesh in the image order (e, x, y, xy)."""
    q = np.asarray(quadrant, dtype=float)
    if sym_x and sym_y:
        return q, meshes.expand_symmetry(q)
    half = np.concatenate([q, reflect(q, 1.0, -1.0)]) if sym_x else np.concatenate([q, reflect(q, -1.0, 1.0)])
    s = (-1.0, 1.0) if sym_x else (1.0, -1.0)
    return half, np.concatenate([half, reflect(half, *s)])


@pytest.fixture(scope="module")
def sessions():
    a, b = hb.Session(0), hb.Session(0)
    yield a, b
    a.close()
    b.close()


CASES = [
    ("hemisphere", float("inf"), 1, 1),
    ("hemisphere", float("inf"), 1, 0),
    ("hemisphere", float("inf"), 0, 1),
    ("cylinder", 3.0, 1, 1),
    ("cylinder", 3.0, 0, 1),
    ("box", 2.0, 1, 0),
    ("cylinder_large", 3.0, 1, 1),  # a 320-panel sector: five 64-column LU panels per character system
    ("cylinder_large", float("inf"), 1, 0),
]


@pytest.mark.parametrize("shape,depth,sym_x,sym_y", CASES)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sector_solve_equals_whole_mesh(sessions, shape, depth, sym_x, sym_y, precision):
    quadrant = {
        "hemisphere": lambda: meshes.hemisphere_panels(1.0, 6, 7, quarter=True),
        "cylinder": lambda: meshes.cylinder_panels(1.0, 0.8, 9, 4, 3, quarter=True),
        "box": lambda: meshes.box_panels(1.0, 0.6, 0.5, (5, 3, 3), quarter=True),
        "cylinder_large": lambda: meshes.cylinder_panels(1.0, 1.0, 16, 12, 8, quarter=True, cosine=True),
    }[shape]()
    sector, whole = sector_and_whole(quadrant, sym_x, sym_y)
    centre = np.array([[0.3, -0.2, -0.1]])  # off the planes: the modes of the whole body still hold
    w = np.array([0.5, 1.1, 2.3])
    heads = [0.0, 30.0, 90.0]
    say(f"case: {shape} sector {len(sector)} of {len(whole)} panels, planes x {sym_x} y {sym_y}, depth {depth}, "
        f"{precision}")
    s, f = sessions
    for x in (s, f):
        x.set_environment(1025.0, 9.81, depth)
    s.set_symmetry(sym_x, sym_y)
    s.set_panels(sector, rotation_centres=centre)
    f.set_symmetry(False, False)
    f.set_panels(whole, rotation_centres=centre)
    got = s.solve(w, heads, precision=precision)
    ref = f.solve(w, heads, precision=precision)
    for key in KEYS:
        e = rel_per_entry(got[key], ref[key])
        say(f"  {key}: sector vs whole {e:.2e}")
        assert e < 1e-9


def test_whole_body_quantities_against_the_oracle(sessions, oracle):
    quadrant = meshes.hemisphere_panels(1.0, 4, 4, quarter=True)
    sector, whole = sector_and_whole(quadrant, 1, 1)
    say(f"case: hemisphere sector {len(sector)} of {len(whole)} panels, both planes, FP64 vs the oracle")
    s = sessions[0]
    s.set_environment(1025.0, 9.81, float("inf"))
    s.set_symmetry(True, True)
    s.set_panels(sector)
    w = np.array([0.7, 1.6])
    got = s.solve(w, [0.0, 45.0])
    orc = oracle.Problem(whole, rho=1025.0, grav=9.81, depth=float("inf")).solve(w, [0.0, 45.0])
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        e = rel_per_entry(got[key], orc[key])
        say(f"  {key}: vs oracle {e:.2e}")
        assert e < 1e-10


def test_symmetry_before_or_after_the_panels_and_switched_off(sessions):
    quadrant = meshes.hemisphere_panels(1.0, 4, 5, quarter=True)
    whole = meshes.expand_symmetry(quadrant)
    s, f = sessions
    for x in (s, f):
        x.set_environment(1025.0, 9.81, float("inf"))
    say("case: hemisphere 20-panel quadrant, symmetry set after the panels, then switched off")
    f.set_symmetry(False, False)
    f.set_panels(whole)
    ref = f.solve([1.0], [20.0])
    s.set_symmetry(False, False)
    s.set_panels(quadrant)
    s.set_symmetry(True, True)
    got = s.solve([1.0], [20.0])
    for key in KEYS:
        assert rel_per_entry(got[key], ref[key]) < 1e-9
    s.set_symmetry(False, False)
    s.set_panels(whole)
    again = s.solve([1.0], [20.0])
    for key in KEYS:
        assert rel_per_entry(again[key], ref[key]) < 1e-12


def test_unsupported_uses_with_symmetry_are_refused(sessions):
    s = sessions[0]
    s.set_environment(1025.0, 9.81, float("inf"))
    s.set_symmetry(True, False)
    s.set_panels(meshes.hemisphere_panels(1.0, 3, 3, quarter=True))
    say("case: symmetry planes with the statics, the assembly")
    try:
        for call in (lambda: s.statics(),
                     lambda: s.assemble(1.0)):
            with pytest.raises(hb.HbgError) as e:
                call()
            assert e.value.status == 3
        with pytest.raises(hb.HbgError):
            hb._check(hb.library().hbg_set_symmetry(s._h, 2, 0))
    finally:
        s.set_symmetry(False, False)


LIMIT_CASES = [
    ("hemisphere", float("inf"), 1, 1),
    ("hemisphere", float("inf"), 0, 1),
    ("cylinder", 3.0, 1, 1),
    ("box", 2.0, 1, 0),
]


@pytest.mark.parametrize("shape,depth,sym_x,sym_y", LIMIT_CASES)
def test_limits_of_a_sector_equal_the_whole_mesh(sessions, shape, depth, sym_x, sym_y):
    """The radiation limits omega = 0 and infinity (and a frequency between) of a sector against the
    whole mesh. In finite depth at omega = 0 the modes with a net volume flux through the whole
    body (heave, and roll and pitch about a centre off the axis) are NaN on both sides."""
    quadrant = {
        "hemisphere": lambda: meshes.hemisphere_panels(1.0, 6, 7, quarter=True),
        "cylinder": lambda: meshes.cylinder_panels(1.0, 0.8, 9, 4, 3, quarter=True),
        "box": lambda: meshes.box_panels(1.0, 0.6, 0.5, (5, 3, 3), quarter=True),
    }[shape]()
    sector, whole = sector_and_whole(quadrant, sym_x, sym_y)
    centre = np.array([[0.3, -0.2, -0.1]])
    w = np.array([0.0, 1.1, np.inf])
    say(f"case: limits of a {shape} sector {len(sector)} of {len(whole)} panels, planes x {sym_x} y {sym_y}, "
        f"depth {depth}")
    s, f = sessions
    for x in (s, f):
        x.set_environment(1025.0, 9.81, depth)
    s.set_symmetry(sym_x, sym_y)
    s.set_panels(sector, rotation_centres=centre)
    f.set_symmetry(False, False)
    f.set_panels(whole, rotation_centres=centre)
    got = s.solve(w, [0.0])["added_mass"]
    ref = f.solve(w, [0.0])["added_mass"]
    s.set_symmetry(False, False)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    if np.isfinite(depth):
        assert nan[0, 2, :].all() and nan[0, 3, :].all() and not nan[0, 0, 0] and not nan[1:].any()
    else:
        assert not nan.any()
    for i, label in enumerate(("omega 0", "omega 1.1", "omega infinity")):
        e = rel_per_entry(np.nan_to_num(got[i]), np.nan_to_num(ref[i]))
        say(f"  {label}: sector vs whole {e:.2e}")
        assert e < 1e-9


@pytest.mark.parametrize("depth", [float("inf"), 3.0])
def test_field_points_and_kochin_of_a_sector_equal_the_whole_mesh(sessions, depth):
    quadrant = meshes.cylinder_panels(1.0, 0.8, 9, 4, 3, quarter=True)
    whole = meshes.expand_symmetry(quadrant)
    pts = np.array([[2.0, 0.5, 0.0], [-1.5, 2.5, -0.3], [0.3, -0.2, -0.4], [-6.0, -4.0, 0.0]])
    w, heads = np.array([0.7, 1.6]), [0.0, 30.0]
    say(f"case: cylinder sector {len(quadrant)} of {len(whole)} panels, both planes, depth {depth}, 4 field points, "
        f"36 Kochin angles")
    s, f = sessions
    res = []
    for x, v, sym in ((s, quadrant, True), (f, whole, False)):
        x.set_environment(1025.0, 9.81, depth)
        x.set_symmetry(sym, sym)
        x.set_panels(v)
        x.set_field_points(pts)
        x.set_kochin(36)
        x.solve(w, heads)
        res.append((x.field_values(2, 2), x.kochin_values(2, 2)))
        x.set_field_points(None)
        x.set_kochin(0)
    s.set_symmetry(False, False)
    for key in ("frad", "finc", "fsca", "ftot"):
        e = rel_per_entry(res[0][0][key], res[1][0][key])
        say(f"  {key}: sector vs whole {e:.2e}")
        assert e < 1e-9
    for key in ("kr", "krt", "kd", "kdt"):
        e = rel_per_entry(res[0][1][key], res[1][1][key])
        say(f"  {key}: sector vs whole {e:.2e}")
        assert e < 1e-9


def test_bench_body_as_its_quadrant_equals_the_whole_mesh(sessions):
    # the 8192-panel finite-depth cylinder of the benchmark and its 2048-panel quadrant with both planes
    quadrant = meshes.cylinder_panels(1.0, 1.0, 32, 32, 32, quarter=True, cosine=True)
    whole = meshes.expand_symmetry(quadrant)
    nu = np.array([0.05, 0.6, 4.0, 20.0]) / 3.0
    w = np.sqrt(9.81 * nu)
    heads = [0.0, 30.0]
    say(f"case: cylinder quadrant {len(quadrant)} of {len(whole)} panels, depth 3, 4 frequencies, FP64 and mixed")
    s, f = sessions
    for x in (s, f):
        x.set_environment(1025.0, 9.81, 3.0)
    f.set_symmetry(False, False)
    f.set_panels(whole)
    ref = f.solve(w, heads)
    s.set_symmetry(True, True)
    s.set_panels(quadrant)
    for precision in ("fp64", "mixed"):
        got = s.solve(w, heads, precision=precision)
        for key in KEYS:
            e = rel_per_entry(got[key], ref[key])
            say(f"  {precision} {key}: quadrant vs whole {e:.2e}")
            assert e < 1e-9
    s.set_symmetry(False, False)
    f.set_panels(meshes.hemisphere_panels(1.0, 2, 4, quarter=False))  # release the large buffers' use
