"""Session state after refused calls, and inputs refused before any launch (GPU: a session needs
a device, though the refused calls themselves launch nothing).

- a refused hbg_set_panels (a bad panel, a panel behind a wall) leaves the session as it was: on a
  fresh session a solve is then a state error, after a valid geometry the old panels still solve
  to the same bits;
- hbg_motions refuses non-finite inputs with HBG_ERR_VALUE before its kernel runs;
- the field and Kochin values take their sizes from the library's record of the last solve, and a
  new geometry invalidates them."""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def bad_mesh(v):
    """The mesh with its 4th panel collapsed to a line (fewer than three distinct vertices)."""
    b = np.array(v, copy=True)
    b[3, :, :] = b[3, 0, :]
    return b


def test_refused_panels_on_a_fresh_session_leave_no_geometry():
    say("case: set_panels refused on a fresh session, then solve")
    with hb.Session(0) as s:
        v = meshes.hemisphere_panels(1.0, 4, 8, quarter=False)
        with pytest.raises(hb.HbgError) as e:
            s.set_panels(bad_mesh(v))
        assert e.value.status == 1 and "panel 4" in e.value.message
        with pytest.raises(hb.HbgError) as e:
            s.solve([1.0], [0.0])
        assert e.value.status == 4  # STATE: no panels set, nothing launched


def panel_cases():
    """Quads that the panel rule (geometry.rs:120-199) refuses, with the words of its message."""
    sq = np.array([[0.0, 0.0, -1.0], [1.0, 0.0, -1.0], [1.0, 1.0, -1.0], [0.0, 1.0, -1.0]])
    nan = sq.copy()
    nan[2, 1] = np.nan
    inf = sq.copy()
    inf[0, 2] = -np.inf
    point = np.repeat(sq[:1], 4, axis=0)
    line = np.array([[0.0, 0.0, -1.0], [1.0, 0.0, -1.0], [2.0, 0.0, -1.0], [3.0, 0.0, -1.0]])
    dart = np.array([[0.0, 0.0, -1.0], [2.0, 0.0, -1.0], [0.5, 0.5, -1.0], [0.0, 2.0, -1.0]])
    bowtie = np.array([[0.0, 0.0, -1.0], [1.0, 1.0, -1.0], [1.0, 0.0, -1.0], [0.0, 1.0, -1.0]])
    return [("nan", nan, "non-finite"), ("inf", inf, "non-finite"), ("point", point, "distinct vertices"),
            ("line", line, "zero area"), ("dart", dart, "convex"), ("bow tie", bowtie, "")]


@pytest.mark.parametrize("label,quad,words", panel_cases(), ids=[c[0] for c in panel_cases()])
def test_panel_rule_refusals_name_the_panel(label, quad, words):
    # failure injection at the ABI: the panel rule's refusals are HBG_ERR_VALUE naming the panel
    # (1-based, as the reference's messages), before anything is built or launched
    say(f"case: set_panels refuses a {label} quad as panel 5")
    with hb.Session(0) as s:
        v = np.array(meshes.hemisphere_panels(1.0, 4, 8, quarter=False), copy=True)
        v[4] = quad
        with pytest.raises(hb.HbgError) as e:
            s.set_panels(v)
        assert e.value.status == 1 and "panel 5" in e.value.message, e.value.message
        assert words in e.value.message, e.value.message
        assert s.n_panels == 0


def test_refused_panels_keep_the_previous_geometry():
    say("case: valid geometry, refused set_panels / set_walls, the old geometry solves to the same bits")
    with hb.Session(0) as s:
        v = meshes.hemisphere_panels(1.0, 6, 12, quarter=False)
        s.set_environment(1025.0, 9.81)
        s.set_panels(v)
        before = s.solve([1.0, 2.0], [0.0])
        big = meshes.hemisphere_panels(1.0, 8, 16, quarter=False)
        with pytest.raises(hb.HbgError) as e:
            s.set_panels(bad_mesh(big))
        assert e.value.status == 1
        # a wall at x = 0 with panels on both sides: refused, the walls stay off
        with pytest.raises(hb.HbgError) as e:
            s.set_walls(wall_x=True)
        assert e.value.status == 1 and "behind the wall" in e.value.message
        after = s.solve([1.0, 2.0], [0.0])
        for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
            assert np.array_equal(before[key], after[key]), key


def test_motions_refuse_non_finite_inputs():
    say("case: hbg_motions with a NaN mass entry, an infinite damping entry")
    with hb.Session(0) as s:
        F, H, D = 2, 1, 6
        res = {"omegas": np.array([1.0, 2.0]), "added_mass": np.ones((F, D, D)),
               "radiation_damping": np.ones((F, D, D)), "excitation_diffraction": np.ones((F, H, D), complex)}
        M = np.eye(D)
        M[2, 2] = np.nan
        with pytest.raises(hb.HbgError) as e:
            s.motions(res, M, np.eye(D))
        assert e.value.status == 1 and "mass" in e.value.message
        res["radiation_damping"][1, 0, 0] = np.inf
        with pytest.raises(hb.HbgError) as e:
            s.motions(res, np.eye(D), np.eye(D))
        assert e.value.status == 1 and "damping" in e.value.message


def test_field_and_kochin_sizes_come_from_the_last_solve():
    say("case: field and Kochin values sized by the library; a new geometry invalidates them")
    with hb.Session(0) as s:
        v = meshes.hemisphere_panels(1.0, 4, 8, quarter=False)
        s.set_environment(1025.0, 9.81)
        s.set_panels(v)
        s.set_field_points(np.array([[3.0, 0.0, 0.0], [0.0, 4.0, -0.5]]))
        s.set_kochin(16)
        s.solve([1.0, 1.5, 2.0], [0.0, 90.0])
        f = s.field_values()
        k = s.kochin_values()
        assert f["frad"].shape == (3, 2, 6) and f["ftot"].shape == (3, 2, 2)
        assert k["kr"].shape == (3, 16, 6) and k["kd"].shape == (3, 16, 2)
        with pytest.raises(ValueError):
            s.field_values(5, 2)  # the caller's counts must match the record
        # two bodies now: the old values (6 radiation columns) are gone, not re-read at 12
        s.set_panels(np.concatenate([v, meshes.place(v, (5.0, 0.0, 0.0))]), body=[0] * len(v) + [1] * len(v))
        for call in (s.field_values, s.kochin_values):
            with pytest.raises(hb.HbgError) as e:
                call()
            assert e.value.status == 4
