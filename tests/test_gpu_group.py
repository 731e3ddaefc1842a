"""The group solve (hbg_solve_group): geometries of uneven panel and body counts, and of their own
environments, solved through one batched LU (GPU), give the results of their own solves."""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    b = np.asarray(b)
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300))


def geometries():
    hemi = meshes.hemisphere_panels(1.0, 10, 20, quarter=False)
    cyl = meshes.cylinder_panels(1.0, 1.0, 22, 7, 7, quarter=False, cosine=True)
    small = meshes.cylinder_panels(0.75, 1.0, 14, 5, 5, quarter=False, cosine=True)
    shift = np.array([4.0, 0.0, 0.0])
    pair = np.concatenate([small, small + shift])
    body = np.repeat([0, 1], len(small)).astype(np.int32)
    centres = np.array([[0.0, 0.0, 0.0], shift])
    # (vertices, body, rotation centres, depth)
    return [(hemi, None, None, np.inf), (cyl, None, None, 3.0), (pair, body, centres, np.inf)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_solve_matches_separate_solves(precision):
    sessions = []
    for v, body, centres, depth in geometries():
        s = hb.Session(0)
        s.set_environment(1025.0, 9.81, depth)
        s.set_panels(v, body, centres)
        sessions.append(s)
    orders = [s.n_panels for s in sessions]
    say(f"case: group solve {precision}, panels {orders}, bodies {[s.n_bodies for s in sessions]}")
    assert len(set(orders)) == len(orders) and any(n % 64 for n in orders)
    w = np.array([0.6, 1.3, 2.2])
    heads = [0.0, 30.0]
    # 9 systems in batches of at most 4, sorted by order: 308 and 280 share a batch (the ratio
    # 0.91 is above the group solve's 0.8), 200 starts its own
    sessions[0].set_batch(4)
    group = hb.solve_group(sessions, w, heads, precision=precision)
    tol = 1e-11 if precision == "fp64" else 1e-9
    for s, g in zip(sessions, group):
        ref = s.solve(w, heads)
        for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
            assert g[key].shape == ref[key].shape
            e = rel(g[key], ref[key])
            say(f"  N {s.n_panels} {key}: {e:.2e}")
            assert e < tol
    if precision == "mixed":  # the FP32 assembly's precision runs as the mixed one in a group
        again = hb.solve_group(sessions, w, heads, precision="mixed32")
        for g, a in zip(group, again):
            for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
                assert np.array_equal(g[key], a[key]), key
    for s in sessions:
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_of_hull_variants_is_bit_identical(precision):
    """A design sweep: six hull variants of one panel count, two frequencies each. Alone, each is a
    batch of two systems; together, one batch of twelve. A system's arithmetic does not depend on
    the batch it is in, so the results are the same bits."""
    sessions = []
    for R in np.linspace(0.8, 1.3, 6):
        s = hb.Session(0)
        s.set_environment(1025.0, 9.81)
        s.set_panels(meshes.cylinder_panels(R, 1.0, 22, 7, 7, quarter=False, cosine=True))
        sessions.append(s)
    say(f"case: group of 6 hull variants, {sessions[0].n_panels} panels each, {precision}")
    w = np.array([0.9, 1.7])
    group = hb.solve_group(sessions, w, [0.0], precision=precision)
    for s, g in zip(sessions, group):
        ref = s.solve(w, [0.0], precision=precision)
        for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
            assert np.array_equal(g[key], ref[key]), (key, rel(g[key], ref[key]))
    for s in sessions:
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_with_symmetry_planes_and_walls(precision):
    """Sessions with symmetry planes (their character systems kept in one batch) and with walls
    (one even character) beside a plain geometry: each gives the results of its own solve."""
    quadrant = meshes.cylinder_panels(1.0, 0.8, 12, 5, 4, quarter=True)
    q = meshes.hemisphere_panels(1.0, 8, 6, quarter=True)
    # the half x >= 0 of the symmetry plane x = 0: the quadrant and its mirror image in y = 0 (the
    # vertex order reversed, so the normals stay in the fluid)
    half = np.concatenate([q, (q * np.array([1.0, -1.0, 1.0]))[:, [0, 3, 2, 1]]])
    box = meshes.place(meshes.box_panels(1.0, 0.5, 0.5, (6, 3, 3), quarter=False), (2.0, 1.5, 0.0))
    plain = meshes.hemisphere_panels(0.9, 8, 14, quarter=False)
    specs = [  # (vertices, symmetry, walls, depth, rotation centre)
        (quadrant, (1, 1), (0, 0), 3.0, [0.0, 0.0, -0.1]),
        (half, (1, 0), (0, 0), np.inf, [0.0, 0.0, -0.2]),
        (box, (0, 0), (1, 1), 3.0, [2.0, 1.5, -0.1]),
        (plain, (0, 0), (0, 0), np.inf, [0.0, 0.0, 0.0]),
    ]
    sessions = []
    for v, sym, walls, depth, c in specs:
        s = hb.Session(0)
        s.set_environment(1025.0, 9.81, depth)
        s.set_symmetry(*sym)
        s.set_panels(v, rotation_centres=np.array([c]))
        s.set_walls(*walls)
        sessions.append(s)
    say(f"case: group of a quadrant (both planes), a half (plane x), a box by two walls and a plain "
        f"hemisphere, panels {[s.n_panels for s in sessions]}, {precision}")
    w = np.array([0.7, 1.6])
    heads = [0.0, 35.0]
    sessions[0].set_batch(6)  # a batch of six slots: the quadrant's four characters never split
    group = hb.solve_group(sessions, w, heads, precision=precision)
    tol = 1e-11 if precision == "fp64" else 1e-9
    for s, g in zip(sessions, group):
        ref = s.solve(w, heads, precision=precision)
        for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
            e = rel(g[key], ref[key])
            say(f"  N {s.n_panels} {key}: {e:.2e}")
            assert e < tol
    for s in sessions:
        s.set_walls(False, False)
        s.set_symmetry(False, False)
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_group_with_an_interior_free_surface(precision):
    """A session with an interior free surface (extra unknowns, free term -4 pi, K S columns) in a
    group beside plain geometries: each gives the results of its own solve, the frequency
    K = 2.88 (the cylinder's first irregular frequency) among them."""
    body = meshes.cylinder_panels(1.0, 0.5, n_phi=16, n_z=4, n_r=4, quarter=False, cosine=True)
    inner = meshes.interior_graded(body)
    sessions = []
    for v, interior in ((body, inner), (body, None), (meshes.hemisphere_panels(1.0, 8, 16, quarter=False), None)):
        s = hb.Session(0)
        s.set_environment(1025.0, 9.81)
        s.set_panels(v)
        if interior is not None:
            s.set_interior(interior)
        sessions.append(s)
    say(f"case: group with an interior surface ({len(body)} + {len(inner)} panels) beside two plain geometries, "
        f"{precision}")
    w = np.sqrt(9.81 * np.array([1.0, 2.88]))
    heads = [0.0, 30.0]
    group = hb.solve_group(sessions, w, heads, precision=precision)
    tol = 1e-11 if precision == "fp64" else 1e-9
    for s, g in zip(sessions, group):
        ref = s.solve(w, heads, precision=precision)
        for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
            e = rel(g[key], ref[key])
            say(f"  N {s.n_panels} {key}: {e:.2e}")
            assert e < tol
    assert rel(group[0]["added_mass"], group[1]["added_mass"]) > 1e-4  # the surface took part
    for s in sessions:
        s.close()
