"""Walls x = 0, y = 0 and both (GPU): the solve in the walled domain against the open sea with the
mirror images of the body as bodies of their own, which is the same problem written out.

In the walled domain the image bodies move as the mirror image of the real one and the incident
wave is the sum of the wave at the reflected points (lo_geometry.rs:947-983, low_order.rs:
1470-1486). In the open sea, with the image body g (reflection diag(sx, sy, 1)) given its own six
rigid modes about the reflected rotation centre, the mirrored motion of mode j is s_g(j) times the
image's mode j: s = (sx, sy, 1) for the translations and det(sigma) (sx, sy, 1) for the rotations
(a reflected moment arm crossed with a reflected normal is det(sigma) sigma of the moment). So

    A_wall(i, j) = sum_g s_g(j) A_open(body 0 mode i, body g mode j), the same for B,
    X_wall(beta) = sum_g X_open(body 0, beta_g), beta_g the heading reflected by g (and FK, SC),

with every image body present as an obstacle in each open-sea problem. The open-sea path itself
is held to the FP64 oracle in the parity tests, and directly here once. The Haskind excitation of
the walled problem is checked against its diffraction excitation (Green's theorem holds with the
walls, where both potentials satisfy the wall condition) at the discretization level."""

import math
import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_per_entry(a, b):
    b = np.asarray(b)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def reflections(wx, wy):
    r = [(1.0, 1.0)]
    if wx:
        r.append((-1.0, 1.0))
    if wy:
        r.append((1.0, -1.0))
    if wx and wy:
        r.append((-1.0, -1.0))
    return r


def reflect(v, sx, sy):
    r = np.asarray(v, dtype=float) * np.array([sx, sy, 1.0])
    return r[:, [0, 3, 2, 1]] if sx * sy < 0 else r


def mirror_signs(sx, sy):
    det = sx * sy
    return np.array([sx, sy, 1.0, det * sx, det * sy, det])


def open_sea_equivalent(v, centre, wx, wy):
    """The body and its wall images as separate bodies (the images about the reflected centres)."""
    refl = reflections(wx, wy)
    verts = np.concatenate([reflect(v, sx, sy) for sx, sy in refl])
    body = np.repeat(np.arange(len(refl), dtype=np.int32), len(v))
    centres = np.array([[sx * centre[0], sy * centre[1], centre[2]] for sx, sy in refl])
    return refl, verts, body, centres


def headings_reflected(heads_deg, refl):
    out = []
    for b in np.radians(heads_deg):
        for sx, sy in refl:
            out.append(math.degrees(math.atan2(sy * math.sin(b), sx * math.cos(b))))
    return out


def wall_from_open(o, refl, H):
    G = len(refl)
    s = np.stack([mirror_signs(*r) for r in refl])  # [G][6]
    out = {}
    for key in ("added_mass", "radiation_damping"):
        a = o[key]  # [F][6G][6G]
        out[key] = sum(a[:, :6, 6 * g:6 * g + 6] * s[g][None, None, :] for g in range(G))
    for key in ("excitation_diffraction", "excitation_froude_krylov", "excitation_scattering"):
        x = o[key][:, :, :6].reshape(len(o[key]), H, G, 6)  # headings [beta][g]
        out[key] = x.sum(axis=2)
    return out


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    yield s
    s.set_walls(False, False)
    s.close()


BODIES = {
    "box": lambda: meshes.place(meshes.box_panels(1.0, 0.5, 0.5, (6, 3, 3), quarter=False), (2.0, 1.5, 0.0)),
    "hemisphere": lambda: meshes.place(meshes.hemisphere_panels(0.8, 6, 16, quarter=False), (1.5, 2.5, 0.0)),
}
CASES = [
    ("box", float("inf"), 1, 0, "fp64"),
    ("box", float("inf"), 0, 1, "mixed"),
    ("box", 3.0, 1, 1, "fp64"),
    ("hemisphere", float("inf"), 1, 1, "mixed"),
    ("hemisphere", 2.5, 1, 0, "mixed"),
    ("hemisphere", 2.5, 0, 1, "fp64"),
]


@pytest.mark.parametrize("shape,depth,wx,wy,precision", CASES)
def test_walls_against_the_open_sea_with_images(session, shape, depth, wx, wy, precision):
    v = BODIES[shape]()
    centre = np.array([v[:, :, 0].mean(), v[:, :, 1].mean(), -0.1])
    heads = [0.0, 40.0]
    w = np.array([0.6, 1.3, 2.2])
    refl, verts, body, centres = open_sea_equivalent(v, centre, wx, wy)
    say(f"case: {shape} {len(v)} panels, walls x {wx} y {wy}, depth {depth}, {precision}; open sea with "
        f"{len(refl)} bodies, {len(refl) * len(heads)} headings")
    s = session
    s.set_environment(1025.0, 9.81, depth)
    s.set_walls(False, False)
    s.set_panels(verts, body=body, rotation_centres=centres)
    o = s.solve(w, headings_reflected(heads, refl), precision=precision)
    ref = wall_from_open(o, refl, len(heads))
    s.set_panels(v, rotation_centres=centre[None])  # the real body first: the images lie behind the walls
    s.set_walls(wx, wy)
    got = s.solve(w, heads, precision=precision)
    for key, r in ref.items():
        e = rel_per_entry(got[key], r)
        say(f"  {key}: walls vs open sea with images {e:.2e}")
        assert e < 1e-9
    say(f"  Haskind vs diffraction excitation: walls "
        f"{rel_per_entry(got['excitation_haskind'], got['excitation_diffraction']):.2e}, open sea with images "
        f"{rel_per_entry(o['excitation_haskind'], o['excitation_diffraction']):.2e}")
    s.set_walls(False, False)


@pytest.mark.parametrize("shape,depth,wx,wy", [("box", float("inf"), 1, 0), ("box", 3.0, 1, 1),
                                                ("hemisphere", 2.5, 0, 1)])
def test_wall_limits_against_the_open_sea_with_images(session, shape, depth, wx, wy):
    """The added mass at omega = 0 and infinity by the walls against the open sea with the image
    bodies. In finite depth at omega = 0 the modes with a net volume flux through the real body
    are NaN on both sides (the flux of the image bodies' modes is the same)."""
    v = BODIES[shape]()
    centre = np.array([v[:, :, 0].mean(), v[:, :, 1].mean(), -0.1])
    w = np.array([0.0, np.inf])
    refl, verts, body, centres = open_sea_equivalent(v, centre, wx, wy)
    say(f"case: limits, {shape} {len(v)} panels, walls x {wx} y {wy}, depth {depth}; open sea with "
        f"{len(refl)} bodies")
    s = session
    s.set_environment(1025.0, 9.81, depth)
    s.set_walls(False, False)
    s.set_panels(verts, body=body, rotation_centres=centres)
    ref = wall_from_open(s.solve(w, []), refl, 0)["added_mass"]
    s.set_panels(v, rotation_centres=centre[None])
    s.set_walls(wx, wy)
    got = s.solve(w, [])["added_mass"]
    s.set_walls(False, False)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert nan[0, 2, :].all() == np.isfinite(depth) and not nan[1].any()
    for i, label in enumerate(("omega 0", "omega infinity")):
        e = rel_per_entry(np.nan_to_num(got[i]), np.nan_to_num(ref[i]))
        say(f"  {label}: walls vs open sea with images {e:.2e}")
        assert e < 1e-9


@pytest.mark.parametrize("wx,wy", [(1, 0), (0, 1), (1, 1)])
def test_wall_haskind_relation_converges(session, wx, wy):
    # the Haskind and the pressure excitation differ by the discretization error only: the gap
    # falls when the panels are halved
    s = session
    s.set_environment(1025.0, 9.81, float("inf"))
    w, heads = np.array([0.8, 1.5]), [40.0]
    gaps = []
    for n in ((4, 2, 2), (8, 4, 4), (16, 8, 8)):
        v = meshes.place(meshes.box_panels(1.0, 0.5, 0.5, n, quarter=False), (2.0, 1.5, 0.0))
        s.set_walls(False, False)
        s.set_panels(v, rotation_centres=np.array([[2.0, 1.5, -0.1]]))
        s.set_walls(wx, wy)
        got = s.solve(w, heads)
        X, HK = got["excitation_diffraction"], got["excitation_haskind"]
        gaps.append(float(np.abs(HK - X).max() / np.abs(X).max()))
    s.set_walls(False, False)
    say(f"case: box, walls x {wx} y {wy}, |HK - X| / max |X| at {len(v) // 16}, {len(v) // 4}, {len(v)} panels: "
        f"{gaps}")
    assert gaps[1] < 0.7 * gaps[0] and gaps[2] < 0.7 * gaps[1]


def test_walls_against_the_oracle(session, oracle):
    v = meshes.place(meshes.box_panels(0.6, 0.4, 0.4, (3, 2, 2), quarter=False), (1.0, 0.8, 0.0))
    centre = np.array([1.0, 0.8, -0.1])
    refl, verts, body, centres = open_sea_equivalent(v, centre, 1, 1)
    heads = [30.0]
    w = np.array([0.9, 1.7])
    say(f"case: box {len(v)} panels, walls x and y, FP64 vs the oracle's open sea with the images")
    orc = oracle.Problem(verts, body=body, rotation_centres=centres, rho=1025.0, grav=9.81).solve(
        w, headings_reflected(heads, refl))
    ref = wall_from_open(orc, refl, len(heads))
    s = session
    s.set_environment(1025.0, 9.81, float("inf"))
    s.set_walls(False, False)
    s.set_panels(v, rotation_centres=centre[None])
    s.set_walls(True, True)
    got = s.solve(w, heads)
    s.set_walls(False, False)
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        e = rel_per_entry(got[key], ref[key])
        say(f"  {key}: vs oracle {e:.2e}")
        assert e < 1e-10


def test_wall_refusals(session):
    s = session
    s.set_environment(1025.0, 9.81, float("inf"))
    s.set_walls(False, False)
    s.set_symmetry(False, False)
    v = meshes.place(meshes.box_panels(0.6, 0.4, 0.4, (3, 2, 2), quarter=False), (0.3, 0.8, 0.0))  # x in [-0.3, 0.9]
    say("case: walls with a body behind the wall, a symmetry plane on a wall, with the statics")
    s.set_panels(v)
    with pytest.raises(hb.HbgError) as e:
        s.set_walls(True, False)
    assert e.value.status == 1
    s.set_walls(False, True)  # y >= 0 holds: accepted
    with pytest.raises(hb.HbgError) as e:
        s.set_symmetry(False, True)  # a symmetry plane where the wall stands
    assert e.value.status == 1
    with pytest.raises(hb.HbgError) as e:
        s.statics()
    assert e.value.status == 3
    s.set_walls(False, False)


@pytest.mark.parametrize("depth,wx,wy", [(float("inf"), 1, 1), (2.5, 0, 1)])
def test_wall_field_points_and_kochin_against_the_open_sea(session, depth, wx, wy):
    v = BODIES["box"]()
    centre = np.array([v[:, :, 0].mean(), v[:, :, 1].mean(), -0.1])
    heads = [0.0, 40.0]
    w = np.array([0.8, 1.7])
    pts = np.array([[4.0, 3.0, 0.0], [0.5, 0.8, -0.3], [6.0, 0.2, -0.1]])
    NK = 24
    refl, verts, body, centres = open_sea_equivalent(v, centre, wx, wy)
    G = len(refl)
    sg = np.stack([mirror_signs(*r) for r in refl])
    say(f"case: box {len(v)} panels, walls x {wx} y {wy}, depth {depth}: field points and Kochin functions "
        f"against the open sea with {G} bodies")
    s = session
    s.set_environment(1025.0, 9.81, depth)
    s.set_walls(False, False)
    s.set_panels(verts, body=body, rotation_centres=centres)
    s.set_field_points(pts)
    s.set_kochin(NK)
    hr = headings_reflected(heads, refl)
    s.solve(w, hr)
    fo, ko = s.field_values(len(w), len(hr)), s.kochin_values(len(w), len(hr))
    s.set_panels(v, rotation_centres=centre[None])
    s.set_walls(wx, wy)
    s.solve(w, heads)
    fw, kw = s.field_values(len(w), len(heads)), s.kochin_values(len(w), len(heads))
    s.set_walls(False, False)
    s.set_field_points(None)
    s.set_kochin(0)
    F = len(w)
    ref = {"frad": sum(fo["frad"][:, :, 6 * g:6 * g + 6] * sg[g] for g in range(G)),
           "kr": sum(ko["kr"][:, :, 6 * g:6 * g + 6] * sg[g] for g in range(G)),
           "krt": sum(ko["krt"][:, :, 6 * g:6 * g + 6] * sg[g] for g in range(G))}
    for key, src in (("finc", fo), ("fsca", fo), ("ftot", fo), ("kd", ko), ("kdt", ko)):
        a = src[key]
        ref[key] = a.reshape(F, a.shape[1], len(heads), G).sum(axis=3)
    got = {**fw, **kw}
    for key, r in ref.items():
        e = rel_per_entry(got[key], r)
        say(f"  {key}: walls vs open sea with images {e:.2e}")
        assert e < 1e-9


@pytest.mark.parametrize("wx,depth,precision", [(1, float("inf"), "fp64"), (1, 2.5, "mixed"), (0, float("inf"), "mixed"),
                                                (0, 2.5, "fp64")])
def test_walls_with_a_symmetry_plane_against_the_whole_body(session, wx, depth, precision):
    # a body symmetric about the plane normal to the wall (y = 0 beside the wall x = 0, or x = 0
    # beside y = 0), stored as its half: the group of both, the two characters even under the wall
    box = meshes.box_panels(1.0, 0.5, 0.5, (6, 4, 3), quarter=False)
    if wx:
        v = meshes.place(box, (2.0, 0.0, 0.0))
        half = v[v.mean(axis=1)[:, 1] > 0.0]
        whole = np.concatenate([half, reflect(half, 1.0, -1.0)])
    else:
        v = meshes.place(box, (0.0, 2.0, 0.0))
        half = v[v.mean(axis=1)[:, 0] > 0.0]
        whole = np.concatenate([half, reflect(half, -1.0, 1.0)])
    centre = np.array([[2.0, 0.0, -0.1]]) if wx else np.array([[0.0, 2.0, -0.1]])
    heads, w = [0.0, 40.0], np.array([0.7, 1.6])
    pts = np.array([[4.0, 1.0, 0.0], [1.0, 4.0, -0.2]])
    say(f"case: box half {len(half)} of {len(whole)} panels, wall {'x' if wx else 'y'} with the symmetry plane "
        f"{'y' if wx else 'x'} = 0, depth {depth}, {precision}; field points and 24 Kochin angles")
    s = session
    s.set_environment(1025.0, 9.81, depth)
    res = []
    for v_, sym in ((half, True), (whole, False)):
        s.set_walls(False, False)
        s.set_symmetry(False, False)
        s.set_panels(v_, rotation_centres=centre)
        s.set_walls(wx, not wx)
        s.set_symmetry(sym and not wx, sym and bool(wx))
        s.set_field_points(pts)
        s.set_kochin(24)
        out = s.solve(w, heads, precision=precision)
        res.append({**out, **s.field_values(len(w), len(heads)), **s.kochin_values(len(w), len(heads))})
    s.set_symmetry(False, False)
    s.set_walls(False, False)
    s.set_field_points(None)
    s.set_kochin(0)
    for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind", "frad", "ftot",
                "kr", "kd", "kdt"):
        e = rel_per_entry(res[0][key], res[1][key])
        say(f"  {key}: half with the symmetry plane vs whole body {e:.2e}")
        # two mixed-precision solves each stop at a backward error of 1e-12
        assert e < (1e-9 if precision == "fp64" else 1e-8)


@pytest.mark.parametrize("depth,wx,wy", [(float("inf"), 1, 0), (3.0, 1, 1)])
def test_walls_with_an_interior_free_surface(session, depth, wx, wy):
    """A body with its interior free surface by the walls against the open sea with the mirror
    bodies, each with its own interior surface (the image of the real one): A, B, X at an ordinary
    frequency and at the body's first irregular frequency, and the limits."""
    body = meshes.place(meshes.cylinder_panels(1.0, 0.5, n_phi=16, n_z=4, n_r=4, quarter=False, cosine=True),
                        (2.0, 1.5, 0.0))
    inner = meshes.place(meshes.interior_graded(meshes.cylinder_panels(1.0, 0.5, n_phi=16, n_z=4, n_r=4,
                                                                       quarter=False, cosine=True)), (2.0, 1.5, 0.0))
    centre = np.array([2.0, 1.5, -0.1])
    refl, verts, bodies, centres = open_sea_equivalent(body, centre, wx, wy)
    inners = np.concatenate([reflect(inner, sx, sy) for sx, sy in refl])
    ibody = np.repeat(np.arange(len(refl), dtype=np.int32), len(inner))
    heads = [0.0, 40.0]
    w = np.sqrt(9.81 * np.array([1.0, 2.88]))
    say(f"case: cylinder {len(body)} + {len(inner)} interior panels by walls x {wx} y {wy}, depth {depth}; open sea "
        f"with {len(refl)} bodies and their interior surfaces")
    s = session
    s.set_environment(1025.0, 9.81, depth)
    s.set_walls(False, False)
    s.set_panels(verts, body=bodies, rotation_centres=centres)
    s.set_interior(inners, ibody)
    o = s.solve(w, headings_reflected(heads, refl))
    ref = wall_from_open(o, refl, len(heads))
    lim_o = wall_from_open(s.solve([0.0, np.inf], []), refl, 0)["added_mass"]
    s.set_panels(body, rotation_centres=centre[None])
    s.set_walls(wx, wy)
    s.set_interior(inner)
    got = s.solve(w, heads)
    lim = s.solve([0.0, np.inf], [])["added_mass"]
    s.set_interior(None)
    s.set_walls(False, False)
    for key, r in ref.items():
        e = rel_per_entry(got[key], r)
        say(f"  {key}: walls vs open sea with images {e:.2e}")
        assert e < 1e-9
    assert np.array_equal(np.isnan(lim), np.isnan(lim_o))
    e = rel_per_entry(np.nan_to_num(lim.reshape(-1, 6)), np.nan_to_num(lim_o.reshape(-1, 6)))
    say(f"  limits: walls vs open sea with images {e:.2e}")
    assert e < 1e-9
