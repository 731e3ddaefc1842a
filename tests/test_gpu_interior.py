"""The interior free surface (irregular-frequency removal; GPU):
- the reference's regression baseline lo_cylinder_interior.npz (tests/data/regression of the
  reference, copied into tests/golden/, loaded with allow_pickle=False): the quadrant cylinder of
  tests/python/regression_problems.py (radius 1, draft 0.5, 8 x 4 side and 8 x 4 bottom panels,
  symmetry planes x = 0 and y = 0) with its graded interior surface, rho 1000, g 9.80665,
  omega = sqrt(g), sqrt(2.5 g), headings 0 and 30 degrees: A, B and the excitations at the rules
  of the other regression baselines (1e-5 of each array's peak, 5e-5 of the per-entry scale;
  test_gpu_regression.py); its field-point velocities are not an output of this path;
- the quadrant with its interior surface and the two symmetry planes against the whole body with
  the whole interior surface (FP64 and mixed);
- the irregular frequency itself: at the first one of the cylinder (K a = j01 coth(j01 d / a),
  about 2.88) the added mass without the interior surface departs from the smooth curve through
  its neighbours, with it it does not; the same for the 8192-panel workload's body in depth 3
  (nu h about 7.3), coarser;
- finite depth: the interior surface on water 60 radii deep against the infinite-depth solve that
  the baseline pins.
"""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from conftest import GOLDEN
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu

RHO, G = 1000.0, 9.80665
PEAK_TOL, ENTRY_TOL = 1e-5, 5e-5
KEYS = ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_froude_krylov",
        "excitation_scattering", "excitation_haskind")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def errors(a, b):
    """(max |a - b| / max |b|, max |a - b| / per-entry scale) (test_regression.py:25-75)."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b).astype(np.complex128)
    assert a.shape == b.shape, (a.shape, b.shape)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    d = np.abs(a - b)
    return float(d.max() / peak), float((d / scale).max())


def quadrant_cylinder():
    return meshes.cylinder_panels(1.0, 0.5, n_phi=8, n_z=4, n_r=4, quarter=True)


def test_graded_interior_of_the_quadrant_cylinder_layout():
    # 8 waterline segments, rays of mean length 1 over segments of 2 sin(pi / 32): M = 6 cells a ray
    q = quadrant_cylinder()
    inner = meshes.interior_graded(q, True, True)
    assert inner.shape == (48, 4, 3) and np.all(inner[:, :, 2] == 0.0)
    n = np.cross(inner[:, 2] - inner[:, 0], inner[:, 3] - inner[:, 1])
    assert np.all(n[:, 2] > 0.0)
    area = 0.5 * n[:, 2].sum()
    assert abs(area - 4.0 * np.sin(np.pi / 16)) < 1e-12  # the polygonal quarter waterplane


def test_interior_cylinder_against_the_reference_baseline():
    ref = np.load(GOLDEN / "lo_cylinder_interior.npz", allow_pickle=False)
    q = quadrant_cylinder()
    inner = meshes.interior_graded(q, True, True)
    say(f"case: lo_cylinder_interior: quadrant {len(q)} + {len(inner)} interior panels, both planes, "
        f"omega {ref['omegas']}, headings {ref['headings']}")
    with hb.Session(0) as s:
        s.set_environment(RHO, G)
        s.set_symmetry(True, True)
        s.set_panels(q)
        s.set_interior(inner)
        res = s.solve(ref["omegas"], ref["headings"])
    for key in KEYS:
        pk, pe = errors(res[key], ref[key])
        say(f"  {key}: {pk:.2e} of the peak, {pe:.2e} of the per-entry scale")
        assert pk <= PEAK_TOL and pe <= ENTRY_TOL, (key, pk, pe)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_interior_quadrant_against_the_whole(precision):
    q = quadrant_cylinder()
    qi = meshes.interior_graded(q, True, True)
    w_body = meshes.expand_symmetry(q)
    w_inner = meshes.interior_graded(w_body)
    assert len(w_inner) == 4 * len(qi)
    w = np.sqrt(G * np.array([1.0, 2.5, 2.88]))
    heads = [0.0, 30.0]
    say(f"case: interior cylinder, quadrant with planes vs whole ({len(w_body)} + {len(w_inner)} panels), {precision}")
    with hb.Session(0) as a, hb.Session(0) as b:
        for s in (a, b):
            s.set_environment(RHO, G)
        a.set_symmetry(True, True)
        a.set_panels(q)
        a.set_interior(qi)
        b.set_panels(w_body)
        b.set_interior(w_inner)
        got, ref = a.solve(w, heads, precision=precision), b.solve(w, heads, precision=precision)
    for key in KEYS:
        pk, pe = errors(got[key], ref[key])
        say(f"  {key}: quadrant vs whole {pe:.2e}")
        assert pe < 1e-9


def test_interior_surface_removes_the_irregular_frequency():
    # first irregular frequency of the truncated cylinder (radius 1, draft 0.5): the interior
    # Dirichlet eigenvalue K = j01 coth(j01 d) with j01 the first zero of J0
    j01 = 2.404825557695773
    K_irr = j01 / np.tanh(j01 * 0.5)
    body = meshes.cylinder_panels(1.0, 0.5, n_phi=32, n_z=8, n_r=8, quarter=False, cosine=True)
    inner = meshes.interior_graded(body)
    Ks = K_irr + np.array([-0.3, 0.0, 0.3])
    w = np.sqrt(G * Ks)
    say(f"case: cylinder {len(body)} panels, added mass at K = {Ks} (irregular {K_irr:.4f}), without and "
        f"with {len(inner)} interior panels")
    with hb.Session(0) as s:
        s.set_environment(RHO, G)
        s.set_panels(body)
        plain = s.solve(w, [0.0])["added_mass"][:, 2, 2]
        s.set_interior(inner)
        removed = s.solve(w, [0.0])["added_mass"][:, 2, 2]
    # the departure of the middle value from the chord through its neighbours, of the values' size
    dev = lambda a: abs(a[1] - 0.5 * (a[0] + a[2])) / np.abs(a).max()
    say(f"  heave added mass without {plain}, with {removed}: departures {dev(plain):.2e}, {dev(removed):.2e}")
    assert dev(removed) < 0.02 and dev(plain) > 5.0 * dev(removed)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_interior_in_deep_finite_depth_against_infinite_depth(precision):
    # the finite-depth kernel (Newman form near the axis, with its own coincident-point limit) on
    # water 60 radii deep against the infinite-depth solve pinned by the reference baseline: the
    # bottom's effect there is e^{-2 K h} (below 1e-50)
    q = quadrant_cylinder()
    inner = meshes.interior_graded(q, True, True)
    w = np.sqrt(G * np.array([1.0, 2.5]))
    heads = [0.0, 30.0]
    say(f"case: interior cylinder (quadrant, planes), depth 60 vs infinite depth, {precision}")
    out = []
    for depth in (60.0, np.inf):
        with hb.Session(0) as s:
            s.set_environment(RHO, G, depth)
            s.set_symmetry(True, True)
            s.set_panels(q)
            s.set_interior(inner)
            out.append(s.solve(w, heads, precision=precision))
    for key in KEYS:
        pk, pe = errors(out[0][key], out[1][key])
        say(f"  {key}: depth 60 vs infinite {pe:.2e}")
        assert pe < 1e-6


def test_interior_removes_the_irregular_frequency_in_finite_depth():
    # the 8192-panel workload's body (radius 1, draft 1) in depth 3, coarser: its first irregular
    # frequency, K = j01 coth(j01 d) (nu h about 7.3), with and without the interior surface
    j01 = 2.404825557695773
    K_irr = j01 / np.tanh(j01 * 1.0)
    body = meshes.cylinder_panels(1.0, 1.0, 32, 8, 8, quarter=False, cosine=True)
    inner = meshes.interior_graded(body)
    Ks = K_irr + np.array([-0.3, 0.0, 0.3])
    w = np.sqrt(G * Ks)
    say(f"case: cylinder {len(body)} panels, depth 3, K = {Ks} (irregular {K_irr:.4f}, nu h {3 * K_irr:.2f}), without "
        f"and with {len(inner)} interior panels")
    with hb.Session(0) as s:
        s.set_environment(RHO, G, 3.0)
        s.set_panels(body)
        plain = s.solve(w, [0.0])["added_mass"][:, 2, 2]
        s.set_interior(inner)
        removed = s.solve(w, [0.0])["added_mass"][:, 2, 2]
    dev = lambda a: abs(a[1] - 0.5 * (a[0] + a[2])) / np.abs(a).max()
    say(f"  heave added mass without {plain}, with {removed}: departures {dev(plain):.2e}, {dev(removed):.2e}")
    assert dev(removed) < 0.02 and dev(plain) > 5.0 * dev(removed)


@pytest.mark.parametrize("depth", [float("inf"), 3.0])
def test_limits_with_the_interior_surface(depth):
    """At omega = 0 and infinity the interior surface's unknowns are not used (identity rows,
    lo_frequency.rs:99-104): the added mass is the body's alone, as without the surface; the sector
    with both planes gives the same. An ordinary frequency between them differs (the surface
    takes part)."""
    q = quadrant_cylinder()
    qi = meshes.interior_graded(q, True, True)
    body = meshes.expand_symmetry(q)
    inner = meshes.interior_graded(body)
    w = np.array([0.0, 2.0, np.inf])
    say(f"case: interior cylinder {len(body)} + {len(inner)} panels, depth {depth}, limits with and without the "
        f"surface, and as the quadrant")
    with hb.Session(0) as a, hb.Session(0) as b:
        for s in (a, b):
            s.set_environment(RHO, G, depth)
        a.set_panels(body)
        plain = a.solve(w, [])["added_mass"]
        a.set_interior(inner)
        with_surface = a.solve(w, [])["added_mass"]
        b.set_symmetry(True, True)
        b.set_panels(q)
        b.set_interior(qi)
        sector = b.solve(w, [])["added_mass"]
    for i, label in ((0, "omega 0"), (2, "omega infinity")):
        e = np.nanmax(np.abs(with_surface[i] - plain[i])) / np.nanmax(np.abs(plain[i]))
        es = np.nanmax(np.abs(sector[i] - plain[i])) / np.nanmax(np.abs(plain[i]))
        say(f"  {label}: with the surface vs without {e:.2e}, quadrant vs without {es:.2e}")
        assert np.array_equal(np.isnan(with_surface[i]), np.isnan(plain[i]))
        assert e < 1e-12 and es < 1e-9
    d = np.abs(with_surface[1] - plain[1]).max() / np.abs(plain[1]).max()
    say(f"  omega 2: with the surface vs without {d:.2e}")
    assert d > 1e-6


@pytest.mark.parametrize("depth", [float("inf"), 3.0])
def test_kochin_and_field_points_with_the_interior_surface(depth):
    """Kochin functions and field values with the interior surface, whose panels are left out of
    their sums (lo_frequency.rs:1311-1313, 1443-1458): the Haskind excitation equals rho g H(beta +
    pi) of the radiation Kochin functions to rounding (the same body panels and values), and the
    field values on a far circle are C H(theta) with the one factor of the far field."""
    body = meshes.cylinder_panels(1.0, 0.5, n_phi=32, n_z=8, n_r=8, quarter=False, cosine=True)
    inner = meshes.interior_graded(body)
    om = 2.0
    k = om * om / G if not np.isfinite(depth) else None
    NK = 72
    heads = [0.0, 45.0]
    say(f"case: cylinder {len(body)} + {len(inner)} interior panels, depth {depth}: Kochin functions and far-field "
        f"points")
    with hb.Session(0) as s:
        s.set_environment(RHO, G, depth)
        s.set_panels(body)
        s.set_interior(inner)
        s.set_kochin(NK)
        if k is not None:
            r = 200.0 / k
            th = 2.0 * np.pi * np.arange(NK) / NK
            s.set_field_points(np.stack([r * np.cos(th), r * np.sin(th), np.zeros(NK)], axis=1))
        out = s.solve([om], heads)
        kv = s.kochin_values(1, len(heads))
        t = [int(round((h + 180.0) / 5.0)) % NK for h in heads]
        hk = RHO * G * kv["kr"][:, t, :]
        ref = out["excitation_haskind"]
        e = np.abs(hk - ref).max() / np.abs(ref).max()
        say(f"  |rho g H(beta + pi) - Haskind excitation| / peak = {e:.2e}")
        assert e < 1e-12
        if k is not None:
            fv = s.field_values(1, len(heads))
            phi = np.concatenate([fv["frad"][0] / k, fv["fsca"][0] / (-1j * om / G)], axis=1)
            H = np.concatenate([kv["kr"][0], kv["kd"][0]], axis=1)
            big = np.abs(H) > 0.05 * np.abs(H).max(axis=0, keepdims=True)
            ratio = phi[big] / H[big]
            C = np.median(ratio.real) + 1j * np.median(ratio.imag)
            spread = np.abs(ratio / C - 1.0).max()
            c_theory = 0.5 * k * np.sqrt(2.0 / (np.pi * k * r))
            say(f"  far circle k r = 200: phi / H = C with spread {spread:.2e}, |C| / theory {abs(C) / c_theory:.6f}")
            assert spread < 2e-2 and abs(abs(C) / c_theory - 1.0) < 1e-2
        s.set_kochin(0)
        s.set_field_points(None)
