"""Field points (GPU): the elevations of lo_frequency.rs:1392-1495 at points in the fluid, on the
free surface and inside the body, infinite and finite depth. No field points in the oracle, so they are
pinned by what the discrete problem and the physics fix:

- the limit at the body: a point at eps along the normal from a panel centroid sees the discrete
  potential of that panel (the jump of its own dipole integral supplies the other 2 pi), so
  -rho sum_k n_i(k) area_k frad_j(x_k) / K reproduces A_ij - i B_ij / w of the forces kernel to
  O(eps / panel size); a dropped term, a wrong sign or scale of D_x or S_x breaks it;
- the incident elevation is the closed form Z(z) e^{-i k (x cos b + y sin b)};
- extinction: the total wave of the scattered formulation vanishes inside the body (Green's
  identity of the incident wave over the body's interior, its free-surface part zero), to the
  discretization error, which falls when the panels are halved;
- energy: the power radiated through a far circle, rho g c_g r / 2 int |eta|^2 d theta with
  c_g the group velocity, equals w^2 B_jj / 2 per unit amplitude of motion, to the far-field
  asymptotics' O(1 / k r) (in finite depth the evanescent modes are gone there)."""

import math
import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu
RHO, G = 1025.0, 9.81


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    s.set_environment(RHO, G)
    yield s
    s.set_field_points(None)
    s.close()


def box_geometry(v):
    """Centroids, outward normals (into the fluid) and areas of the rectangular panels of a box."""
    v = np.asarray(v)
    c = v.mean(axis=1)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 1])
    a = np.linalg.norm(n, axis=1)
    return c, n / a[:, None], a


def wavenumber(om, depth):
    K = om * om / G
    if not np.isfinite(depth):
        return K
    k = max(K, math.sqrt(K / depth))
    for _ in range(100):  # Newton on k tanh(k h) = K
        t = math.tanh(k * depth)
        k -= (k * t - K) / (t + k * depth * (1.0 - t * t))
    return k


def depth_factor(k, z, depth):
    if not np.isfinite(depth):
        return np.exp(k * z)
    return np.cosh(k * (z + depth)) / np.cosh(k * depth)


@pytest.mark.parametrize("depth", [float("inf"), 2.0])
def test_body_limit_reproduces_added_mass_and_damping(session, depth):
    s = session
    s.set_environment(RHO, G, depth)
    v = meshes.box_panels(1.0, 0.6, 0.5, (8, 4, 4), quarter=False)
    c, n, a = box_geometry(v)
    size = np.sqrt(a).max()
    eps = 1e-7 * size
    pts = c + eps * n
    rc = np.array([0.1, -0.05, -0.2])
    s.set_panels(v, rotation_centres=rc[None])
    s.set_field_points(pts)
    w = np.array([0.7, 1.5, 2.6])
    say(f"case: box {len(v)} panels, depth {depth}, field points at 1e-7 panel sizes off every centroid, "
        f"3 frequencies")
    out = s.solve(w, [0.0, 30.0])
    fv = s.field_values(len(w), 2)
    nj = np.concatenate([n, np.cross(c - rc, n)], axis=1)  # [N][6]
    for f, om in enumerate(w):
        K = om * om / G
        phi = fv["frad"][f] / K  # [N][6]
        z = -RHO * (nj * a[:, None]).T @ phi  # [6][6] = A - i B / w
        ref = out["added_mass"][f] - 1j * out["radiation_damping"][f] / om
        e = np.abs(z - ref).max() / np.abs(ref).max()
        say(f"  omega {om}: |(A - iB/w) from the field values - the forces| / peak = {e:.2e}")
        assert e < 1e-5
        k = wavenumber(om, depth)
        x = pts[:, 0][:, None] * np.cos(np.radians([0.0, 30.0]))[None] + pts[:, 1][:, None] * np.sin(
            np.radians([0.0, 30.0]))[None]
        inc = depth_factor(k, pts[:, 2], depth)[:, None] * np.exp(-1j * k * x)
        assert np.abs(fv["finc"][f] - inc).max() < 1e-11
        assert np.abs(fv["ftot"][f] - fv["finc"][f] - fv["fsca"][f]).max() < 1e-13
    s.set_environment(RHO, G)


@pytest.mark.parametrize("depth", [float("inf"), 2.0])
def test_total_wave_vanishes_inside_the_body(session, depth):
    s = session
    s.set_environment(RHO, G, depth)
    errs = []
    inside = np.array([[0.0, 0.0, -0.3], [0.3, 0.2, -0.5], [-0.4, -0.3, -0.2]])
    for nt, nph in ((8, 16), (16, 32)):
        v = meshes.hemisphere_panels(1.0, nt, nph, quarter=False)
        s.set_panels(v)
        s.set_field_points(inside)
        s.solve([1.2, 2.0], [0.0, 60.0])
        fv = s.field_values(2, 2)
        errs.append(float(np.abs(fv["ftot"]).max() / np.abs(fv["finc"]).max()))
    say(f"case: hemisphere 128 and 512 panels, depth {depth}, |total elevation| / |incident| inside the body: "
        f"{errs}")
    assert errs[1] < 0.5 * errs[0] and errs[1] < 2e-2
    s.set_environment(RHO, G)


@pytest.mark.parametrize("depth", [float("inf"), 3.0])
def test_radiated_energy_through_a_far_circle_is_the_damping(session, depth):
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    s = session
    s.set_environment(RHO, G, depth)
    s.set_panels(v)
    om = 1.6
    k = wavenumber(om, depth)
    r = 150.0 / k
    th = np.linspace(0.0, 2.0 * np.pi, 128, endpoint=False)
    s.set_field_points(np.stack([r * np.cos(th), r * np.sin(th), np.zeros_like(th)], axis=1))
    out = s.solve([om], [])
    fv = s.field_values(1, 0)
    kh = k * depth if np.isfinite(depth) else np.inf
    cg = om / k / 2.0 * (1.0 + (2.0 * kh / np.sinh(2.0 * kh) if np.isfinite(kh) else 0.0))
    say(f"case: hemisphere 512 panels, depth {depth}, omega 1.6, 128 points on a circle at k r = 150")
    for j in (0, 2):
        flux = RHO * G * cg * r / 2.0 * np.sum(np.abs(fv["frad"][0, :, j]) ** 2) * (2.0 * np.pi / len(th))
        B = out["radiation_damping"][0, j, j]
        e = abs(flux - om * om * B / 2.0) / (om * om * B / 2.0)
        say(f"  mode {j + 1}: radiated power {flux:.6e}, w^2 B / 2 {om * om * B / 2:.6e}, rel {e:.2e}")
        assert e < 2e-2
    s.set_field_points(None)
    s.set_environment(RHO, G)


def test_field_point_refusals(session):
    s = session
    s.set_panels(meshes.hemisphere_panels(1.0, 4, 8, quarter=False))
    with pytest.raises(hb.HbgError) as e:
        s.set_field_points(np.array([[0.0, 0.0, 0.1]]))
    assert e.value.status == 1
    s.set_field_points(np.array([[3.0, 0.0, 0.0]]))
    say("case: field points at the frequency limits, below the sea floor")
    try:
        with pytest.raises(hb.HbgError) as e:
            s.solve([0.0], [])
        assert e.value.status == 3
        s.set_environment(RHO, G, 3.0)
        s.set_field_points(np.array([[3.0, 0.0, -3.5]]))
        with pytest.raises(hb.HbgError) as e:
            s.solve([1.0], [])
        assert e.value.status == 1
    finally:
        s.set_environment(RHO, G)
        s.set_field_points(None)
