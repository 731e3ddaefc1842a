"""Kochin functions (GPU), infinite and finite depth. No Kochin functions in the oracle, so they are
pinned by identities of the discrete problem and of the far field:

- Haskind: the excitation the forces kernel integrates as i w rho int (n_j phi_I - phi_j
  dphi_I/dn) dS, with phi_I = (i g / w) Phi of the direction beta + pi, is rho g H_j(beta + pi)
  over the same panels and centroid values, so the two agree to rounding (the Haskind excitation
  is held to the oracle in the parity tests); a wrong sign, factor or direction fails it;
- the theta derivatives against central differences of H on a fine grid of angles;
- the far field: at k r >> 1 the potential of each solution is C(r) H(theta) with one factor C
  for every radiation column and heading, |C| = (k / 2) sqrt(2 / (pi k r)) in infinite depth
  (the far field of the wave Green function, G ~ -2 pi i k e^{k (z + zeta)} H0^(2)(k R)), here
  against the field-point elevations at 128 angles on a circle."""

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
    yield s
    s.set_kochin(0)
    s.set_field_points(None)
    s.close()


@pytest.mark.parametrize("depth", [float("inf"), 2.5])
def test_haskind_excitation_is_the_radiation_kochin_function(session, depth):
    s = session
    s.set_environment(RHO, G, depth)
    v = meshes.place(meshes.box_panels(1.0, 0.6, 0.5, (6, 4, 3), quarter=False), (0.4, -0.3, 0.0))
    s.set_panels(v, rotation_centres=np.array([[0.2, 0.1, -0.2]]))
    NK = 72
    s.set_kochin(NK)
    heads = [0.0, 45.0, 100.0]
    w = np.array([0.6, 1.4, 2.5])
    say(f"case: box {len(v)} panels off the origin, depth {depth}, {NK} Kochin angles, FP64 and mixed")
    for precision in ("fp64", "mixed"):
        out = s.solve(w, heads, precision=precision)
        kv = s.kochin_values(len(w), len(heads))
        t = [int(round((h + 180.0) / 5.0)) % NK for h in heads]
        hk = RHO * G * kv["kr"][:, t, :]  # [F][H][6]
        ref = out["excitation_haskind"]
        e = np.abs(hk - ref).max() / np.abs(ref).max()
        say(f"  {precision}: |rho g H(beta + pi) - Haskind excitation| / peak = {e:.2e}")
        assert e < 1e-12
    s.set_kochin(0)
    s.set_environment(RHO, G)


def test_theta_derivatives_against_central_differences(session):
    s = session
    s.set_environment(RHO, G, 3.0)
    v = meshes.place(meshes.hemisphere_panels(1.0, 8, 16, quarter=False), (0.5, 0.3, 0.0))
    s.set_panels(v)
    NK = 1440
    s.set_kochin(NK)
    w = np.array([0.9, 2.0])
    s.solve(w, [30.0])
    kv = s.kochin_values(len(w), 1)
    d = 2.0 * np.pi / NK
    worst = 0.0
    for a, at in (("kr", "krt"), ("kd", "kdt")):
        H, Ht = kv[a], kv[at]
        fd = (np.roll(H, -1, axis=1) - np.roll(H, 1, axis=1)) / (2.0 * d)
        e = np.abs(fd - Ht).max() / np.abs(Ht).max()
        worst = max(worst, e)
        say(f"case: hemisphere 256 panels, depth 3, {NK} angles: {at} vs central differences of {a}: {e:.2e}")
    assert worst < 1e-4
    s.set_kochin(0)
    s.set_environment(RHO, G)


def test_far_field_is_the_kochin_function(session):
    s = session
    s.set_environment(RHO, G)
    v = meshes.place(meshes.box_panels(1.0, 0.6, 0.5, (8, 4, 4), quarter=False), (0.3, 0.2, 0.0))
    s.set_panels(v)
    om = 1.5
    k = om * om / G
    r = 200.0 / k
    NK = 128
    th = 2.0 * np.pi * np.arange(NK) / NK
    s.set_kochin(NK)
    s.set_field_points(np.stack([r * np.cos(th), r * np.sin(th), np.zeros(NK)], axis=1))
    heads = [0.0, 70.0]
    s.solve([om], heads)
    kv = s.kochin_values(1, 2)
    fv = s.field_values(1, 2)
    phi_rad = fv["frad"][0] / k  # [P][6]
    phi_sca = fv["fsca"][0] / (-1j * om / G)  # [P][H]
    H = np.concatenate([kv["kr"][0], kv["kd"][0]], axis=1)  # [NK][6 + H]
    phi = np.concatenate([phi_rad, phi_sca], axis=1)
    big = np.abs(H) > 0.05 * np.abs(H).max(axis=0, keepdims=True)
    ratio = phi[big] / H[big]
    C = np.median(ratio.real) + 1j * np.median(ratio.imag)
    spread = np.abs(ratio / C - 1.0).max()
    c_theory = 0.5 * k * np.sqrt(2.0 / (np.pi * k * r))
    say(f"case: box {len(v)} panels, omega {om}, k r = 200, {NK} angles: phi / H = C with spread {spread:.2e}, "
        f"|C| / (k/2) sqrt(2 / (pi k r)) = {abs(C) / c_theory:.6f}")
    assert spread < 2e-2
    assert abs(abs(C) / c_theory - 1.0) < 1e-2
    s.set_kochin(0)
    s.set_field_points(None)
