"""The radiation limits omega = 0 and omega = infinity in infinite depth (GPU): the floating
hemisphere's surge and heave added mass against the exact sphere series (tests/sphere_limits.py)
on two meshes (discretization error O(h^2): it must fall with the mesh), the limit approached by
small frequencies, limits interleaved with ordinary frequencies. Finite depth:
tests/test_gpu_limits_finite.py."""

import math
import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes
from sphere_limits import hemisphere_added_mass

pytestmark = pytest.mark.gpu
RHO, G = 1000.0, 9.81


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    s.set_environment(RHO, G)
    yield s
    s.close()


def test_hemisphere_limits_against_the_exact_series(session):
    exact = np.array([hemisphere_added_mass(m, lim, terms=480, points=800)
                      for lim in ("zero", "infinity") for m in ("surge", "heave")])
    m0 = RHO * 2.0 * math.pi / 3.0
    errs = []
    for n_theta, n_phi in ((12, 48), (24, 96)):
        v = meshes.hemisphere_panels(1.0, n_theta, n_phi, quarter=False)
        session.set_panels(v)
        out = session.solve([0.0, math.inf], [])
        a0, ai = out["added_mass"][0] / m0, out["added_mass"][1] / m0
        got = np.array([a0[0, 0], a0[2, 2], ai[0, 0], ai[2, 2]])
        e = np.abs(got - exact) / np.abs(exact).max()
        say(f"case: hemisphere {n_theta * n_phi} panels, limits: got {np.round(got, 5)}, exact {np.round(exact, 5)}, "
            f"errors {e}")
        assert np.all(out["radiation_damping"] == 0.0)
        errs.append(e)
    assert np.all(errs[1] < errs[0])  # O(h^2): the finer mesh is closer in every entry
    assert errs[1].max() < 3e-3


def test_limit_approached_by_small_frequencies(session):
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    session.set_panels(v)
    w = np.sqrt(np.array([1e-4, 1e-3]) * G)  # K a = 1e-4, 1e-3
    out = session.solve([0.0, *w], [])
    a = out["added_mass"]
    d = np.abs(a[1:] - a[0][None]).max(axis=(1, 2)) / np.abs(a[0]).max()
    say(f"case: hemisphere 512 panels, |A(K) - A(0)| / peak at K a = 1e-4, 1e-3: {d}")
    # heave A(K) - A(0) = O(K a log K a) in infinite depth
    assert d[0] < d[1] < 1e-2


def test_limits_interleaved_with_ordinary_frequencies(session):
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    session.set_panels(v)
    ordinary = session.solve([1.0, 2.0], [0.0, 90.0], precision="mixed")
    limits = session.solve([math.inf, 0.0], [0.0, 90.0])
    both = session.solve([1.0, 0.0, math.inf, 2.0], [0.0, 90.0], precision="mixed")
    say("case: hemisphere 512 panels, omegas 1, 0, inf, 2 in one mixed-precision solve")
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        assert np.array_equal(both[key][[0, 3]], ordinary[key])
    assert np.array_equal(both["added_mass"][1], limits["added_mass"][1])
    assert np.array_equal(both["added_mass"][2], limits["added_mass"][0])
    assert np.all(both["excitation_diffraction"][[1, 2]] == 0.0)
