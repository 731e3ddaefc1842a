"""Systems smaller than one 64-column panel of the LU, or a partial panel past one: the FP64 solve
against the FP64 oracle and the mixed-precision solve against the FP64 one, finite and infinite
depth (GPU). The blocked kernels then run with a single partial diagonal block, no trailing update
or a one-column one, and launch grids of one workgroup per system."""

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


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    yield s
    s.close()


@pytest.mark.parametrize("n_theta,n_phi", [(2, 4), (5, 12), (5, 13), (10, 13)])
@pytest.mark.parametrize("depth", [float("inf"), 2.0])
def test_small_orders_fp64_and_mixed(session, oracle, n_theta, n_phi, depth):
    v = meshes.hemisphere_panels(1.0, n_theta, n_phi, quarter=False)
    n = np.asarray(v).reshape(-1, 4, 3).shape[0]
    say(f"case: hemisphere {n} panels, depth {depth}, 3 frequencies, 2 headings, FP64 and mixed")
    w = np.array([0.6, 1.4, 2.5])
    heads = [0.0, 45.0]
    session.set_environment(1025.0, 9.81, depth)
    session.set_panels(v)
    plain = session.solve(w, heads)
    mixed = session.solve(w, heads, precision="mixed")
    orc = oracle.Problem(v, rho=1025.0, grav=9.81, depth=depth).solve(w, heads)
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        e = rel_per_entry(plain[key], orc[key])
        em = rel_per_entry(mixed[key], plain[key])
        say(f"  {key}: fp64 vs oracle {e:.2e}, mixed vs fp64 {em:.2e}")
        assert e < 1e-10
        assert em < 1e-9
