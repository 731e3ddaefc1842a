"""The precision path with the FP32 far-field assembly (HBG_PRECISION_MIXED32) on the GPU.

- the FP32 evaluators (truncated FP32 tables, FP32 arithmetic) against the FP64 ones at points:
  infinite depth over [0, 48) x (0, 48), finite depth over the Newman form and the series;
- whole solves against the plain FP64 path: A, B, X within 1e-5 of the per-entry scale (the
  reference's regression rule, tests/python/test_regression.py:25-75), on a hemisphere, a
  finite-depth cylinder, the 8192-panel finite-depth cylinder of the learned configuration at
  frequencies across its sweep (its first irregular frequency, nu h about 7.3, included) and the
  12288-panel barge of the precision configuration (omega about 1.5 rad/s, its first irregular
  frequency, included).
"""

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


def test_fp32_wave_kernel_against_fp64(session):
    say("case: FP32 infinite-depth kernel vs FP64 at 20000 points")
    rng = np.random.default_rng(11)
    X = np.concatenate([rng.uniform(0.0, 47.99, 16000), rng.uniform(0.0, 2.0, 4000)])
    Y = np.concatenate([rng.uniform(1e-3, 47.99, 16000), rng.uniform(1e-3, 2.0, 4000)])
    f32 = session.wave_kernel32(X, Y)
    f64 = session.wave_kernel(X, Y)[:, [0, 1, 3, 4]]
    # the error against each value's own size, floored at 1 (Ft is O(1) away from the origin, where
    # its log singularity is handled by the statics' log correction)
    err = np.abs(f32 - f64) / np.maximum(np.abs(f64), 1.0)
    say(f"  max error / max(|v|, 1): Ft {err[:, 0].max():.2e}, Ft_X {err[:, 1].max():.2e}, "
        f"Gi {err[:, 2].max():.2e}, Gi_X {err[:, 3].max():.2e}")
    assert err.max() < 2e-6


@pytest.mark.parametrize("nuh", [0.05, 0.7, 3.0, 7.3, 20.0])
def test_fp32_finite_kernel_against_fp64(session, nuh):
    h = 3.0
    K = nuh / h
    say(f"case: FP32 finite-depth kernel vs FP64, h = 3, nu h = {nuh}")
    rng = np.random.default_rng(int(nuh * 100))
    n = 6000
    R = h * np.exp(rng.uniform(np.log(1e-3), np.log(2.0), n))  # Newman form below R / h = 1/2, series above
    z = -h * rng.uniform(1e-3, 1.0, n)
    zeta = -h * rng.uniform(1e-3, 1.0, n)
    f32 = session.finite_kernel32(h, K, R, z, zeta)
    f64 = session.finite_kernel(h, K, R, z, zeta)
    # each quantity against its size over the sample (the regular part is O(1/h), its derivatives
    # O(1/h^2) near the axis): the reference's scale columns of green_fin_reference.txt
    scale = np.abs(f64).max(axis=0)
    err = (np.abs(f32 - f64) / scale).max(axis=0)
    say("  max error / scale: " + ", ".join(f"{e:.1e}" for e in err))
    assert err.max() < 2e-5


@pytest.mark.parametrize("depth", [np.inf, 3.0])
def test_mixed32_matrices_against_fp64(session, depth):
    say(f"case: FP32-assembled S, A vs FP64, cylinder 1024 panels, depth {depth}")
    session.set_environment(1025.0, 9.81, depth)
    session.set_panels(meshes.cylinder_panels(1.0, 1.0, 32, 16, 16, quarter=False, cosine=True))
    for w in (0.8, 2.5):
        S64, A64 = session.assemble(w)
        S32, A32 = session.assemble(w, precision="mixed32")
        eS = np.abs(S32 - S64).max() / np.abs(S64).max()
        eA = np.abs(A32 - A64).max() / np.abs(A64).max()
        d = np.abs(np.diag(A32) - np.diag(A64)).max()
        say(f"  omega {w}: max |dS| / max |S| {eS:.2e}, max |dA| / max |A| {eA:.2e}, diagonal {d:.2e}")
        assert eS < 2e-6 and eA < 2e-6 and d < 1e-12


def solve_both(session, omegas, heads):
    a = session.solve(omegas, heads, precision="fp64")
    session.set_profiling(True)
    b = session.solve(omegas, heads, precision="mixed32")
    prof = session.profile()
    session.set_profiling(False)
    errs = {k: rel_per_entry(b[k], a[k]) for k in ("added_mass", "radiation_damping", "excitation_diffraction",
                                                    "excitation_haskind")}
    say("  " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()) +
        f"; refinement steps {prof['mixed_refinement_steps']:.0f} over {prof['frequencies']:.0f} frequencies, "
        f"final backward error {prof['mixed_max_backward_error']:.1e}, FP64 fallbacks {prof['mixed_fp64_fallbacks']:.0f}")
    # the FP32 path itself ran: no frequency fell back to the FP64 solve, and the results differ
    # from the FP64 ones (the FP32 evaluation is visible in them)
    assert prof["mixed_fp64_fallbacks"] == 0 and prof["mixed_refinement_steps"] > 0
    assert max(errs.values()) > 0.0
    return errs


def test_mixed32_hemisphere_infinite_depth(session):
    say("case: mixed32 vs fp64, hemisphere 512 panels, 8 frequencies, 2 headings")
    session.set_environment(1025.0, 9.81)
    session.set_panels(meshes.hemisphere_panels(1.0, 16, 32, quarter=False))
    errs = solve_both(session, np.sqrt(9.81 * np.linspace(0.1, 3.0, 8)), [0.0, 30.0])
    assert max(errs.values()) < 1e-5


def test_mixed32_cylinder_finite_depth(session):
    say("case: mixed32 vs fp64, cylinder 1024 panels, depth 3, nu h 0.05..20")
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(meshes.cylinder_panels(1.0, 1.0, 32, 16, 16, quarter=False, cosine=True))
    nu = np.array([0.05, 0.3, 1.0, 2.4, 4.0, 7.3, 12.0, 20.0]) / 3.0
    errs = solve_both(session, np.sqrt(9.81 * nu), [0.0])
    assert max(errs.values()) < 1e-5


def test_mixed32_headline_body(session):
    say("case: mixed32 vs fp64, 8192-panel cylinder, depth 3, 6 frequencies nu h 0.05..20 (7.3: irregular)")
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(meshes.cylinder_panels(1.0, 1.0, 128, 32, 32, quarter=False, cosine=True))
    nu = np.array([0.05, 0.8, 3.0, 7.3, 12.0, 20.0]) / 3.0
    errs = solve_both(session, np.sqrt(9.81 * nu), [0.0])
    assert max(errs.values()) < 1e-5


def test_mixed32_barge_12288(session):
    say("case: mixed32 vs fp64, 12288-panel barge, infinite depth, omega 0.2, 1.0, 1.5, 2.0, 2 headings")
    session.set_environment(1025.0, 9.81)
    session.set_panels(meshes.box_panels(40.0, 10.0, 5.0, (168, 43, 12), quarter=False))
    errs = solve_both(session, [0.2, 1.0, 1.5, 2.0], [0.0, 45.0])
    assert max(errs.values()) < 1e-5
