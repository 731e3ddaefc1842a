"""Parity of the HIP library with the FP64 oracle on the same inputs (GPU).

The bar: the plain FP64 path reproduces the oracle's A, B, X (and the intermediate statics,
matrices and kernel values) to 1e-10 relative; the reference's own regression baseline to its
1e-5. Everything calls through the C ABI (ctypes)."""

import math
import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from conftest import GOLDEN
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    yield s
    s.close()


def rel_per_entry(a, b):
    """max |a - b| / scale, scale = each entry's largest magnitude over the first axis, at least 1e-3
    of the peak (test_regression.py:53-56)."""
    b = np.asarray(b)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def test_wave_kernel_golden_and_oracle(session, oracle):
    say("case: wave kernel vs green_inf_reference.txt and oracle")
    d = np.loadtxt(GOLDEN / "green_inf_reference.txt")
    d = d[(d[:, 0] > 0) | (d[:, 1] > 0)]
    out = session.wave_kernel(d[:, 0], d[:, 1])
    R0 = np.hypot(d[:, 0], d[:, 1])
    errF = np.abs(out[:, 0] - d[:, 2]) / np.maximum(np.abs(d[:, 2]), 2e-3 / R0)
    errFX = np.abs(out[:, 1] - d[:, 3]) / np.maximum(np.abs(d[:, 3]), 1e-3)
    say(f"  golden: Ft {errF.max():.2e}, Ft_X {errFX.max():.2e}")
    assert errF.max() < 1e-9 and errFX.max() < 1e-8
    rng = np.random.default_rng(7)
    X = np.concatenate([rng.uniform(0, 60, 300), rng.uniform(0, 3, 200), [0.0, 47.9, 48.0, 48.1, 10.0]])
    Y = np.concatenate([rng.uniform(0, 60, 300), rng.uniform(0, 3, 200), [0.5, 3.0, 3.0, 3.0, 49.0]])
    out = session.wave_kernel(X, Y)
    ref = np.array([oracle.wave_kernel(x, y) for x, y in zip(X, Y)])
    scale = np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max(axis=0))
    err = (np.abs(out - ref) / scale).max(axis=0)
    say(f"  vs oracle per output: {err}")
    assert err.max() < 1e-9


@pytest.fixture(scope="module")
def cylinder():
    return meshes.expand_symmetry(meshes.cylinder_panels(1.0, 0.5, n_phi=8, n_z=4, n_r=4, quarter=True))


def test_statics(session, oracle, cylinder):
    say("case: statics, cylinder 256 panels")
    session.set_environment(1000.0, 9.80665)
    session.set_panels(cylinder)
    g = session.statics()
    o = oracle.Problem(cylinder, rho=1000.0, grav=9.80665).statics()
    for q in range(4):
        scale = np.abs(o[:, :, q]).max()
        err = np.abs(g[:, :, q] - o[:, :, q]).max() / scale
        say(f"  component {q}: {err:.2e}")
        assert err < 1e-12


def test_assembly(session, oracle, cylinder):
    say("case: assembly, cylinder 256 panels, omega 1.7")
    session.set_environment(1000.0, 9.80665)
    session.set_panels(cylinder)
    S, A = session.assemble(1.7)
    So, Ao = oracle.Problem(cylinder, rho=1000.0, grav=9.80665).assemble(1.7)
    eS = np.abs(S - So).max() / np.abs(So).max()
    eA = np.abs(A - Ao).max() / np.abs(Ao).max()
    say(f"  S {eS:.2e}, A {eA:.2e}")
    assert eS < 1e-12 and eA < 1e-12


@pytest.mark.parametrize("n", [64, 300, 700, 1100])
def test_lu_known_answer_and_random(session, n):
    say(f"case: lu n={n}")
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    t = (i * 7 + j * 13).astype(float)
    a = 0.1 * (np.sin(0.3 * t) + 1j * np.cos(0.17 * t))
    a[np.arange(n), np.arange(n)] += n * 0.2 + 1.0 + 0.5j
    x = 1.0 + 1j * np.arange(n) / n
    sol = session.lu_solve(a, (a @ x).reshape(n, 1))
    assert np.abs(sol[:, 0] - x).max() < 1e-10  # linalg.rs:459
    rng = np.random.default_rng(n)
    m = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = rng.normal(size=(n, 5)) + 1j * rng.normal(size=(n, 5))
    xs = session.lu_solve(m, b)
    ref = np.linalg.solve(m, b)
    be = np.abs(m @ xs - b).max() / (np.abs(m).max() * np.abs(xs).max() * n)
    say(f"  random: backward {be:.2e}, forward {np.abs(xs - ref).max() / np.abs(ref).max():.2e}")
    assert be < 1e-14
    assert np.abs(xs - ref).max() < 1e-8 * np.abs(ref).max()


def test_lu_many_rhs_and_late_pivoting(session):
    # diagonally dominant except rows 200..259 (swapped in pairs), so only the fourth panel needs
    # interchanges; 45 right-hand sides (more than one chunk of the back substitution)
    say("case: lu n=260, 45 rhs, interchanges in the fourth panel only")
    n, r = 260, 45
    rng = np.random.default_rng(7)
    a = 0.05 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    a[np.arange(n), np.arange(n)] += 4.0 + 1.0j
    perm = np.arange(n)
    perm[200:260] = np.arange(200, 260).reshape(-1, 2)[:, ::-1].reshape(-1)
    a = a[perm]
    x = rng.normal(size=(n, r)) + 1j * rng.normal(size=(n, r))
    sol = session.lu_solve(a, a @ x)
    err = np.abs(sol - x).max() / np.abs(x).max()
    say(f"  forward error {err:.2e}")
    assert err < 1e-12


def test_lu_singular_reports_zero_pivot(session):
    say("case: lu singular")
    a = np.zeros((3, 3), complex)
    a[0, 0] = 1.0
    with pytest.raises(hb.HbgError) as e:
        session.lu_solve(a, np.ones((3, 1), complex))
    assert e.value.status == 8 and "zero pivot" in e.value.message


def test_lu_overflow_to_nan_reports_numerical(session):
    """Finite entries whose elimination overflows: the panel needs interchanges (|a10| > |a00|),
    the pivoted elimination turns every remaining row into NaN by its fourth step, and the pivot
    search must still choose a row (NaN counts as the largest magnitude) so that the solve ends
    with a numerical failure, not with an interchange taken from no row."""
    say("case: lu n=64, elimination overflowing to NaN in the tournament")
    n = 64
    a = np.eye(n, dtype=complex)
    a[1:, 0] = 2.0
    a[1, 1:] = 1e308
    a[2:, 1:] = -1e308
    a[np.arange(2, n), np.arange(2, n)] = -1e308
    with pytest.raises(hb.HbgError) as e:
        session.lu_solve(a, np.ones((n, 1), complex))
    assert e.value.status == 8
    # the device and the workspace serve the next solve
    rng = np.random.default_rng(3)
    b = np.eye(n) * 4.0 + 0.1 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    x = rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))
    assert np.abs(session.lu_solve(b, b @ x) - x).max() < 1e-12


def test_solve_regression_baseline_and_oracle(session, oracle, cylinder):
    say("case: solve lo_cylinder_drift (256 panels) vs baseline and oracle")
    ref = np.load(GOLDEN / "lo_cylinder_drift.npz", allow_pickle=False)
    session.set_environment(1000.0, 9.80665)
    session.set_panels(cylinder)
    out = session.solve(ref["omegas"], ref["headings"])
    orc = oracle.Problem(cylinder, rho=1000.0, grav=9.80665).solve(ref["omegas"], ref["headings"])
    orc["excitation_haskind"] = np.transpose(orc["excitation_haskind"], (0, 2, 1))
    for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_froude_krylov",
                "excitation_scattering", "excitation_haskind"):
        e_base = rel_per_entry(out[key], ref[key])
        e_orc = rel_per_entry(out[key], orc[key])
        say(f"  {key}: baseline {e_base:.2e}, oracle {e_orc:.2e}")
        assert e_base < 1e-5 and e_orc < 1e-10


def test_solve_hemisphere_512(session, oracle):
    # the CPU-runnable configuration: hemisphere R = 1, 32 x 16 panels, rho 1025, g 9.81,
    # 16 kR in 0.1..3.0 (omega = sqrt(g k)), heading 0, 6 radiation + 1 diffraction
    say("case: solve hemisphere 512 panels, 16 frequencies")
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    w = np.sqrt(9.81 * np.linspace(0.1, 3.0, 16))
    session.set_environment(1025.0, 9.81)
    session.set_panels(v)
    out = session.solve(w, [0.0])
    orc = oracle.Problem(v, rho=1025.0, grav=9.81).solve(w, [0.0])
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        e = rel_per_entry(out[key], orc[key])
        say(f"  {key}: {e:.2e}")
        assert e < 1e-10


def test_solve_multi_leaf_tournament(session, oracle):
    # 1536 panels: six tournament leaves in the first panels of the LU
    say("case: solve cylinder 1536 panels (multi-leaf tournament pivoting)")
    v = meshes.cylinder_panels(1.0, 1.0, 48, 16, 16, quarter=False, cosine=True)
    session.set_environment(1025.0, 9.81)
    session.set_panels(v)
    w = [0.8, 2.2]
    out = session.solve(w, [0.0, 45.0])
    orc = oracle.Problem(v, rho=1025.0, grav=9.81).solve(w, [0.0, 45.0])
    for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
        o = orc[key] if key != "excitation_haskind" else np.transpose(orc[key], (0, 2, 1))
        e = rel_per_entry(out[key], o)
        say(f"  {key}: {e:.2e}")
        assert e < 1e-10


def test_batched_frequencies_match_one_at_a_time(session, oracle):
    # frequencies in flight share the LU launches; each system's arithmetic is the same
    say("case: batch 1 vs batch 3 on 5 frequencies, hemisphere 512 panels")
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    w = np.sqrt(9.81 * np.linspace(0.2, 2.8, 5))
    session.set_environment(1025.0, 9.81)
    session.set_panels(v)
    session.set_batch(1)
    one = session.solve(w, [0.0, 30.0])
    session.set_batch(3)
    three = session.solve(w, [0.0, 30.0])
    session.set_batch(0)
    for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
        assert np.array_equal(one[key], three[key]), key


def test_progress_and_cancel(session):
    say("case: progress callback and cancel")
    v = meshes.hemisphere_panels(1.0, 6, 24, quarter=False)
    session.set_environment(1025.0, 9.81)
    session.set_panels(v)
    seen = []
    session.solve([1.0, 2.0], [0.0], progress=lambda st, d, t: seen.append((st, d, t)) or 1)
    assert (0, 0, 1) in seen and (2, 1, 2) in seen
    with pytest.raises(hb.HbgError) as e:
        session.solve([1.0, 2.0], [0.0], progress=lambda st, d, t: 0 if st == 2 else 1)
    assert e.value.status == 5
