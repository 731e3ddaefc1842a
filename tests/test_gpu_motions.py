"""The motions (RAOs) of hbg_motions against numpy's LAPACK solve of the same systems (GPU)."""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def reference(result, M, C, Bx):
    w = result["omegas"]
    out = np.zeros_like(result["excitation_diffraction"])
    for f, om in enumerate(w):
        Z = C - om ** 2 * (M + result["added_mass"][f]) + 1j * om * (result["radiation_damping"][f] + Bx)
        out[f] = np.linalg.solve(Z, result["excitation_diffraction"][f].T).T
    return out


def test_motions_floating_hemisphere():
    say("case: RAOs of a floating hemisphere (512 panels), 8 frequencies, 2 headings")
    s = hb.Session(0)
    rho, g, R = 1025.0, 9.81, 1.0
    v = meshes.hemisphere_panels(R, 16, 32, quarter=False)
    s.set_environment(rho, g)
    s.set_panels(v)
    res = s.solve(np.sqrt(g * np.linspace(0.2, 2.5, 8)), [0.0, 30.0])
    m = rho * 2.0 / 3.0 * np.pi * R ** 3
    M = np.diag([m, m, m, 0.4 * m * R ** 2, 0.4 * m * R ** 2, 0.4 * m * R ** 2])
    C = np.zeros((6, 6))
    C[2, 2] = rho * g * np.pi * R ** 2
    C[3, 3] = C[4, 4] = rho * g * np.pi * R ** 4 / 4 * 0.3
    Bx = np.diag([0, 0, 0, 0, 0, 50.0]) + 1e3 * np.eye(6)
    rao = s.motions(res, M, C, Bx)
    ref = reference(res, M, C, Bx)
    err = np.abs(rao - ref).max() / np.abs(ref).max()
    say(f"  max error / max |xi| {err:.2e}; heave RAO {np.abs(rao[:, 0, 2]).round(3)}")
    assert err < 1e-12
    s.close()


def test_motions_random_systems_and_zero_pivot():
    say("case: RAOs of random 18-mode systems; a singular system")
    s = hb.Session(0)
    rng = np.random.default_rng(3)
    F, H, D = 5, 3, 18
    res = {"omegas": rng.uniform(0.3, 2.0, F), "added_mass": rng.normal(size=(F, D, D)),
           "radiation_damping": rng.normal(size=(F, D, D)),
           "excitation_diffraction": rng.normal(size=(F, H, D)) + 1j * rng.normal(size=(F, H, D))}
    M, C, Bx = rng.normal(size=(D, D)), rng.normal(size=(D, D)) + 5 * np.eye(D), np.zeros((D, D))
    rao = s.motions(res, M, C, Bx)
    ref = reference(res, M, C, Bx)
    err = np.abs(rao - ref).max() / np.abs(ref).max()
    say(f"  max error / max |xi| {err:.2e}")
    assert err < 1e-10
    res0 = {k: (np.zeros_like(v) if k != "omegas" else v) for k, v in res.items()}
    with pytest.raises(hb.HbgError) as e:
        s.motions(res0, np.zeros((D, D)), np.zeros((D, D)))
    assert e.value.status == 8 and "motion equation at omega" in e.value.message and "zero pivot 1" in e.value.message
    s.close()


def test_motions_with_more_headings_than_one_launch():
    say("case: RAOs of random 12-mode systems, 4 frequencies, 70 headings (groups of 32)")
    s = hb.Session(0)
    rng = np.random.default_rng(11)
    F, H, D = 4, 70, 12
    res = {"omegas": rng.uniform(0.3, 2.0, F), "added_mass": rng.normal(size=(F, D, D)),
           "radiation_damping": rng.normal(size=(F, D, D)),
           "excitation_diffraction": rng.normal(size=(F, H, D)) + 1j * rng.normal(size=(F, H, D))}
    M, C, Bx = rng.normal(size=(D, D)), rng.normal(size=(D, D)) + 5 * np.eye(D), rng.normal(size=(D, D))
    rao = s.motions(res, M, C, Bx)
    ref = reference(res, M, C, Bx)
    err = np.abs(rao - ref).max() / np.abs(ref).max()
    say(f"  max error / max |xi| {err:.2e}")
    assert err < 1e-10
    # each heading's motions do not depend on the other headings of its launch
    part = {k: (v[:, 33:34] if k == "excitation_diffraction" else v) for k, v in res.items()}
    one = s.motions(part, M, C, Bx)
    assert np.array_equal(one[:, 0], rao[:, 33])
    s.close()
