"""Parity of the finite-depth path (GPU) with the oracle and the reference's vectors."""

import sys
from collections import defaultdict

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
    b = np.asarray(b)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def test_finite_kernel_golden(session):
    # tests/data/green_fin_reference.txt (mpmath, 30 digits), ordinary frequencies; the reference
    # allows 2e-6 of max(|v|, scale) (core.rs:155-195)
    say("case: finite-depth kernel vs green_fin_reference.txt")
    groups = defaultdict(list)
    for line in (GOLDEN / "green_fin_reference.txt").read_text().splitlines():
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        if f[0] == "0":
            groups[(float(f[1]), float(f[2]))].append([float(x) for x in f[3:]])
    worst_near = worst_far = worst_im = 0.0
    for (h, K), rows in groups.items():
        rows = np.array(rows)
        R, z, ze = rows[:, 0], rows[:, 1], rows[:, 2]
        ok = ~((R == 0) & (z + ze == 0))
        rows, R, z, ze = rows[ok], R[ok], z[ok], ze[ok]
        g = session.finite_kernel(h, K, R, z, ze)
        ref, sc, scD = rows[:, 3:11], rows[:, 11], rows[:, 12]
        den = np.maximum(np.abs(ref), np.where(np.arange(8) % 4 == 0, sc[:, None], scD[:, None]))
        e = np.abs(g - ref) / den
        near = R < 0.5 * h
        if near.any():
            worst_near = max(worst_near, e[near, :4].max())
        if (~near).any():
            worst_far = max(worst_far, e[~near, :4].max())
        worst_im = max(worst_im, e[:, 4:].max())
    say(f"  near {worst_near:.2e}, far {worst_far:.2e}, imaginary {worst_im:.2e}")
    assert worst_near < 1e-9 and worst_far < 2e-6 and worst_im < 1e-12


def test_finite_kernel_vs_oracle(session, oracle):
    say("case: finite-depth kernel vs oracle at random points, h = 3")
    rng = np.random.default_rng(3)
    h = 3.0
    for K in (0.0167, 0.5, 2.2, 6.67):
        n = 60
        R = np.concatenate([rng.uniform(0, 1.5, n // 2), rng.uniform(1.5, 6.0, n // 2)])
        z = -rng.uniform(0, h, n)
        ze = -rng.uniform(0, h, n)
        g = session.finite_kernel(h, K, R, z, ze)
        o = np.array([oracle.finite_kernel(h, K, a, b, c) for a, b, c in zip(R, z, ze)])
        scale = np.maximum(np.abs(o), 1e-3 * np.abs(o).max(axis=0))
        err = (np.abs(g - o) / scale).max()
        say(f"  K {K}: {err:.2e}")
        assert err < 1e-10


@pytest.fixture(scope="module")
def cyl_fd():
    # cylinder R = 1, draft 1, 256 panels; depth 3: pairs on both sides of R = h/2
    return meshes.cylinder_panels(1.0, 1.0, 16, 8, 8, quarter=False, cosine=True)


def test_statics_and_assembly_finite(session, oracle, cyl_fd):
    say("case: finite-depth statics and assembly, 256 panels, h = 3")
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(cyl_fd)
    P = oracle.Problem(cyl_fd, rho=1025.0, grav=9.81, depth=3.0)
    g, o = session.statics(), P.statics()
    for q in range(4):
        err = np.abs(g[:, :, q] - o[:, :, q]).max() / np.abs(o[:, :, q]).max()
        say(f"  statics component {q}: {err:.2e}")
        assert err < 1e-12
    for w in (0.5, 2.0, 6.0):
        S, A = session.assemble(w)
        So, Ao = P.assemble(w)
        eS = np.abs(S - So).max() / np.abs(So).max()
        eA = np.abs(A - Ao).max() / np.abs(Ao).max()
        say(f"  omega {w}: S {eS:.2e}, A {eA:.2e}")
        assert eS < 1e-11 and eA < 1e-11


def test_solve_finite_vs_oracle(session, oracle, cyl_fd):
    say("case: finite-depth solve vs oracle, 256 panels, h = 3")
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(cyl_fd)
    w = [0.45, 1.8, 4.0]
    out = session.solve(w, [0.0, 30.0])
    orc = oracle.Problem(cyl_fd, rho=1025.0, grav=9.81, depth=3.0).solve(w, [0.0, 30.0])
    orc["excitation_haskind"] = np.transpose(orc["excitation_haskind"], (0, 2, 1))
    for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
        e = rel_per_entry(out[key], orc[key])
        say(f"  {key}: {e:.2e}")
        assert e < 1e-10


def test_cylinder_finite_depth_eigenfunction_matching(session):
    # tests/data/cylinder_finite_depth.txt: truncated cylinder a = 1, draft 1, depths 2 and 4,
    # Ka 0.4, 0.8, 1.5 by eigenfunction matching (converged to 1e-5); the reference's A5 test
    # (test_finite_depth_and_limits.py:65-80) requires its graded meshes to converge to it at
    # O(h^p), p about 1.8
    say("case: A5 cylinder in finite depth vs eigenfunction matching")
    ref = np.loadtxt(GOLDEN / "cylinder_finite_depth.txt")
    G, RHO = 9.80665, 1000.0
    errs = {}
    for n in (12, 24):
        v = meshes.expand_symmetry(meshes.cylinder_panels(1.0, 1.0, n_phi=n, n_z=n, n_r=n // 2, quarter=True,
                                                          cosine=True))
        session.set_panels(v)
        rows = []
        for h in (2.0, 4.0):
            sel = ref[ref[:, 0] == h]
            w = np.sqrt(sel[:, 2] * G)
            session.set_environment(RHO, G, h)
            r = session.solve(w, [])
            A, B = r["added_mass"] / RHO, r["radiation_damping"] / RHO
            rows.append(np.c_[A[:, 2, 2], B[:, 2, 2] / w, A[:, 0, 0], B[:, 0, 0] / w])
        vals = np.vstack(rows)
        errs[n] = np.abs(vals - ref[:, 3:]).max(axis=0) / np.abs(ref[:, 3:]).max(axis=0)
        say(f"  n = {n} ({v.shape[0]} panels): {errs[n]}")
    order = np.log(errs[12] / errs[24]) / np.log(2.0)
    say(f"  observed order {order}")
    assert np.all(errs[24] <= 6e-3)
    assert np.all((order >= 1.4) & (order <= 2.4))
