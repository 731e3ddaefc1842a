"""The radiation limits omega = 0 and omega = infinity in finite depth (GPU): the limit kernels
against the reference's mpmath rows (tests/data/green_fin_reference.txt modes 1 and 2) and the
oracle, the limit solves against the oracle's, the net-volume-flux rule at omega = 0, the limits
approached by small and large frequencies, and deep water against infinite depth."""

import math
import sys
from pathlib import Path

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes
from oracle import oracle

pytestmark = pytest.mark.gpu
RHO, G = 1000.0, 9.81
GOLDEN = Path(__file__).resolve().parent / "golden"


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    yield s
    s.close()


def limit_rows():
    rows = []
    for line in (GOLDEN / "green_fin_reference.txt").read_text().splitlines():
        f = line.split()
        if f and f[0] in ("1", "2"):
            rows.append([float(x) for x in f])
    return np.array(rows)


def test_limit_kernels_against_mpmath_and_the_oracle(session):
    d = limit_rows()
    for mode in (1, 2):
        r = d[d[:, 0] == mode]
        h = r[0, 1]
        assert np.all(r[:, 1] == h)
        K = 0.0 if mode == 1 else math.inf
        got = session.finite_kernel(h, K, r[:, 3], r[:, 4], r[:, 5])
        ref = r[:, 6:14]
        sc, scD = r[:, 14], r[:, 15]
        err = np.abs(got[:, :4] - ref[:, :4]) / np.column_stack([sc, scD, scD, scD])
        near = r[:, 3] < 0.5 * h
        say(f"case: limit kernel mode {mode}, {len(r)} mpmath rows: near {err[near].max():.2e}, "
            f"series {err[~near].max():.2e}")
        # the image sums below R = h/2; the series with the reference's stop at k_n R >= 18 above
        assert err[near].max() < 1e-11
        assert err[~near].max() < 2e-6
        assert np.all(got[:, 4:] == 0.0)
        # the oracle at random points, both forms (same algorithm, independent code)
        rng = np.random.default_rng(mode)
        n = 400
        hh = 3.0
        R = hh * np.exp(rng.uniform(np.log(1e-3), np.log(8.0), n))
        z, ze = -hh * rng.uniform(0, 1, n), -hh * rng.uniform(0, 1, n)
        g = session.finite_kernel(hh, K, R, z, ze)
        o = np.array([oracle.finite_limit_kernel(mode, hh, R[i], z[i], ze[i]) for i in range(n)])
        e = np.abs(g - o).max(axis=0) / (np.abs(o).max(axis=0) + 1e-300)
        say(f"  against the oracle at {n} random points (depth 3): {e[:4]}")
        assert e[:4].max() < 1e-9


def body(n_theta=6, n_phi=24):
    return meshes.hemisphere_panels(1.0, n_theta, n_phi, quarter=False)


def test_limit_solves_against_the_oracle(session):
    v = body()
    depth = 3.0
    session.set_environment(RHO, G, depth)
    session.set_panels(v)
    out = session.solve([0.0, math.inf], [0.0])
    ref = oracle.Problem(v, rho=RHO, grav=G, depth=depth).solve([0.0, math.inf], [0.0])
    a, ar = out["added_mass"], ref["added_mass"]
    ok = np.ones((6, 6), bool)
    ok[2, :] = ok[:, 2] = False
    e0 = np.abs(a[0][ok] - ar[0][ok]).max() / np.abs(ar[0][ok]).max()
    ei = np.abs(a[1] - ar[1]).max() / np.abs(ar[1]).max()
    say(f"case: hemisphere {len(v)} panels, depth {depth}: omega = 0 {e0:.2e} (heave NaN), infinity {ei:.2e}")
    assert e0 < 1e-10 and ei < 1e-10
    # omega = 0: heave has a net volume flux, its row and column are not defined
    assert np.all(np.isnan(a[0][2, :])) and np.all(np.isnan(a[0][:, 2]))
    assert np.all(np.isfinite(a[0][ok]))
    for key in ("radiation_damping", "excitation_diffraction", "excitation_haskind"):
        assert np.all(out[key] == 0.0)


def test_limits_approached_by_small_and_large_frequencies(session):
    # omega -> 0: A(w) - A(0) = O((kh)^2) in finite depth for the modes without net flux (the
    # reference's test: 1.1e-7 of A(0) at w = 1e-3); omega -> inf: the free-surface effect decays
    # with K on a submerged box (no waterline, no irregular frequencies)
    depth = 3.0
    session.set_environment(RHO, G, depth)
    session.set_panels(body(8, 32))
    out = session.solve([0.0, 1e-3, 1e-2], [])
    a = out["added_mass"]
    d = np.abs(a[1:, 0, 0] - a[0, 0, 0]) / a[0, 0, 0]
    say(f"case: hemisphere 256 panels, depth 3, surge |A(w) - A(0)| / A(0) at w = 1e-3, 1e-2: {d}")
    assert d[0] < d[1] and d[0] < 1e-6
    box = meshes.box_panels(1.0, 0.5, 0.5, (8, 4, 4), quarter=False) + np.array([0.0, 0.0, -1.0])
    session.set_panels(box)
    ws = [1.0, 2.0, 4.0, 6.0]
    out = session.solve(ws + [math.inf], [])
    ai = out["added_mass"][-1]
    err = [np.abs(out["added_mass"][f] - ai).max() / np.abs(ai).max() for f in range(len(ws))]
    say(f"case: submerged box {len(box)} panels, depth 3, |A(w) - A(inf)| / peak at w = {ws}: {np.round(err, 4)}")
    assert all(e1 < e0 for e0, e1 in zip(err, err[1:])) and err[-1] < 0.05


def test_deep_water_limits_tend_to_infinite_depth(session):
    v = body()
    session.set_environment(RHO, G)
    session.set_panels(v)
    inf_depth = session.solve([0.0, math.inf], [])["added_mass"]
    session.set_environment(RHO, G, 400.0)
    session.set_panels(v)
    deep = session.solve([0.0, math.inf], [])["added_mass"]
    ok = np.ones((6, 6), bool)
    ok[2, :] = ok[:, 2] = False
    e0 = np.abs(deep[0][ok] - inf_depth[0][ok]).max() / np.abs(inf_depth[0][ok]).max()
    ei = np.abs(deep[1] - inf_depth[1]).max() / np.abs(inf_depth[1]).max()
    say(f"case: hemisphere {len(v)} panels, depth 400 against infinite depth: omega = 0 {e0:.2e}, inf {ei:.2e}")
    assert e0 < 1e-4 and ei < 1e-6
