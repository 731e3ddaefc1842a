"""The learned surrogate on the GPU: its backward pass against the FP64 CPU restatement
(oracle/mlp.py), its BF16 inference against the CPU forward with the same BF16 roundings, its
targets against the finite-depth evaluator, a short training run, and the solve with and without
it (the ablation)."""

import sys
import time

import numpy as np
import pytest
from scipy import special

import hydrobem_amd as hb
from hydrobem_amd import meshes, train_surrogate
from oracle import mlp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def session():
    s = hb.Session(0)
    yield s
    s.close()


def test_gradient_matches_cpu(session):
    say("case: surrogate gradient, batch 1000, vs FP64 CPU")
    assert hb.Session.surrogate_size() == mlp.SIZE
    p = mlp.init(11).astype(np.float32)
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, size=(1000, 4)).astype(np.float32)
    Y = rng.normal(size=(1000, 4)).astype(np.float32)
    loss, g = session.surrogate_gradient(p, X, Y)
    lref, gref = mlp.loss_and_gradient(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    err = np.abs(g - gref).max() / np.abs(gref).max()
    say(f"  loss {loss:.6e} vs {lref:.6e}; gradient max error / max |g| {err:.2e}")
    assert abs(loss - lref) <= 1e-5 * lref
    assert err < 1e-4
    # per block of parameters (each layer's weights and biases)
    i = 0
    for name, shape in mlp.SHAPES:
        n = int(np.prod(shape))
        e = np.abs(g[i:i + n] - gref[i:i + n]).max() / max(np.abs(gref[i:i + n]).max(), 1e-30)
        assert e < 1e-3, name
        i += n


def test_bf16_inference_matches_cpu(session):
    say("case: surrogate BF16 inference vs CPU forward with BF16 roundings")
    p = mlp.init(12).astype(np.float32)
    session.surrogate_set(p)
    assert np.array_equal(session.surrogate_get(), p)
    rng = np.random.default_rng(4)
    n = 1000
    pts = np.column_stack([np.exp(rng.uniform(np.log(0.01), np.log(20), n)), -rng.uniform(0, 1, n),
                           -rng.uniform(0, 1, n), np.exp(rng.uniform(np.log(0.05), np.log(20), n))])
    out = session.surrogate_eval(pts)
    u = np.column_stack([2 * (np.log(pts[:, 0]) - np.log(0.01)) / (np.log(20) - np.log(0.01)) - 1,
                         2 * pts[:, 1] + 1, 2 * pts[:, 2] + 1,
                         2 * (np.log(pts[:, 3]) - np.log(0.05)) / (np.log(20) - np.log(0.05)) - 1])
    s0 = 1 + pts[:, 3]
    s1 = s0 ** 2 / np.minimum(1.0, pts[:, 0])
    ref = mlp.forward(p.astype(np.float64), u.astype(np.float32), round_bf16=True)
    ref = ref * np.column_stack([s0, s1, s1, s1])
    err = (np.abs(out - ref) / np.abs(ref).max(0)).max()
    say(f"  max error / max |y| {err:.2e}")
    assert err < 1e-2


def test_remainder_is_the_evaluator_minus_the_propagating_mode(session):
    say("case: evanescent remainder vs the finite-depth kernel and the analytic propagating mode")
    rng = np.random.default_rng(9)
    h, n = 3.0, 400
    for nuh in (0.05, 0.7, 6.0, 20.0):
        K = nuh / h
        R = h * np.exp(rng.uniform(np.log(0.01), np.log(20), n))
        z = -h * rng.uniform(0, 1, n)
        zeta = -h * rng.uniform(0, 1, n)
        rem = session.finite_remainder(h, K, R, z, zeta)
        ker = session.finite_kernel(h, K, R, z, zeta)
        k0 = hb.dispersion_roots(K * h, 1)[0] / h
        A0 = 2 * np.pi * k0 / (-np.expm1(-4 * k0 * h) + 4 * k0 * h * np.exp(-2 * k0 * h))
        v = np.array([-(z + zeta), z - zeta + 2 * h, -z + zeta + 2 * h, z + zeta + 4 * h])
        e = np.exp(-k0 * v)
        E = e.sum(0)
        Ez = k0 * (e[0] + e[1] - e[2] - e[3])
        y0, y1 = special.y0(k0 * R), special.y1(k0 * R)
        exp_T = (ker[:, 0] + A0 * E * y0) * h
        exp_TR = (ker[:, 1] - A0 * E * k0 * y1) * h * h
        exp_Tzeta = (ker[:, 2] + A0 * Ez * y0) * h * h
        scale = np.abs(exp_T).max() + 1.0
        err = max(np.abs(rem[:, 0] - exp_T).max(), np.abs(rem[:, 1] - exp_TR).max() / h,
                  np.abs(rem[:, 2] - exp_Tzeta).max() / h) / scale
        say(f"  nu h {nuh}: {err:.2e}")
        assert err < 1e-9
        # Im G is the propagating mode alone: Wi = -A0 E J0(k0 R)
        assert np.abs(ker[:, 4] + A0 * E * special.j0(k0 * R)).max() < 1e-9 * (np.abs(ker[:, 4]).max() + 1)


def test_short_training_learns(session):
    say("case: short surrogate training (200k samples, 3 epochs)")
    rep = session.surrogate_train(n_train=200_000, n_holdout=20_000, epochs=3, batch=2048, learning_rate=2e-3, seed=7)
    say("  " + ", ".join(f"{k} {v:.3g}" for k, v in rep.items()))
    assert rep["steps"] == 3 * (200_000 // 2048)
    assert rep["nonfinite_targets"] == 0 and rep["nonfinite_params"] == 0
    # standardized targets: an untrained network sits near 1; 291 steps bring it well below
    assert np.isfinite(rep["train_loss"]) and rep["train_loss"] < 0.3
    assert rep["holdout_rms"] < 0.6
    # the held-out judge: the network against the FP64 evaluator at fresh points
    rng = np.random.default_rng(21)
    h, nuh, n = 1.0, 1.3, 300
    R = np.exp(rng.uniform(np.log(0.1), np.log(20), n))
    z, zeta = -rng.uniform(0, 1, n), -rng.uniform(0, 1, n)
    ref = session.finite_remainder(h, nuh / h, R, z, zeta)
    got = session.surrogate_eval(np.column_stack([R / h, z / h, zeta / h, np.full(n, nuh)]))
    err = np.abs(got - ref).max(0) / (np.abs(ref).max(0) + 1e-12)
    trained = session.surrogate_get().copy()
    session.surrogate_set(mlp.init(7).astype(np.float32))
    got0 = session.surrogate_eval(np.column_stack([R / h, z / h, zeta / h, np.full(n, nuh)]))
    session.surrogate_set(trained)
    err0 = np.abs(got0 - ref).max(0) / (np.abs(ref).max(0) + 1e-12)
    rms = np.sqrt(((got - ref) ** 2).mean(0) / (ref ** 2).mean(0))
    rms0 = np.sqrt(((got0 - ref) ** 2).mean(0) / (ref ** 2).mean(0))
    say(f"  relative max error per output at nu h {nuh}: {err}; an untrained network: {err0}")
    say(f"  relative RMS error per output: {rms}; an untrained network: {rms0}")
    # 291 steps: the outputs together far closer to the evaluator than an untrained network's, in
    # the RMS the loss measures (measured 0.14-0.61 per output against 0.37-3.7 untrained)
    assert np.sum(rms ** 2) < 0.25 * np.sum(rms0 ** 2)
    assert rms.max() < 0.8


@pytest.mark.skipif(not train_surrogate.WEIGHTS.exists(), reason="no trained surrogate shipped")
def test_shipped_network_held_out(session):
    """The shipped network against the FP64 evaluator at fresh points of its domain: the held-out
    judge of the training, at three depths and frequencies."""
    say("case: shipped surrogate, held-out points against the FP64 evaluator")
    train_surrogate.load(session)
    rng = np.random.default_rng(21)
    for h, nuh in ((1.0, 1.3), (3.0, 0.1), (3.0, 10.0)):
        n = 2000
        R = h * np.exp(rng.uniform(np.log(0.1), np.log(20), n))
        z, zeta = -h * rng.uniform(0, 1, n), -h * rng.uniform(0, 1, n)
        ref = session.finite_remainder(h, nuh / h, R, z, zeta)
        got = session.surrogate_eval(np.column_stack([R / h, z / h, zeta / h, np.full(n, nuh)]))
        err = np.abs(got - ref).max(0) / np.abs(ref).max(0)
        say(f"  h {h}, nu h {nuh}: max error / max |output| {err}")
        # measured on MI355X: 0.5-7 % per output (BF16 weights and activations), the floor that
        # keeps the surrogate's A, B, X about a per cent off the evaluator's
        assert np.all(err < 0.1)


def _reflect(v, sx, sy):
    r = np.asarray(v, dtype=float) * np.array([sx, sy, 1.0])
    return r[:, [0, 3, 2, 1]] if sx * sy < 0 else r


@pytest.mark.skipif(not train_surrogate.WEIGHTS.exists(), reason="no trained surrogate shipped")
@pytest.mark.parametrize("sym_x,sym_y", [(1, 1), (0, 1)])
def test_surrogate_on_a_sector_against_the_whole_mesh(sym_x, sym_y):
    """The surrogate over the image blocks of a symmetric body (the network and its fallback queue
    in every block) against the whole mesh. With the network's lower R/h bound above every pair
    all pairs take the fallback queue, and the sector gives the evaluator's results of the whole
    mesh to the solver's rounding (the queue's image blocks, statics and free term). With the
    network on, a pair (k, g n) of the sector is the entry (k, m) of the whole mesh, which the
    whole mesh may take as the mirror of its pair (m, k): the network at (z_m, z_k) instead of
    (z_k, z_m), equal only to the network's own error, so the two differ by less than either
    differs from the evaluator."""
    q = meshes.cylinder_panels(1.0, 0.8, 12, 5, 4, quarter=True)
    if sym_x and sym_y:
        sector, whole = q, meshes.expand_symmetry(q)
    else:  # the plane y = 0: the half y >= 0, the quadrant and its mirror in x = 0
        sector = np.concatenate([q, _reflect(q, -1.0, 1.0)])
        whole = np.concatenate([sector, _reflect(sector, 1.0, -1.0)])
    say(f"case: surrogate on a sector of {len(sector)} of {len(whole)} panels, planes x {sym_x} y {sym_y}, depth 3")
    a, b = hb.Session(0), hb.Session(0)
    try:
        for s in (a, b):
            train_surrogate.load(s)
            s.set_environment(1025.0, 9.81, 3.0)
        a.set_symmetry(sym_x, sym_y)
        a.set_panels(sector)
        b.set_panels(whole)
        w, heads = np.array([1.0, 2.0]), [0.0, 30.0]
        plain = b.solve(w, heads)
        a.set_tuning(a.TUNE_NN_RMIN, 100)  # R/h >= 1: no pair of this body (R/h <= 0.67) to the network
        fb = a.solve(w, heads, surrogate=True)
        a.set_tuning(a.TUNE_NN_RMIN, 10)
        got = a.solve(w, heads, surrogate=True)
        ref = b.solve(w, heads, surrogate=True)
        for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
            def rel(x, y):
                return np.abs(x[key] - y[key]).max() / np.abs(y[key]).max()
            e0, e, d, ds = rel(fb, plain), rel(got, ref), rel(ref, plain), rel(got, plain)
            say(f"  {key}: fallback only vs evaluator {e0:.2e}; with the network sector vs whole {e:.2e}, "
                f"whole vs evaluator {d:.2e}, sector vs evaluator {ds:.2e}")
            assert e0 < 1e-9
            assert e < 0.5 * min(d, ds) and d > 1e-5 and ds > 1e-5
    finally:
        a.close()
        b.close()
