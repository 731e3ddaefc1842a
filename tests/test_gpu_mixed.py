"""The mixed-precision solve (complex64 LU on the matrix cores + FP64 iterative refinement) against
the FP64 oracle and the plain GPU path (GPU). The refinement stops at a normwise backward error of
1e-12, so the forces agree with the FP64 results far inside the 1e-5 the mixed path must meet."""

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


def test_mixed_hemisphere_vs_oracle(session, oracle):
    say("case: mixed precision, hemisphere 512 panels, 6 frequencies")
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    w = np.sqrt(9.81 * np.linspace(0.1, 3.0, 6))
    session.set_environment(1025.0, 9.81)
    session.set_panels(v)
    session.set_profiling(True)
    out = session.solve(w, [0.0, 60.0], precision="mixed")
    prof = session.profile()
    session.set_profiling(False)
    orc = oracle.Problem(v, rho=1025.0, grav=9.81).solve(w, [0.0, 60.0])
    say(f"  refinement steps {prof['mixed_refinement_steps']:.0f}, backward error "
        f"{prof['mixed_max_backward_error']:.2e}, FP64 fallbacks {prof['mixed_fp64_fallbacks']:.0f}")
    assert prof["mixed_fp64_fallbacks"] == 0 and prof["mixed_refinement_steps"] >= 6
    assert prof["mixed_max_backward_error"] <= 1e-12
    for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
        o = orc[key] if key != "excitation_haskind" else np.transpose(orc[key], (0, 2, 1))
        e = rel_per_entry(out[key], o)
        say(f"  {key}: {e:.2e}")
        assert e < 1e-9


def test_mixed_matches_plain_cylinder(session):
    say("case: mixed vs plain, cylinder 2304 panels, finite depth 3 m, 3 frequencies")
    v = meshes.cylinder_panels(1.0, 1.0, 48, 24, 24, quarter=False, cosine=True)
    w = np.sqrt(9.81 * np.array([0.05, 1.0, 6.0]) / 3.0)
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(v)
    plain = session.solve(w, [0.0])
    mixed = session.solve(w, [0.0], precision="mixed")
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        e = rel_per_entry(mixed[key], plain[key])
        say(f"  {key}: {e:.2e}")
        assert e < 1e-9


@pytest.mark.parametrize("variant", [3, 4, 5, 6, 7])
def test_mixed_split_fp16_gemm(session, variant):
    """The complex64 LU's GEMM on the FP16 matrix cores with split operands: the refined results
    stay those of the plain path, in at most two refinement steps per frequency and without an
    FP64 fallback."""
    say(f"case: mixed with GEMM variant {variant} (split FP16), cylinder 2304 panels, finite depth, "
        "4 frequencies, 4 headings")
    v = meshes.cylinder_panels(1.0, 1.0, 48, 24, 24, quarter=False, cosine=True)
    w = np.sqrt(9.81 * np.array([0.05, 0.7, 4.0, 15.0]) / 3.0)
    heads = [0.0, 45.0, 90.0, 135.0]
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(v)
    plain = session.solve(w, heads)
    old = session.get_tuning(session.TUNE_LU_GEMM_MIXED)
    session.set_tuning(session.TUNE_LU_GEMM_MIXED, variant)
    try:
        for graphs in (1, 0):  # replayed graph, then launched one by one with stage timing
            session.set_tuning(session.TUNE_LU_GRAPHS, graphs)
            session.set_profiling(graphs == 0)
            mixed = session.solve(w, heads, precision="mixed")
            if graphs == 0:
                prof = session.profile()
                say(f"  refinement steps {prof['mixed_refinement_steps']:.0f}, backward error "
                    f"{prof['mixed_max_backward_error']:.2e} (first solution "
                    f"{prof['mixed_max_first_backward_error']:.2e}), FP64 fallbacks {prof['mixed_fp64_fallbacks']:.0f}")
                assert prof["mixed_fp64_fallbacks"] == 0
                # the complex64 LU itself is FP32-accurate (normwise): a wrong product in the GEMM
                # that the refinement would still repair shows here
                assert 0.0 < prof["mixed_max_first_backward_error"] < 1e-5
                assert prof["mixed_refinement_steps"] <= 2 * len(w)
                assert prof["mixed_max_backward_error"] <= 1e-12
            for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
                e = rel_per_entry(mixed[key], plain[key])
                say(f"  graphs {graphs} {key}: {e:.2e}")
                assert e < 1e-9
    finally:
        session.set_profiling(False)
        session.set_tuning(session.TUNE_LU_GRAPHS, 1)
        session.set_tuning(session.TUNE_LU_GEMM_MIXED, old)


@pytest.mark.parametrize("depth,sub,sub2", [(4, 2, 0), (8, 3, 0), (32, 8, 0), (16, 8, 3), (36, 16, 4)])
def test_lu_sub_blocks(session, depth, sub, sub2):
    """Sub-blocks inside the LU's blocks (the in-block updates from the panel's sub-block, an
    L-shaped update at each sub-block's end): FP64 to round-off of the plain blocked LU, and the
    complex64 LU (split-FP16 GEMM) FP32-accurate before its refinement."""
    say(f"case: LU depth {depth}, sub-blocks of {sub} and {sub2} panels, cylinder 2304 panels, finite depth, "
        "3 frequencies")
    v = meshes.cylinder_panels(1.0, 1.0, 48, 24, 24, quarter=False, cosine=True)
    w = np.sqrt(9.81 * np.array([0.05, 1.0, 6.0]) / 3.0)
    session.set_environment(1025.0, 9.81, 3.0)
    session.set_panels(v)
    keys = ("added_mass", "radiation_damping", "excitation_diffraction")
    old = [session.get_tuning(k) for k in (session.TUNE_LU_DEPTH, session.TUNE_LU_DEPTH_MIXED,
                                            session.TUNE_LU_GEMM_MIXED)]
    plain = session.solve(w, [0.0])
    try:
        session.set_tuning(session.TUNE_LU_DEPTH, depth)
        session.set_tuning(session.TUNE_LU_SUB, sub)
        session.set_tuning(session.TUNE_LU_SUB2, sub2)
        assert session.get_tuning(session.TUNE_LU_SUB_MIXED) == sub
        assert session.get_tuning(session.TUNE_LU_SUB2_MIXED) == sub2
        sb = session.solve(w, [0.0])
        for key in keys:
            e = rel_per_entry(sb[key], plain[key])
            say(f"  fp64 {key}: {e:.2e}")
            assert e < 1e-11
        session.set_tuning(session.TUNE_LU_GEMM_MIXED, 6)
        session.set_profiling(True)
        mixed = session.solve(w, [0.0], precision="mixed32")
        prof = session.profile()
        say(f"  mixed32: first-solution backward error {prof['mixed_max_first_backward_error']:.2e}, "
            f"refinement steps {prof['mixed_refinement_steps']:.0f}")
        assert 0.0 < prof["mixed_max_first_backward_error"] < 1e-5
        assert prof["mixed_fp64_fallbacks"] == 0
        for key in keys:
            e = rel_per_entry(mixed[key], plain[key])
            say(f"  mixed32 {key}: {e:.2e}")
            assert e < 1e-5
    finally:
        session.set_profiling(False)
        session.set_tuning(session.TUNE_LU_SUB, 0)
        session.set_tuning(session.TUNE_LU_SUB2, 0)
        session.set_tuning(session.TUNE_LU_DEPTH, old[0])
        session.set_tuning(session.TUNE_LU_DEPTH_MIXED, old[1])
        session.set_tuning(session.TUNE_LU_GEMM_MIXED, old[2])
