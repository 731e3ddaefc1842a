"""The batched LU replayed from captured hipGraphs against the same launches made one by one, over
several replays with other GPU work between them; the autotuner's sweep (GPU)."""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_graph_replays_match_eager():
    say("case: LU hipGraph replays vs eager launches, hemisphere 512, 5 frequencies in batches of 3")
    s = hb.Session(0)
    v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
    s.set_environment(1025.0, 9.81)
    s.set_panels(v)
    s.set_batch(3)
    w = np.sqrt(9.81 * np.linspace(0.3, 2.7, 5))
    s.set_tuning(s.TUNE_LU_GRAPHS, 0)
    eager = s.solve(w, [0.0, 45.0])
    s.set_tuning(s.TUNE_LU_GRAPHS, 1)
    for rep in range(3):
        g = s.solve(w, [0.0, 45.0])
        for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
            assert np.array_equal(g[key], eager[key]), (rep, key)
        # other work on the GPU between the replays
        s.set_tuning(s.TUNE_LU_GRAPHS, 0)
        s.solve(w[:2], [10.0])
        s.set_tuning(s.TUNE_LU_GRAPHS, 1)
    mixed_e = None
    s.set_tuning(s.TUNE_LU_GRAPHS, 0)
    mixed_e = s.solve(w, [0.0], precision="mixed")
    s.set_tuning(s.TUNE_LU_GRAPHS, 1)
    for rep in range(2):
        mixed_g = s.solve(w, [0.0], precision="mixed")
        for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
            assert np.array_equal(mixed_g[key], mixed_e[key]), (rep, key)
    s.close()


def test_autotune_sweep_and_solve(oracle):
    say("case: autotune n = 1024, 7 rhs, FP64 and mixed; solve afterwards vs oracle")
    s = hb.Session(0)
    rows = s.autotune(1024, 7, "fp64")
    say("  " + "; ".join(f"v{int(r[0])} d{int(r[1])} s{int(r[2])}/{int(r[3])} b{int(r[4])} {r[5]:.3f} ms"
                         for r in rows))
    # two variants x five depths at batch 48, six nested (depth, sub-blocks) at the fastest
    # variant, then batches 4..32 at the fastest of them
    assert len(rows) == 2 * 5 + 6 + 4 and all(r[5] > 0 for r in rows)
    assert {int(r[1]) for r in rows} == {2, 4, 8, 12, 16, 24, 32, 64}
    assert {int(r[2]) for r in rows} == {0, 8, 12, 16, 32} and {int(r[3]) for r in rows} == {0, 8, 16}
    best = min(rows, key=lambda r: r[5])
    assert s.get_tuning(s.TUNE_LU_GEMM) == int(best[0]) and s.get_tuning(s.TUNE_LU_DEPTH) == int(best[1])
    assert s.get_tuning(s.TUNE_LU_SUB) == int(best[2]) and s.get_tuning(s.TUNE_LU_SUB2) == int(best[3])
    rows32 = s.autotune(1024, 7, "mixed")
    say("  " + "; ".join(f"v{int(r[0])} d{int(r[1])} s{int(r[2])}/{int(r[3])} b{int(r[4])} {r[5]:.3f} ms"
                         for r in rows32))
    assert len(rows32) == 7 * 5 + 6 + 4 and all(r[5] > 0 for r in rows32)
    best32 = min(rows32, key=lambda r: r[5])
    assert s.get_tuning(s.TUNE_LU_GEMM_MIXED) == int(best32[0])
    assert s.get_tuning(s.TUNE_LU_DEPTH_MIXED) == int(best32[1])
    assert s.get_tuning(s.TUNE_LU_SUB_MIXED) == int(best32[2])
    assert s.get_tuning(s.TUNE_LU_SUB2_MIXED) == int(best32[3])
    # the complex128 settings stay those of its own sweep
    assert s.get_tuning(s.TUNE_LU_GEMM) == int(best[0]) and s.get_tuning(s.TUNE_LU_DEPTH) == int(best[1])
    assert s.get_tuning(s.TUNE_LU_SUB) == int(best[2]) and s.get_tuning(s.TUNE_LU_SUB2) == int(best[3])
    v = meshes.cylinder_panels(1.0, 1.0, 32, 16, 16, quarter=False, cosine=True)
    s.set_environment(1025.0, 9.81)
    s.set_panels(v)
    w = [0.7, 1.9]
    out = s.solve(w, [0.0])
    orc = oracle.Problem(v, rho=1025.0, grav=9.81).solve(w, [0.0])
    for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
        e = np.abs(out[key] - orc[key]).max() / np.abs(orc[key]).max()
        say(f"  {key}: {e:.2e}")
        assert e < 1e-10
    s.set_tuning(s.TUNE_LU_GEMM, 1)
    s.set_tuning(s.TUNE_LU_DEPTH, 4)
    s.set_tuning(s.TUNE_LU_SUB, 0)
    s.set_tuning(s.TUNE_LU_SUB2, 0)
    s.close()


def test_graph_replays_after_sessions_are_freed():
    """Sessions freed and made again get their buffers at the same addresses, so a new session's
    LU meets graphs captured for the old one's matrices: the cache replays a graph only within the
    epoch of device frees it was captured in, and the results equal the eager launches'."""
    say("case: three sessions solved, freed and made again (same sizes), graphs vs eager launches")
    geoms = [meshes.hemisphere_panels(1.0, 10, 20, quarter=False),
             meshes.cylinder_panels(1.0, 1.0, 22, 7, 7, quarter=False, cosine=True),
             meshes.hemisphere_panels(0.8, 8, 16, quarter=False)]
    w, heads = np.array([0.6, 1.3, 2.2]), [0.0, 30.0]

    def make():
        out = []
        for v in geoms:
            s = hb.Session(0)
            s.set_environment(1025.0, 9.81)
            s.set_panels(v)
            out.append(s)
        return out

    for _ in range(2):
        ss = make()
        for s in ss:
            s.solve(w, heads)
        for s in ss:
            s.close()
    ss = make()
    graphs = [s.solve(w, heads) for s in ss]
    ss[0].set_tuning(ss[0].TUNE_LU_GRAPHS, 0)
    eager = [s.solve(w, heads) for s in ss]
    ss[0].set_tuning(ss[0].TUNE_LU_GRAPHS, 1)
    for g, e in zip(graphs, eager):
        for key in ("added_mass", "radiation_damping", "excitation_diffraction"):
            assert np.array_equal(g[key], e[key]), key
    for s in ss:
        s.close()
