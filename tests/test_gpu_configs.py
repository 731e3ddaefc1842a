"""The benchmark configurations at their own sizes (GPU):
- the plain FP64 configuration: the 4096-panel cylinder (radius 1, draft 1, side 64 x 32
  cosine-spaced, bottom 64 x 32 rings, infinite depth) against the oracle's solve at two of its
  frequencies and its four headings (tests/golden/c2_cylinder4096_oracle.npz, written by
  tests/golden/write_c2_oracle.py), and its quadrant with the two symmetry planes against the
  whole mesh;
- the precision configuration's barge (80 x 20 m, draft 5 m, infinite depth) as a quadrant with
  the two symmetry planes against the whole body, FP64 and mixed, at frequencies around its first
  irregular frequency (omega about 1.5 rad/s). The 12288-panel mesh of that configuration has 43
  panels across the beam, so no quadrant of it exists; the mesh here is the nearest one that has
  one (168 x 44 bottom panels, 12 down the sides: 12480 panels), and the 12288-panel mesh itself
  is solved against the FP64 path in tests/test_gpu_mixed32.py;
- the batching configuration: the cylinders of 1536, 2816 and 6144 panels and their array (10496
  panels, three bodies) solved as one group (their systems sharing the batched LU) against their
  own solves.
"""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from conftest import GOLDEN
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu

KEYS = ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_froude_krylov",
        "excitation_scattering", "excitation_haskind")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_per_entry(a, b):
    """max |a - b| / scale, scale = each entry's largest magnitude over the first axis, at least 1e-3
    of the peak (test_regression.py:53-56)."""
    b = np.asarray(b)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def test_cylinder4096_fp64_against_the_oracle():
    ref = np.load(GOLDEN / "c2_cylinder4096_oracle.npz", allow_pickle=False)
    w, heads = ref["omegas"], ref["headings"]
    say(f"case: cylinder 4096 panels, infinite depth, FP64 vs the oracle, omega {w}, headings {heads}")
    with hb.Session(0) as s:
        s.set_environment(1025.0, 9.81)
        s.set_panels(meshes.cylinder_panels(1.0, 1.0, 64, 32, 32, quarter=False, cosine=True))
        assert s.n_panels == 4096
        got = s.solve(w, heads)
    for key in KEYS:
        o = ref[key] if key != "excitation_haskind" else np.transpose(ref[key], (0, 2, 1))
        e = rel_per_entry(got[key], o)
        say(f"  {key}: {e:.2e}")
        assert e < 1e-10


def test_cylinder4096_quadrant_against_the_whole():
    w = np.array([0.3, 1.35, 3.0])
    heads = [0.0, 45.0, 90.0, 135.0]
    say("case: cylinder 4096 panels as its quadrant (1024) with both planes vs the whole mesh, FP64")
    with hb.Session(0) as q, hb.Session(0) as f:
        for s in (q, f):
            s.set_environment(1025.0, 9.81)
        # n_phi counts the divisions of the mesh's own arc: 16 over the quadrant = 64 over the circle
        quadrant = meshes.cylinder_panels(1.0, 1.0, 16, 32, 32, quarter=True, cosine=True)
        q.set_symmetry(True, True)
        q.set_panels(quadrant)
        f.set_panels(meshes.cylinder_panels(1.0, 1.0, 64, 32, 32, quarter=False, cosine=True))
        assert (len(quadrant), f.n_panels) == (1024, 4096)
        got, ref = q.solve(w, heads), f.solve(w, heads)
    for key in KEYS:
        e = rel_per_entry(got[key], ref[key])
        say(f"  {key}: quadrant vs whole {e:.2e}")
        assert e < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_barge_quadrant_against_the_whole(precision):
    quadrant = meshes.box_panels(40.0, 10.0, 5.0, (84, 22, 12), quarter=True)
    whole = meshes.expand_symmetry(quadrant)
    w = np.array([1.0, 1.45, 1.5, 1.55])
    heads = [0.0, 45.0]
    say(f"case: barge 80 x 20 x 5, quadrant {len(quadrant)} of {len(whole)} panels, both planes, {precision}, "
        f"omega {w}")
    with hb.Session(0) as q, hb.Session(0) as f:
        for s in (q, f):
            s.set_environment(1025.0, 9.81)
        q.set_symmetry(True, True)
        q.set_panels(quadrant)
        f.set_panels(whole)
        got, ref = q.solve(w, heads, precision=precision), f.solve(w, heads, precision=precision)
    for key in KEYS:
        e = rel_per_entry(got[key], ref[key])
        say(f"  {key}: quadrant vs whole {e:.2e}")
        assert e < 1e-9


def test_batching_geometries_group_against_separate_solves():
    bodies = [meshes.cylinder_panels(R, 1.0, *d, quarter=False, cosine=True)
              for R, d in ((1.0, (48, 16, 16)), (0.75, (64, 22, 22)), (1.5, (96, 32, 32)))]
    shifts = [np.array([4.0 * i, 0.0, 0.0]) for i in range(3)]
    array = np.concatenate([b + sft for b, sft in zip(bodies, shifts)])
    body = np.concatenate([np.full(len(b), i, np.int32) for i, b in enumerate(bodies)])
    geoms = [(b, None, None) for b in bodies] + [(array, body, np.array(shifts))]
    w = np.sqrt(9.81 * np.array([0.05, 1.3, 4.0]))
    heads = [0.0]
    sessions = []
    try:
        for v, bd, rc in geoms:
            s = hb.Session(0)
            s.set_environment(1025.0, 9.81)
            s.set_panels(v, bd, rc)
            sessions.append(s)
        say(f"case: batching geometries, panels {[s.n_panels for s in sessions]}, group vs separate, mixed")
        assert [s.n_panels for s in sessions] == [1536, 2816, 6144, 10496]
        group = hb.solve_group(sessions, w, heads, precision="mixed")
        for s, g in zip(sessions, group):
            ref = s.solve(w, heads, precision="mixed")
            for key in ("added_mass", "radiation_damping", "excitation_diffraction", "excitation_haskind"):
                assert g[key].shape == ref[key].shape
                e = rel_per_entry(g[key], ref[key])
                say(f"  N {s.n_panels} {key}: {e:.2e}")
                assert e < 1e-9
    finally:
        for s in sessions:
            s.close()
