"""More headings than one system's 32 right-hand sides (GPU): the solve takes the headings in
groups of 32 - 6 bodies, each with the radiation columns; its outputs equal those of the groups
solved one by one (the same systems), for one body (26 headings a group) and two (20)."""

import sys

import numpy as np
import pytest

import hydrobem_amd as hb
from hydrobem_amd import meshes

pytestmark = pytest.mark.gpu
KEYS = ("excitation_diffraction", "excitation_froude_krylov", "excitation_scattering", "excitation_haskind")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("bodies,precision", [(1, "fp64"), (2, "mixed")])
def test_many_headings_in_groups(bodies, precision):
    v = meshes.hemisphere_panels(1.0, 6, 12, quarter=False)
    body, centres = None, None
    if bodies == 2:
        v = np.concatenate([v, v + np.array([3.0, 0.0, 0.0])])
        body = np.repeat([0, 1], len(v) // 2).astype(np.int32)
        centres = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    heads = np.linspace(0.0, 350.0, 36)
    group = 32 - 6 * bodies
    w = np.array([0.8, 1.6])
    say(f"case: {bodies} bodies, {len(heads)} headings ({6 * bodies + len(heads)} right-hand sides), groups of "
        f"{group}, {precision}")
    with hb.Session(0) as s:
        s.set_environment(1025.0, 9.81)
        s.set_panels(v, body, centres)
        whole = s.solve(w, heads, precision=precision)
        parts = [s.solve(w, heads[h0:h0 + group], precision=precision) for h0 in range(0, len(heads), group)]
    for key in ("added_mass", "radiation_damping"):
        assert np.array_equal(whole[key], parts[0][key]), key
    for key in KEYS:
        ref = np.concatenate([p[key] for p in parts], axis=1)
        assert whole[key].shape == ref.shape
        assert np.array_equal(whole[key], ref), key
    # the symmetry of the sweep: heading b and 360 - b differ only by the sign of the y terms
    X = whole["excitation_diffraction"]
    assert np.abs(X).max() > 0.0
