"""The workload meshes: panel counts and orientation of the configurations' bodies."""

import numpy as np
import pytest

from hydrobem_amd import meshes


def normals(v):
    n = np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 1])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


@pytest.mark.parametrize("make,count", [
    (lambda: meshes.hemisphere_panels(1.0, 16, 32, quarter=False), 512),
    (lambda: meshes.cylinder_panels(1.0, 1.0, 64, 32, 32, quarter=False, cosine=True), 4096),
    (lambda: meshes.cylinder_panels(1.0, 1.0, 128, 32, 32, quarter=False, cosine=True), 8192),
    (lambda: meshes.box_panels(40, 10, 5, n=(168, 43, 12)), 12288),
])
def test_counts_and_outward_normals(make, count):
    v = make()
    assert v.shape == (count, 4, 3)
    c = v.mean(axis=1)
    n = normals(v)
    # every panel's normal points away from the body axis or down (into the fluid)
    out = n[:, 0] * c[:, 0] + n[:, 1] * c[:, 1] - n[:, 2] * 1e-3 * np.sign(c[:, 2])
    assert np.all((out > -1e-12) | (n[:, 2] < -0.5))


def test_expand_symmetry_keeps_orientation():
    q = meshes.cylinder_panels(1.0, 0.5, n_phi=8, n_z=4, n_r=4, quarter=True)
    v = meshes.expand_symmetry(q)
    assert v.shape[0] == 4 * q.shape[0]
    c = v.mean(axis=1)
    n = normals(v)
    side = np.abs(n[:, 2]) < 0.5
    assert np.all(n[side, 0] * c[side, 0] + n[side, 1] * c[side, 1] > 0)
    assert np.all(n[~side, 2] < 0)
