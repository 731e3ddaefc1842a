"""The regression rule of tests/test_gpu_regression.py on the CPU: it catches an error of 2e-5 of
an entry's own scale in one entry and passes the baseline itself."""

import numpy as np

from conftest import GOLDEN
from test_gpu_regression import regression_error


def test_regression_rule_fails_a_planted_error():
    ref = np.load(GOLDEN / "lo_cylinder_drift.npz", allow_pickle=False)["kochin_radiation"]
    bad = ref.astype(np.complex128).copy()
    bad[1, 17, 2] += 2e-5 * np.abs(ref[:, 17, 2]).max()
    assert not regression_error(bad, ref)[1]
    assert regression_error(ref.astype(np.complex128), ref)[1]
