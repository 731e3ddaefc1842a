"""The C-ABI library on the CPU side: it loads, exports every symbol include/hydrobem_amd.h
declares, refuses bad arguments before touching a GPU, and reports a missing GPU as a state
error (no compute calls here)."""

import ctypes
import re

import pytest

import hydrobem_amd as hb
from conftest import ROOT


def declared_symbols():
    text = (ROOT / "include" / "hydrobem_amd.h").read_text()
    return sorted(set(re.findall(r"\b(hbg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = hb.library()
    names = declared_symbols()
    assert len(names) >= 12
    for name in names:
        assert hasattr(L, name), name
        assert name in hb.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_errors():
    L = hb.library()
    assert b"gfx950" in L.hbg_version()
    h = ctypes.c_void_p()
    assert L.hbg_session_create(0, None) == 7  # ARGUMENT: out-pointer checked first
    assert b"null" in L.hbg_last_error_message()
    assert L.hbg_set_environment(None, 1025.0, 9.81, float("inf")) == 7
    assert L.hbg_session_free(None) == 0


def test_argument_errors_of_the_geometry_and_output_entry_points():
    # checked before any device call, so they run without a GPU
    L = hb.library()
    assert L.hbg_set_symmetry(None, 1, 0) == 7
    assert L.hbg_set_walls(None, 0, 1) == 7
    assert L.hbg_set_field_points(None, None, 0) == 7
    assert L.hbg_field_values(None, None, None, None, None) == 7
    assert L.hbg_set_kochin(None, 8) == 7
    assert L.hbg_kochin_values(None, None, None, None, None) == 7
    assert b"null" in L.hbg_last_error_message()


@pytest.mark.skipif(__import__("torch").cuda.is_available(), reason="GPU present: covered by the GPU tests")
def test_session_without_gpu_is_a_state_error():
    with pytest.raises(hb.HbgError) as e:
        hb.Session()
    assert e.value.status in (4,)


def test_wamit_files_on_host(tmp_path):
    import numpy as np

    F, H, D = 2, 1, 6
    rng = np.random.default_rng(0)
    res = {"omegas": np.array([0.5, 1.0]), "headings": np.array([0.0]),
           "added_mass": rng.normal(size=(F, D, D)) * 1e3, "radiation_damping": rng.normal(size=(F, D, D)) * 1e2,
           "excitation_diffraction": rng.normal(size=(F, H, D)) + 1j * rng.normal(size=(F, H, D)),
           "excitation_haskind": rng.normal(size=(F, H, D)) + 1j * rng.normal(size=(F, H, D))}
    stem = tmp_path / "body"
    hb.wamit(stem, res, density=1025.0, gravity=9.81, stamp="2026-01-01 00:00:00")
    t1 = (tmp_path / "body.1").read_bytes().decode()
    lines = t1.split("\r\n")
    assert lines[0] == " HydroBEM 0.1.0-amd   body.1   2026-01-01 00:00:00"
    assert len(lines) == 1 + F * D * D + 1 and lines[-1] == ""
    # PER I J Abar Bbar: E14.6 (0P), I6, I6, E14.6 (1P), E14.6 (1P)
    row = lines[1]
    assert len(row) == 14 + 6 + 6 + 14 + 14
    per = float(row[:14])
    assert per == pytest.approx(2 * np.pi / 0.5, rel=5e-6)  # 0P: six significant digits
    assert row[:14].strip().startswith("0.")
    assert int(row[14:20]) == 1 and int(row[20:26]) == 1
    assert float(row[26:40]) == pytest.approx(res["added_mass"][0, 0, 0] / 1025.0, rel=1e-6)
    assert float(row[40:54]) == pytest.approx(res["radiation_damping"][0, 0, 0] / (0.5 * 1025.0), rel=1e-6)
    t3 = (tmp_path / "body.3").read_bytes().decode().split("\r\n")
    r3 = t3[1]
    assert len(r3) == 14 * 2 + 6 + 4 * 14
    z = res["excitation_diffraction"][0, 0, 0] / (1025.0 * 9.81)
    assert float(r3[34:48]) == pytest.approx(abs(z), rel=1e-6)
    assert float(r3[48:62]) == pytest.approx(np.degrees(np.angle(z)), rel=1e-6, abs=1e-5)
    # overwrite refused
    with pytest.raises(hb.HbgError) as e:
        hb.wamit(stem, res, density=1025.0, gravity=9.81, overwrite=False)
    assert e.value.status == 10


def test_wamit_limit_rows_on_host(tmp_path):
    # the radiation limits in the order of the frequencies: period -1 (omega = 0) and 0 (omega =
    # infinity), the added mass alone; the force files list the ordinary frequencies only
    import math

    import numpy as np

    F, H, D = 3, 1, 6
    rng = np.random.default_rng(1)
    res = {"omegas": np.array([0.0, 0.8, math.inf]), "headings": np.array([0.0]),
           "added_mass": rng.normal(size=(F, D, D)) * 1e3, "radiation_damping": np.zeros((F, D, D)),
           "excitation_diffraction": rng.normal(size=(F, H, D)) + 1j * rng.normal(size=(F, H, D)),
           "excitation_haskind": rng.normal(size=(F, H, D)) + 1j * rng.normal(size=(F, H, D))}
    res["radiation_damping"][1] = rng.normal(size=(D, D)) * 1e2
    stem = tmp_path / "lim"
    hb.wamit(stem, res, density=1025.0, gravity=9.81, stamp="2026-01-01 00:00:00")
    lines = (tmp_path / "lim.1").read_bytes().decode().split("\r\n")[1:-1]
    assert len(lines) == F * D * D
    zero, ordinary, inf = lines[:36], lines[36:72], lines[72:]
    assert all(len(r) == 14 + 6 + 6 + 14 for r in zero + inf)
    assert all(len(r) == 14 + 6 + 6 + 14 + 14 for r in ordinary)
    assert float(zero[0][:14]) == -1.0 and float(inf[0][:14]) == 0.0
    assert float(ordinary[0][:14]) == pytest.approx(2 * np.pi / 0.8, rel=5e-6)
    assert float(zero[7][26:40]) == pytest.approx(res["added_mass"][0, 1, 1] / 1025.0, rel=1e-6)
    assert float(inf[7][26:40]) == pytest.approx(res["added_mass"][2, 1, 1] / 1025.0, rel=1e-6)
    t3 = (tmp_path / "lim.3").read_bytes().decode().split("\r\n")[1:-1]
    assert len(t3) == 1 * H * D and all(float(r[:14]) == pytest.approx(2 * np.pi / 0.8, rel=5e-6) for r in t3)


def test_wamit_limit_rows_skip_undefined_entries(tmp_path):
    # finite depth, omega = 0: the heave row and column are NaN (a net volume flux): no rows for
    # them, every other entry written as usual
    import math

    import numpy as np

    D = 6
    A = np.arange(1.0, 1.0 + 3 * D * D).reshape(3, D, D) * 10.0
    A[0, 2, :] = A[0, :, 2] = math.nan
    res = {"omegas": np.array([0.0, 0.8, math.inf]), "headings": np.array([0.0]), "added_mass": A,
           "radiation_damping": np.zeros((3, D, D)), "excitation_diffraction": np.ones((3, 1, D), complex),
           "excitation_haskind": np.ones((3, 1, D), complex)}
    hb.wamit(tmp_path / "fd", res, density=1025.0, gravity=9.81, stamp="2026-01-01 00:00:00")
    lines = (tmp_path / "fd.1").read_bytes().decode().split("\r\n")[1:-1]
    zero = [r for r in lines if float(r[:14]) == -1.0]
    assert len(zero) == D * D - (2 * D - 1) and len(lines) == 3 * D * D - (2 * D - 1)
    assert all(int(r[14:20]) != 3 and int(r[20:26]) != 3 for r in zero)


def test_wamit_field_point_files_on_host(tmp_path):
    # .fpt: the points once; .6p: per ordinary frequency the radiation block (points in order, Re,
    # Im of eta / L^r per column) then the diffraction block (per heading the points, Re, Im)
    import math

    import numpy as np

    F, H, D, P, L = 2, 2, 6, 3, 2.0
    rng = np.random.default_rng(3)
    c = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)  # noqa: E731
    res = {"omegas": np.array([0.7, 1.4]), "headings": np.array([0.0, 45.0]), "added_mass": rng.normal(size=(F, D, D)),
           "radiation_damping": rng.normal(size=(F, D, D)), "excitation_diffraction": c(F, H, D),
           "excitation_haskind": c(F, D, H)}
    pts = np.array([[3.0, 0.0, 0.0], [0.0, 4.5, -1.0], [-2.0, -2.0, -0.25]])
    field = {"frad": c(F, P, D), "ftot": c(F, P, H)}
    hb.wamit(tmp_path / "fp", res, density=1025.0, gravity=9.81, reference_length=L, stamp="2026-01-01 00:00:00",
             field_points=pts, field=field)
    fpt = (tmp_path / "fp.fpt").read_bytes().decode().split("\r\n")[1:-1]
    assert len(fpt) == P and all(len(r) == 10 + 3 * 14 for r in fpt)
    assert [int(r[:10]) for r in fpt] == [1, 2, 3]
    assert np.allclose([[float(r[10 + 14 * j:24 + 14 * j]) for j in range(3)] for r in fpt], pts, rtol=1e-6)
    rows = (tmp_path / "fp.6p").read_bytes().decode().split("\r\n")[1:-1]
    assert len(rows) == F * (P + H * P)
    for f in range(F):
        block = rows[f * (P + H * P):(f + 1) * (P + H * P)]
        per = 2 * math.pi / res["omegas"][f]
        for q in range(P):
            r = block[q]
            assert len(r) == 14 + 10 + 2 * 14 * D and int(r[14:24]) == q + 1
            assert float(r[:14]) == pytest.approx(per, rel=5e-6)
            for k in range(D):
                z = field["frad"][f, q, k] / L ** (1 if k >= 3 else 0)
                assert float(r[24 + 28 * k:38 + 28 * k]) == pytest.approx(z.real, rel=1e-6, abs=1e-12)
                assert float(r[38 + 28 * k:52 + 28 * k]) == pytest.approx(z.imag, rel=1e-6, abs=1e-12)
        for h in range(H):
            for q in range(P):
                r = block[P + h * P + q]
                assert len(r) == 14 + 14 + 10 + 28 and int(r[28:38]) == q + 1
                assert float(r[14:28]) == res["headings"][h]
                z = field["ftot"][f, q, h]
                assert float(r[38:52]) == pytest.approx(z.real, rel=1e-6) and float(r[52:66]) == pytest.approx(
                    z.imag, rel=1e-6)


def test_wamit_force_parts_and_motions_on_host(tmp_path):
    # .2fk / .2sc: the Froude-Krylov force by solved column and Haskind - Froude-Krylov; .3fk / .3sc:
    # the Froude-Krylov and scattering parts; .4: the motions times L^ri (rotations), not divided by rho g
    import numpy as np

    F, H, D, L = 2, 2, 6, 2.0
    rng = np.random.default_rng(2)

    def c(*shape):
        return rng.normal(size=shape) + 1j * rng.normal(size=shape)

    res = {"omegas": np.array([0.5, 1.0]), "headings": np.array([0.0, 90.0]),
           "added_mass": rng.normal(size=(F, D, D)), "radiation_damping": rng.normal(size=(F, D, D)),
           "excitation_diffraction": c(F, H, D), "excitation_haskind": c(F, H, D),
           "excitation_froude_krylov": c(F, H, D), "excitation_scattering": c(F, H, D)}
    rao = c(F, H, D)
    stem = tmp_path / "parts"
    hb.wamit(stem, res, density=1000.0, gravity=10.0, reference_length=L, rao=rao, stamp="s")

    def rows(ext):
        lines = (tmp_path / ("parts" + ext)).read_bytes().decode().split("\r\n")[1:-1]
        assert len(lines) == F * H * D
        return [complex(float(r[62:76]), float(r[76:90])) for r in lines]

    def scale(i):
        return 1000.0 * 10.0 * L ** (2 + (1 if i >= 3 else 0))

    q = 1 * D + 4  # frequency 0, heading 1, row 5 (a rotation)
    hk, fk, sc = res["excitation_haskind"][0, 1, 4], res["excitation_froude_krylov"][0, 1, 4], \
        res["excitation_scattering"][0, 1, 4]
    assert rows(".2fk")[q] == pytest.approx(fk / scale(4), rel=1e-5)
    assert rows(".2sc")[q] == pytest.approx((hk - fk) / scale(4), rel=1e-5)
    assert rows(".3fk")[q] == pytest.approx(fk / scale(4), rel=1e-5)
    assert rows(".3sc")[q] == pytest.approx(sc / scale(4), rel=1e-5)
    assert rows(".4")[q] == pytest.approx(rao[0, 1, 4] * L, rel=1e-5)
    assert rows(".4")[2] == pytest.approx(rao[0, 0, 2], rel=1e-5)


# the reference's formatter table (crates/hydrobem/tests/io.rs:37-65): value, 1P, field text
FORTRAN_E_TABLE = [
    (1.0, False, "  0.100000E+01"),
    (-123.4567, False, " -0.123457E+03"),
    (0.0, False, "  0.000000E+00"),
    (-0.0, False, "  0.000000E+00"),
    (6.2831853, False, "  0.628319E+01"),
    (0.99999951, False, "  0.100000E+01"),
    (1e-5, False, "  0.100000E-04"),
    (9.0e98, False, "  0.900000E+99"),
    (1.0, True, "  1.000000E+00"),
    (-1.0, True, " -1.000000E+00"),
    (9.9999996, True, "  1.000000E+01"),
    (0.0, True, "  0.000000E+00"),
    (123456.78, True, "  1.234568E+05"),
    (-2.5e-99, True, " -2.500000E-99"),
    (float(__import__("numpy").float32(0.1)), True, "  1.000000E-01"),
]


def fortran_e(x, one_p, width=14, digits=6):
    buf = ctypes.create_string_buffer(80)
    st = hb.library().hbg_fortran_e(float(x), int(one_p), width, digits, buf, 80)
    return st, buf.value.decode()


@pytest.mark.parametrize("x,one_p,text", FORTRAN_E_TABLE)
def test_fortran_e_matches_the_reference_table(x, one_p, text):
    st, got = fortran_e(x, one_p)
    assert st == 0 and got == text, (x, one_p, got)


def test_fortran_fields_width_integers_and_refusals():
    assert fortran_e(1.0, True, 15) == (0, "   1.000000E+00")
    buf = ctypes.create_string_buffer(16)
    assert hb.library().hbg_fortran_i(12, 6, buf, 16) == 0 and buf.value == b"    12"
    # a three-digit exponent does not fit, nor a non-finite value or a long integer
    for x in (1e100, -3e101, 1e-100, float("nan"), float("inf")):
        assert fortran_e(x, True)[0] == 1, x
    assert fortran_e(1e99, False)[0] == 1
    assert hb.library().hbg_fortran_i(1234567, 6, buf, 16) == 1
    assert hb.library().hbg_fortran_e(1.0, 1, 14, 6, buf, 4) == 7  # capacity too small
