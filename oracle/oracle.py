"""ctypes front of the FP64 CPU oracle (oracle/hb_oracle.cpp).

Test infrastructure only: the oracle is the checker of the HIP library. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and never as the
thing measured or shipped. Each function restates the reference algorithm cited in
hb_oracle.cpp's header.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libhb_oracle.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> Path:
    """Compile the oracle (g++ and OpenMP, no GPU) into oracle/_build/."""
    subprocess.run(["make", "-s", "-f", str(HERE / "Makefile")], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        for name, args in [
            ("orc_wave_kernel", [ctypes.c_double, ctypes.c_double, _dp]),
            ("orc_v_vx", [ctypes.c_double, ctypes.c_double, _dp]),
            ("orc_yt", [ctypes.c_double, _dp]),
            ("orc_struve01", [ctypes.c_double, _dp]),
            ("orc_gauss_legendre", [ctypes.c_int, _dp, _dp]),
            ("orc_make_panel", [_dp, _dp]),
            ("orc_rankine", [_dp, _dp, _dp]),
            ("orc_log_singularity", [_dp, _dp, _dp]),
            ("orc_statics", [ctypes.c_void_p, _dp]),
            ("orc_assemble", [ctypes.c_void_p, ctypes.c_double, _dp, _dp]),
            ("orc_lu_solve", [ctypes.c_int, _dp, ctypes.c_int, _dp]),
            ("orc_finite_kernel", [ctypes.c_double] * 5 + [_dp]),
            ("orc_finite_phi", [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]),
            ("orc_finite_limit_kernel", [ctypes.c_int] + [ctypes.c_double] * 4 + [_dp]),
            ("orc_solve", [ctypes.c_void_p, _dp, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.orc_dispersion_x0.argtypes = [ctypes.c_double]
        L.orc_dispersion_x0.restype = ctypes.c_double
        L.orc_dispersion_xn.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_dispersion_xn.restype = ctypes.c_double
        L.orc_problem_create.argtypes = [_dp, ctypes.c_int, _ip, ctypes.c_int, _dp, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, _ip]
        L.orc_problem_create.restype = ctypes.c_void_p
        L.orc_problem_free.argtypes = [ctypes.c_void_p]
        L.orc_problem_free.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(st, what):
    if st != 0:
        raise RuntimeError(f"oracle {what} failed with status {st}")


def wave_kernel(X: float, Y: float) -> np.ndarray:
    """(Ft, Ft_X, Ft_Y, Gi, Gi_X, Gi_Y, exp(-Y) J0) at (X, Y) = (K R, K |z + zeta|) (green_inf.rs:316)."""
    out = np.zeros(7)
    _check(lib().orc_wave_kernel(float(X), float(Y), _p(out)), "wave_kernel")
    return out


def v_vx(X: float, Y: float) -> np.ndarray:
    out = np.zeros(2)
    _check(lib().orc_v_vx(float(X), float(Y), _p(out)), "v_vx")
    return out


def yt(X: float) -> np.ndarray:
    out = np.zeros(2)
    _check(lib().orc_yt(float(X), _p(out)), "yt")
    return out


def struve01(x: float) -> np.ndarray:
    out = np.zeros(2)
    _check(lib().orc_struve01(float(x), _p(out)), "struve01")
    return out


def dispersion_x0(H: float) -> float:
    """x0 tanh x0 = H (dispersion.rs:19-52)."""
    return lib().orc_dispersion_x0(float(H))


def dispersion_xn(H: float, n: int) -> float:
    """xn tan xn = -H, (n - 1/2) pi < xn < n pi (dispersion.rs:62-95)."""
    return lib().orc_dispersion_xn(float(H), int(n))


def finite_kernel(h, K, R, z, zeta) -> np.ndarray:
    """(Wr, WrR, WrZ, Wrz, Wi, WiR, WiZ, Wiz) of the finite-depth regular part (green_fin.rs:1007)."""
    out = np.zeros(8)
    _check(lib().orc_finite_kernel(float(h), float(K), float(R), float(z), float(zeta), _p(out)), "finite_kernel")
    return out


def finite_limit_kernel(mode, h, R, z, zeta) -> np.ndarray:
    """The finite-depth regular part at a frequency limit (mode 1: omega = 0, 2: omega = infinity,
    green_fin.rs FreqMode::Zero / Infinite); the eight values of finite_kernel, Wi = 0."""
    out = np.zeros(8)
    _check(lib().orc_finite_limit_kernel(int(mode), float(h), float(R), float(z), float(zeta), _p(out)),
           "finite_limit_kernel")
    return out


def finite_phi(H, which, xb, vb) -> np.ndarray:
    out = np.zeros(3)
    _check(lib().orc_finite_phi(float(H), int(which), float(xb), float(vb), _p(out)), "finite_phi")
    return out


def gauss_legendre(n: int):
    x, w = np.zeros(n), np.zeros(n)
    _check(lib().orc_gauss_legendre(n, _p(x), _p(w)), "gauss_legendre")
    return x, w


def make_panel(v12) -> dict:
    v = np.ascontiguousarray(v12, dtype=np.float64).reshape(12)
    out = np.zeros(20)
    st = lib().orc_make_panel(_p(v), _p(out))
    if st != 0:
        return {"status": st}
    return {"status": 0, "centroid": out[0:3].copy(), "normal": out[3:6].copy(), "area": out[6],
            "distinct": int(out[7]), "proj": out[8:20].reshape(4, 3).copy()}


def rankine(v12, x) -> np.ndarray:
    v = np.ascontiguousarray(v12, dtype=np.float64).reshape(12)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(2)
    _check(lib().orc_rankine(_p(v), _p(xx), _p(out)), "rankine")
    return out


def log_singularity(v12, image) -> np.ndarray:
    v = np.ascontiguousarray(v12, dtype=np.float64).reshape(12)
    xx = np.ascontiguousarray(image, dtype=np.float64)
    out = np.zeros(2)
    _check(lib().orc_log_singularity(_p(v), _p(xx), _p(out)), "log_singularity")
    return out


def lu_solve(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Partial-pivot complex LU solve (linalg.rs:321-328), a (n, n) and b (n, r) complex."""
    n = a.shape[0]
    af = np.asfortranarray(a, dtype=np.complex128).copy(order="F")
    bb = np.asfortranarray(b.reshape(n, -1), dtype=np.complex128).copy(order="F")
    st = lib().orc_lu_solve(n, af.ctypes.data_as(_dp), bb.shape[1], bb.ctypes.data_as(_dp))
    _check(st, "lu_solve")
    return bb.reshape(b.shape)


class Problem:
    """One body (or several) of wetted panels in infinite depth, as the oracle solves it."""

    def __init__(self, vertices, body=None, rotation_centres=None, rho=1025.0, grav=9.81,
                 depth=float("inf"), logint=True, log_tolerance=1e-4):
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 4, 3)
        self.N = v.shape[0]
        b = np.zeros(self.N, np.int32) if body is None else np.ascontiguousarray(body, dtype=np.int32)
        self.nb = int(b.max()) + 1
        rc = np.zeros((self.nb, 3)) if rotation_centres is None else np.ascontiguousarray(rotation_centres, float)
        st = ctypes.c_int(0)
        self._v, self._b, self._rc = v, b, rc
        self.h = lib().orc_problem_create(_p(v), self.N, b.ctypes.data_as(_ip), self.nb, _p(rc), rho, grav,
                                          float(depth), int(bool(logint)), float(log_tolerance), ctypes.byref(st))
        _check(st.value, "problem_create")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_free(self.h)
            self.h = None

    def statics(self) -> np.ndarray:
        out = np.zeros((self.N, self.N, 4))
        _check(lib().orc_statics(self.h, _p(out)), "statics")
        return out

    def assemble(self, omega: float):
        """(S, A) column-major complex N x N as numpy (N, N) arrays indexed [k, n]."""
        S = np.zeros((self.N, self.N), np.complex128, order="F")
        A = np.zeros((self.N, self.N), np.complex128, order="F")
        _check(lib().orc_assemble(self.h, float(omega), S.ctypes.data_as(_dp), A.ctypes.data_as(_dp)), "assemble")
        return S, A

    def solve(self, omegas, headings_deg):
        w = np.ascontiguousarray(omegas, dtype=np.float64)
        h = np.ascontiguousarray(headings_deg, dtype=np.float64)
        F, H, D = len(w), len(h), 6 * self.nb
        A = np.zeros((F, D, D))
        B = np.zeros((F, D, D))
        X = np.zeros((F, H, D), np.complex128)
        FK = np.zeros((F, H, D), np.complex128)
        SC = np.zeros((F, H, D), np.complex128)
        HK = np.zeros((F, D, H), np.complex128)
        st = lib().orc_solve(self.h, _p(w), F, _p(h), H, _p(A), _p(B), X.ctypes.data_as(_dp),
                             FK.ctypes.data_as(_dp), SC.ctypes.data_as(_dp), HK.ctypes.data_as(_dp))
        _check(st, "solve")
        return {"added_mass": A, "radiation_damping": B, "excitation_diffraction": X,
                "excitation_froude_krylov": FK, "excitation_scattering": SC, "excitation_haskind": HK}


if __name__ == "__main__":  # pragma: no cover
    build()
    print(os.fspath(LIB))
