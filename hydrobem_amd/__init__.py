"""hydrobem_amd: HydroBEM's low-order frequency-domain solve on one AMD MI355X (gfx950).

The work happens in the HIP library hydrobem_amd/_build/libhydrobem_amd.so behind the C ABI of
include/hydrobem_amd.h. This module is its Python host mirror: `Session.solve` returns the
quantities of the reference's results (solve_driver.rs:336-600 common_results: added_mass and
radiation_damping [F, D, K], excitation_diffraction / _froude_krylov / _scattering [F, H, D],
excitation_haskind [F, H, K]) for wetted bodies in infinite depth. There is no CPU fallback: the
library must be built (`python -m hydrobem_amd.build`) and a GPU present, or the calls fail.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import meshes  # noqa: F401

__all__ = ["Session", "HbgError", "library", "LIB_PATH", "wamit", "solve_group"]

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libhydrobem_amd.so"

STATUS = {0: "OK", 1: "VALUE", 2: "TYPE", 3: "UNSUPPORTED", 4: "STATE", 5: "CANCELLED", 6: "INTERNAL",
          7: "ARGUMENT", 8: "NUMERICAL", 9: "IO", 10: "EXISTS"}


class HbgError(RuntimeError):
    """A failed library call: `status` is the HbgStatus code, the message the library's own."""

    def __init__(self, status: int, message: str, location: str = ""):
        super().__init__(f"{STATUS.get(status, status)}: {message}" + (f" ({location})" if location else ""))
        self.status = status
        self.message = message


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
PROGRESS = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32)


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("logarithmic", ctypes.c_int32), ("log_tolerance", ctypes.c_double),
                ("surrogate", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SurrogateTraining(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("n_train", ctypes.c_int32), ("n_holdout", ctypes.c_int32),
                ("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("learning_rate", ctypes.c_double)]


_fp = ctypes.POINTER(ctypes.c_float)


_lib = None

# the C ABI of include/hydrobem_amd.h: name -> (restype, argtypes)
SIGNATURES = {
    "hbg_last_error_message": (ctypes.c_char_p, []),
    "hbg_last_error_location": (ctypes.c_char_p, []),
    "hbg_version": (ctypes.c_char_p, []),
    "hbg_session_create": (ctypes.c_int32, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "hbg_session_free": (ctypes.c_int32, [ctypes.c_void_p]),
    "hbg_set_environment": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "hbg_set_panels": (ctypes.c_int32, [ctypes.c_void_p, _dp, ctypes.c_int32, _ip, ctypes.c_int32, _dp]),
    "hbg_set_interior": (ctypes.c_int32, [ctypes.c_void_p, _dp, ctypes.c_int32, _ip]),
    "hbg_set_lids": (ctypes.c_int32, [ctypes.c_void_p, _dp, ctypes.c_int32, _ip, _dp]),
    "hbg_set_symmetry": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "hbg_set_walls": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "hbg_set_field_points": (ctypes.c_int32, [ctypes.c_void_p, _dp, ctypes.c_int32]),
    "hbg_field_values": (ctypes.c_int32, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "hbg_set_kochin": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "hbg_kochin_values": (ctypes.c_int32, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "hbg_field_shape": (ctypes.c_int32, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int32)] * 4),
    "hbg_kochin_shape": (ctypes.c_int32, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int32)] * 4),
    "hbg_solve": (ctypes.c_int32, [ctypes.c_void_p, _dp, ctypes.c_int32, _dp, ctypes.c_int32,
                                   ctypes.POINTER(Options), _dp, _dp, _dp, _dp, _dp, _dp, PROGRESS,
                                   ctypes.c_void_p]),
    "hbg_solve_group": (ctypes.c_int32, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, _dp, ctypes.c_int32, _dp,
                                         ctypes.c_int32, ctypes.POINTER(Options)] + [ctypes.POINTER(_dp)] * 6),
    "hbg_wave_kernel": (ctypes.c_int32, [ctypes.c_void_p, _dp, _dp, ctypes.c_int64, _dp]),
    "hbg_finite_kernel": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _dp,
                                           _dp, _dp, _dp]),
    "hbg_dispersion_roots": (ctypes.c_int32, [ctypes.c_double, ctypes.c_int32, _dp]),
    "hbg_wave_kernel32": (ctypes.c_int32, [ctypes.c_void_p, _dp, _dp, ctypes.c_int32, _dp]),
    "hbg_finite_kernel32": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _dp,
                                             _dp, _dp, _dp]),
    "hbg_statics": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, _dp]),
    "hbg_assemble": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(Options), _dp, _dp]),
    "hbg_lu_solve": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.c_int32, _dp]),
    "hbg_lu_solve_noalloc": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.c_int32, _dp]),
    "hbg_fortran_e": (ctypes.c_int32, [ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_char_p, ctypes.c_int32]),
    "hbg_fortran_i": (ctypes.c_int32, [ctypes.c_int64, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]),
    "hbg_set_profiling": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "hbg_motions": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                     _dp, _dp, _dp, _dp, _dp]),
    "hbg_surrogate_size": (ctypes.c_int32, []),
    "hbg_surrogate_train": (ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(SurrogateTraining), _dp]),
    "hbg_surrogate_set": (ctypes.c_int32, [ctypes.c_void_p, _fp, ctypes.c_int32]),
    "hbg_surrogate_get": (ctypes.c_int32, [ctypes.c_void_p, _fp, ctypes.c_int32]),
    "hbg_surrogate_eval": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, _dp, _dp]),
    "hbg_finite_remainder": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                              _dp, _dp, _dp, _dp]),
    "hbg_surrogate_gradient": (ctypes.c_int32, [ctypes.c_void_p, _fp, ctypes.c_int32, _fp, _fp, _fp, _dp]),
    "hbg_set_batch": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "hbg_autotune": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp,
                                      ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "hbg_set_tuning": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "hbg_get_tuning": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "hbg_profile_read": (ctypes.c_int32, [ctypes.c_void_p, _dp, ctypes.c_int32]),
    "hbg_write_wamit": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _dp, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32,
                                         _dp, _dp, _dp, _dp, ctypes.c_char_p]),
    "hbg_write_wamit_files": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, _dp, ctypes.c_int32, _dp, ctypes.c_int32,
                                               ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_char_p]),
    "hbg_write_wamit_all": (ctypes.c_int32, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, _dp, ctypes.c_int32, _dp, ctypes.c_int32,
                                             ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int32, _dp,
                                             _dp, _dp, ctypes.c_char_p]),
}


def library():
    """The loaded HIP library (fails loudly if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hydrobem_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int):
    if status != 0:
        L = library()
        raise HbgError(status, L.hbg_last_error_message().decode(), L.hbg_last_error_location().decode())


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


PRECISIONS = {"fp64": 0, "mixed": 1, "mixed32": 2}


def _options(logarithmic=True, log_tolerance=1e-4, precision=0, surrogate=False) -> Options:
    return Options(int(precision), int(bool(logarithmic)), float(log_tolerance), int(bool(surrogate)), 0)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Session:
    """One geometry on one GPU: panels, environment, statics cached on the device.

    The Python face of the reference's solve (hb.Model.solve, python/hydrobem/engine.py:286-301)
    for wetted panels in infinite or finite depth, with symmetry planes or walls, and field points.
    """

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(library().hbg_session_create(int(device), ctypes.byref(self._h)))
        self.n_panels = 0
        self.n_bodies = 0
        self.n_points = 0
        self.n_kochin = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            library().hbg_session_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_environment(self, density=1000.0, gravity=9.80665, depth=float("inf")):
        _check(library().hbg_set_environment(self._h, float(density), float(gravity), float(depth)))

    def set_panels(self, vertices, body=None, rotation_centres=None):
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 4, 3)
        n = v.shape[0]
        b = None if body is None else np.ascontiguousarray(body, dtype=np.int32)
        nb = 1 if b is None else int(b.max()) + 1
        rc = None if rotation_centres is None else np.ascontiguousarray(rotation_centres, dtype=np.float64)
        _check(library().hbg_set_panels(self._h, _p(v), n, b.ctypes.data_as(_ip) if b is not None else None, nb,
                                        _p(rc)))
        self.n_panels, self.n_bodies = n, nb
        self.n_interior = 0
        self.n_lids = 0

    def set_interior(self, vertices=None, body=None):
        """The interior free surface (panels in z = 0 inside the waterlines, in the stored symmetry
        form of the body panels; meshes.interior_graded makes them): extra unknowns that remove the
        irregular frequencies (None: none)."""
        v = np.zeros((0, 4, 3)) if vertices is None else np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 4, 3)
        b = None if body is None else np.ascontiguousarray(body, dtype=np.int32)
        _check(library().hbg_set_interior(self._h, _p(v) if len(v) else None, len(v),
                                          b.ctypes.data_as(_ip) if b is not None else None))
        self.n_interior = len(v)

    def set_lids(self, vertices=None, damping=0.0, body=None):
        """Damping lids (hbg_set_lids): panels in z = 0 on the outer free surface with the damping
        factor eps (a scalar or one per panel; kappa = w^2 / (g + i w eps)); None: none."""
        v = np.zeros((0, 4, 3)) if vertices is None else np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 4, 3)
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(damping, dtype=np.float64), (len(v),)))
        b = None if body is None else np.ascontiguousarray(body, dtype=np.int32)
        _check(library().hbg_set_lids(self._h, _p(v) if len(v) else None, len(v),
                                      b.ctypes.data_as(_ip) if b is not None else None, _p(e) if len(v) else None))
        self.n_lids = len(v)

    def set_symmetry(self, sym_x=False, sym_y=False):
        """Symmetry planes x = 0 (sym_x) and y = 0 (sym_y): the panels set are one sector of the
        body and the solves return the whole body's quantities."""
        _check(library().hbg_set_symmetry(self._h, int(bool(sym_x)), int(bool(sym_y))))

    def set_field_points(self, points=None):
        """Points [n][3] (z <= 0) whose elevations every later solve evaluates (None: none)."""
        p = np.zeros((0, 3)) if points is None else np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        _check(library().hbg_set_field_points(self._h, _p(p) if len(p) else None, len(p)))
        self.n_points = len(p)

    def _shape(self, name, n_omegas, n_headings):
        dims = [ctypes.c_int32() for _ in range(4)]
        _check(getattr(library(), name)(self._h, *[ctypes.byref(d) for d in dims]))
        F, P, K, H = (d.value for d in dims)
        if (n_omegas is not None and n_omegas != F) or (n_headings is not None and n_headings != H):
            raise ValueError(f"the last solve had {F} frequencies and {H} headings, not {n_omegas} and {n_headings}")
        return F, P, K, H

    def field_values(self, n_omegas=None, n_headings=None):
        """The field values of the last solve: frad [F][P][K], finc, fsca, ftot [F][P][H]; the sizes
        are the library's record of that solve (n_omegas, n_headings, when given, must match)."""
        F, P, K, H = self._shape("hbg_field_shape", n_omegas, n_headings)
        frad = np.zeros((F, P, K), np.complex128)
        finc, fsca, ftot = (np.zeros((F, P, H), np.complex128) for _ in range(3))
        _check(library().hbg_field_values(self._h, _p(frad), _p(finc), _p(fsca), _p(ftot)))
        return {"frad": frad, "finc": finc, "fsca": fsca, "ftot": ftot}

    def set_kochin(self, n_angles=0):
        """Kochin functions on n_angles angles 2 pi t / n_angles in every later solve (0: none)."""
        _check(library().hbg_set_kochin(self._h, int(n_angles)))
        self.n_kochin = int(n_angles)

    def kochin_values(self, n_omegas=None, n_headings=None):
        """The Kochin functions of the last solve: kr, krt [F][NK][K], kd, kdt [F][NK][H] (sizes as
        field_values)."""
        F, NK, K, H = self._shape("hbg_kochin_shape", n_omegas, n_headings)
        kr, krt = (np.zeros((F, NK, K), np.complex128) for _ in range(2))
        kd, kdt = (np.zeros((F, NK, H), np.complex128) for _ in range(2))
        _check(library().hbg_kochin_values(self._h, _p(kr), _p(krt), _p(kd), _p(kdt)))
        return {"kr": kr, "krt": krt, "kd": kd, "kdt": kdt}

    def set_walls(self, wall_x=False, wall_y=False):
        """Reflecting walls x = 0 (wall_x) and y = 0 (wall_y), the bodies on their positive side:
        the solves return the real bodies' quantities in the walled domain."""
        _check(library().hbg_set_walls(self._h, int(bool(wall_x)), int(bool(wall_y))))

    def solve(self, omegas, headings_deg=(), logarithmic=True, log_tolerance=1e-4, progress=None, precision="fp64",
              surrogate=False):
        """The quantities of one solve (the reference's result names and axes). precision "fp64"
        (the plain path), "mixed" (complex64 LU on the matrix cores + FP64 refinement) or "mixed32"
        (the same with the FP32 far-field assembly);
        surrogate: the finite-depth evanescent remainder from the trained network."""
        w = np.ascontiguousarray(omegas, dtype=np.float64).reshape(-1)
        h = np.ascontiguousarray(headings_deg, dtype=np.float64).reshape(-1)
        F, H, D = len(w), len(h), 6 * self.n_bodies
        A = np.zeros((F, D, D))
        B = np.zeros((F, D, D))
        X = np.zeros((F, H, D), np.complex128)
        FK = np.zeros((F, H, D), np.complex128)
        SC = np.zeros((F, H, D), np.complex128)
        HK = np.zeros((F, D, H), np.complex128)
        cb = PROGRESS(lambda user, stage, done, total: int(progress(stage, done, total)) if progress else 1)
        opt = _options(logarithmic, log_tolerance, PRECISIONS[precision], surrogate)
        _check(library().hbg_solve(self._h, _p(w), F, _p(h), H, ctypes.byref(opt), _p(A), _p(B), _p(X), _p(FK),
                                   _p(SC), _p(HK), cb, None))
        return {"omegas": w, "headings": h, "added_mass": A, "radiation_damping": B,
                "excitation_diffraction": X, "excitation_froude_krylov": FK, "excitation_scattering": SC,
                "excitation_haskind": np.transpose(HK, (0, 2, 1)).copy()}

    PROFILE_KEYS = ("statics_ms", "assembly_ms", "rhs_ms", "lu_ms", "lu_gemm_ms", "lu_gemm_flops",
                    "lu_gemm_launches", "forces_ms", "frequencies", "lu_flops", "assembly_bytes", "statics_runs",
                    "lu_fallback_panels", "mixed_refinement_steps", "mixed_max_backward_error",
                    "mixed_fp64_fallbacks", "mixed_refinement_ms", "mixed_max_first_backward_error",
                    "lu_big_gemm_ms", "lu_big_gemm_flops", "lu_big_gemm_launches")

    def set_batch(self, max_batch: int):
        """Frequencies in flight in solve (0: as many as fit, at most 48)."""
        _check(library().hbg_set_batch(self._h, int(max_batch)))

    def motions(self, result, mass, stiffness, external_damping=None) -> np.ndarray:
        """RAOs [F][H][D] of a solve's result (hbg_motions): mass, stiffness, external damping [D][D]."""
        w = np.ascontiguousarray(result["omegas"], dtype=np.float64)
        A = np.ascontiguousarray(result["added_mass"], dtype=np.float64)
        B = np.ascontiguousarray(result["radiation_damping"], dtype=np.float64)
        X = np.ascontiguousarray(result["excitation_diffraction"], dtype=np.complex128)
        F, H, D = X.shape
        M = np.ascontiguousarray(mass, dtype=np.float64)
        C = np.ascontiguousarray(stiffness, dtype=np.float64)
        Bx = None if external_damping is None else np.ascontiguousarray(external_damping, dtype=np.float64)
        out = np.zeros((F, H, D), np.complex128)
        _check(library().hbg_motions(self._h, F, _p(w), H, D, _p(M), _p(C), _p(Bx), _p(A), _p(B),
                                     X.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
        return out

    # ---- the learned surrogate
    @staticmethod
    def surrogate_size() -> int:
        return int(library().hbg_surrogate_size())

    REPORT_KEYS = ("train_loss", "holdout_rms", "holdout_max_T", "holdout_max_TR", "holdout_max_Tz",
                   "holdout_max_Tzeta", "data_seconds", "train_seconds", "steps", "nonfinite_targets",
                   "nonfinite_params", "reserved")

    def surrogate_train(self, n_train=2_000_000, n_holdout=100_000, epochs=4, batch=4096, learning_rate=2e-3,
                        seed=1) -> dict:
        t = SurrogateTraining(int(seed), int(n_train), int(n_holdout), int(epochs), int(batch), float(learning_rate))
        rep = np.zeros(12)
        _check(library().hbg_surrogate_train(self._h, ctypes.byref(t), _p(rep)))
        return {k: float(rep[i]) for i, k in enumerate(self.REPORT_KEYS)}

    def surrogate_set(self, params):
        p = _f32(params)
        _check(library().hbg_surrogate_set(self._h, p.ctypes.data_as(_fp), int(p.size)))

    def surrogate_get(self) -> np.ndarray:
        p = np.zeros(self.surrogate_size(), np.float32)
        _check(library().hbg_surrogate_get(self._h, p.ctypes.data_as(_fp), int(p.size)))
        return p

    def surrogate_eval(self, points) -> np.ndarray:
        """points [n][4] = (R/h, z/h, zeta/h, nu h) -> [n][4] = (h T, h^2 T_R, h^2 T_zeta, h^2 T_z)."""
        x = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 4)
        out = np.zeros_like(x)
        _check(library().hbg_surrogate_eval(self._h, x.shape[0], _p(x), _p(out)))
        return out

    def finite_remainder(self, h, K, R, z, zeta) -> np.ndarray:
        R, z, zeta = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (R, z, zeta))
        out = np.zeros((R.size, 4))
        _check(library().hbg_finite_remainder(self._h, float(h), float(K), R.size, _p(R), _p(z), _p(zeta), _p(out)))
        return out

    def surrogate_gradient(self, params, inputs, targets):
        p, x, y = _f32(params), _f32(inputs), _f32(targets)
        g = np.zeros_like(p)
        loss = np.zeros(1)
        _check(library().hbg_surrogate_gradient(self._h, p.ctypes.data_as(_fp), x.shape[0], x.ctypes.data_as(_fp),
                                                y.ctypes.data_as(_fp), g.ctypes.data_as(_fp), _p(loss)))
        return float(loss[0]), g

    TUNE_LU_GEMM = 0
    TUNE_LU_DEPTH = 1
    TUNE_LU_GRAPHS = 2
    TUNE_LU_GEMM_MIXED = 3
    TUNE_LU_DEPTH_MIXED = 4
    TUNE_MIXED32_TOL = 5
    TUNE_NN_RMIN = 6
    TUNE_LU_SUB = 7
    TUNE_LU_SUB_MIXED = 8
    TUNE_LU_SUB2 = 9
    TUNE_LU_SUB2_MIXED = 10

    def autotune(self, n, nrhs, precision="fp64") -> list:
        """Time the batched LU variants on synthetic systems and keep the fastest (hbg_autotune);
        returns the rows (variant, depth, sub-block, its sub-block, batch, ms per system)."""
        rep = np.zeros((128, 6))
        rows = ctypes.c_int32(0)
        _check(library().hbg_autotune(self._h, int(n), int(nrhs), PRECISIONS[precision], _p(rep), len(rep),
                                      ctypes.byref(rows)))
        return [tuple(r) for r in rep[:rows.value]]

    def set_tuning(self, key: int, value: int):
        _check(library().hbg_set_tuning(self._h, int(key), int(value)))

    def get_tuning(self, key: int) -> int:
        v = ctypes.c_int32(0)
        _check(library().hbg_get_tuning(self._h, int(key), ctypes.byref(v)))
        return v.value

    def set_profiling(self, enable: bool):
        _check(library().hbg_set_profiling(self._h, int(bool(enable))))

    def profile(self) -> dict:
        out = np.zeros(len(self.PROFILE_KEYS))
        _check(library().hbg_profile_read(self._h, _p(out), len(self.PROFILE_KEYS)))
        return {k: float(out[i]) for i, k in enumerate(self.PROFILE_KEYS)}

    def wave_kernel(self, X, Y):
        x = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        y = np.ascontiguousarray(Y, dtype=np.float64).reshape(-1)
        out = np.zeros((len(x), 7))
        _check(library().hbg_wave_kernel(self._h, _p(x), _p(y), len(x), _p(out)))
        return out

    def finite_kernel(self, h, K, R, z, zeta):
        """(Wr, WrR, WrZ, Wrz, Wi, WiR, WiZ, Wiz) of the finite-depth kernel at point pairs."""
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
        zeta = np.ascontiguousarray(zeta, dtype=np.float64).reshape(-1)
        out = np.zeros((len(R), 8))
        _check(library().hbg_finite_kernel(self._h, float(h), float(K), len(R), _p(R), _p(z), _p(zeta), _p(out)))
        return out

    def wave_kernel32(self, X, Y):
        """(Ft, Ft_X, Gi, Gi_X) of the FP32 evaluator (the precision path's) at 0 <= X, Y < 48."""
        x = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        y = np.ascontiguousarray(Y, dtype=np.float64).reshape(-1)
        out = np.zeros((len(x), 4))
        _check(library().hbg_wave_kernel32(self._h, _p(x), _p(y), len(x), _p(out)))
        return out

    def finite_kernel32(self, h, K, R, z, zeta):
        """finite_kernel's eight values from the FP32 evaluator."""
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
        zeta = np.ascontiguousarray(zeta, dtype=np.float64).reshape(-1)
        out = np.zeros((len(R), 8))
        _check(library().hbg_finite_kernel32(self._h, float(h), float(K), len(R), _p(R), _p(z), _p(zeta), _p(out)))
        return out

    def statics(self, logarithmic=True, log_tolerance=1e-4):
        out = np.zeros((self.n_panels, self.n_panels, 4))
        _check(library().hbg_statics(self._h, int(bool(logarithmic)), float(log_tolerance), _p(out)))
        return out

    def assemble(self, omega, logarithmic=True, log_tolerance=1e-4, precision="fp64"):
        """S and A of one frequency: the FP64 assembly, or (precision "mixed32") the FP32 one."""
        N = self.n_panels
        S = np.zeros((N, N), np.complex128)
        A = np.zeros((N, N), np.complex128)
        opt = _options(logarithmic, log_tolerance, PRECISIONS[precision])
        _check(library().hbg_assemble(self._h, float(omega), ctypes.byref(opt), _p(S), _p(A)))
        return S, A

    def lu_solve(self, a, b):
        """x = a^-1 b on the GPU (complex128); a (n, n), b (n,) or (n, r)."""
        n = a.shape[0]
        af = np.array(a, dtype=np.complex128, order="F", copy=True)
        bb = np.array(b, dtype=np.complex128, order="F", copy=True).reshape(n, -1, order="F")
        _check(library().hbg_lu_solve(self._h, n, af.ctypes.data_as(_dp), bb.shape[1], bb.ctypes.data_as(_dp)))
        return bb.reshape(b.shape, order="F")


def solve_group(sessions, omegas, headings_deg=(), logarithmic=True, log_tolerance=1e-4, precision="fp64",
                surrogate=False):
    """Session.solve for several geometries at once (hbg_solve_group): their (geometry, frequency)
    systems of uneven order share the launches of the batched LU. Returns one result per session."""
    w = np.ascontiguousarray(omegas, dtype=np.float64).reshape(-1)
    h = np.ascontiguousarray(headings_deg, dtype=np.float64).reshape(-1)
    F, H = len(w), len(h)
    outs, ptrs = [], {k: [] for k in ("A", "B", "X", "FK", "SC", "HK")}
    for s in sessions:
        D = 6 * s.n_bodies
        o = {"A": np.zeros((F, D, D)), "B": np.zeros((F, D, D)), "X": np.zeros((F, H, D), np.complex128),
             "FK": np.zeros((F, H, D), np.complex128), "SC": np.zeros((F, H, D), np.complex128),
             "HK": np.zeros((F, D, H), np.complex128)}
        outs.append(o)
        for k, v in o.items():
            ptrs[k].append(v.ctypes.data_as(_dp))
    arr = {k: (_dp * len(v))(*v) for k, v in ptrs.items()}
    hs = (ctypes.c_void_p * len(sessions))(*[s._h.value for s in sessions])
    opt = _options(logarithmic, log_tolerance, PRECISIONS[precision], surrogate)
    _check(library().hbg_solve_group(hs, len(sessions), _p(w), F, _p(h), H, ctypes.byref(opt), arr["A"], arr["B"],
                                     arr["X"], arr["FK"], arr["SC"], arr["HK"]))
    return [{"omegas": w, "headings": h, "added_mass": o["A"], "radiation_damping": o["B"],
             "excitation_diffraction": o["X"], "excitation_froude_krylov": o["FK"], "excitation_scattering": o["SC"],
             "excitation_haskind": np.transpose(o["HK"], (0, 2, 1)).copy()} for o in outs]


def dispersion_roots(H, n):
    """[x0, x1, ..., xn] of x0 tanh x0 = H and xi tan xi = -H."""
    out = np.zeros(n + 1)
    _check(library().hbg_dispersion_roots(float(H), int(n), _p(out)))
    return out


def wamit(stem, result, density, gravity, reference_length=1.0, overwrite=True, stamp="", n_bodies=1, rao=None,
          field_points=None, field=None):
    """The WAMIT files of a solve (the C ABI's hbg_write_wamit_all): .1, .2, .3, and .2fk, .2sc,
    .3fk, .3sc when the result holds the Froude-Krylov and scattering forces, .4 with rao (the
    motions [F][H][D] of Session.motions), .fpt and .6p with field_points [P][3] and field (the
    dict of Session.field_values)."""
    w = np.ascontiguousarray(result["omegas"], dtype=np.float64)
    h = np.ascontiguousarray(result["headings"], dtype=np.float64)
    A = np.ascontiguousarray(result["added_mass"], dtype=np.float64)
    B = np.ascontiguousarray(result["radiation_damping"], dtype=np.float64)
    HK = np.ascontiguousarray(np.transpose(result["excitation_haskind"], (0, 2, 1)), dtype=np.complex128)
    X = np.ascontiguousarray(result["excitation_diffraction"], dtype=np.complex128)

    def opt(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.complex128)

    FK, SC, RAO = opt(result.get("excitation_froude_krylov")), opt(result.get("excitation_scattering")), opt(rao)
    ptr = (lambda a: None if a is None else a.ctypes.data_as(_dp))
    pts = None if field_points is None else np.ascontiguousarray(field_points, dtype=np.float64).reshape(-1, 3)
    frad = opt(field["frad"]) if field is not None else None
    fdif = opt(field["ftot"]) if field is not None else None
    _check(library().hbg_write_wamit_all(str(stem).encode(), int(bool(overwrite)), float(density), float(gravity),
                                         float(reference_length), _p(w), len(w), _p(h), len(h), int(n_bodies),
                                         _p(A), _p(B), HK.ctypes.data_as(_dp), X.ctypes.data_as(_dp), ptr(FK),
                                         ptr(SC), ptr(RAO), 0 if pts is None else len(pts), ptr(pts), ptr(frad),
                                         ptr(fdif), stamp.encode()))
