#!/usr/bin/env python3
"""Throughput of the low-order frequency-domain solve on one MI355X.

A step is one complete solve of the workload: the panels set (the statics recomputed), then every
frequency of the sweep (boundary values, matrices, right-hand sides S U, LU with the right-hand
sides carried along, back substitution, forces). `value` is frequency solves per second over the
timed steps, the inputs resident in HBM. The dominant kernel's roofline is measured live with HIP
events on the library's stream during the timed steps; the CPU baseline is the FP64 oracle (a
C++/OpenMP restatement of the reference's algorithm) timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cylinder8192fd] [--paths fp64,mixed]
                    [--batch B] [--no-cpu-baseline]
    python bench.py --workload batching4    (uneven geometries through one group solve)
    python bench.py --workload variants8    (eight hull variants, four frequencies each, one group solve)

By default four paths run on the 8192-panel finite-depth cylinder (the workload the metric is
quoted on): plain FP64, mixed (complex64 LU + FP64 refinement), mixed+NN (the learned surrogate in
the assembly) and mixed32 (the FP32 far-field assembly with FP64 near entries, complex64 LU, FP64
refinement). `value` is the throughput of the fastest path whose A, B and X stay within 1e-5
of the plain FP64 GPU results (which match the FP64 oracle to 1e-10 in the tests); every path's
throughput, stage times, speed-up and errors are in `config`.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "freq solves/s @8192 panels: mixed+NN vs plain FP64 GPU, A/B/X rel err<=1e-5"
PEAK_FP64 = 78.6e12  # MI355X dense FP64 (matrix and vector), FLOP/s
PEAK_FP32 = 157.3e12  # MI355X dense FP32 matrix, FLOP/s
PEAK_FP16 = 2.5e15    # MI355X dense FP16 matrix, FLOP/s
# LU GEMM variants: tiles, kernel, arithmetic of the products
VARIANTS = {1: ("64x64", "gemm_mfma_kernel<T, 64, 64>", None), 2: ("128x64", "gemm_mfma_kernel<T, 128, 64>", None),
            3: ("64x64 split-fp16", "gemm_f16s_kernel<64, 64>", "f16x3"),
            4: ("128x64 split-fp16", "gemm_f16s_kernel<128, 64>", "f16x3"),
            5: ("128x128 split-fp16", "gemm_f16s_kernel<128, 128, 2, 4>", "f16x3"),
            6: ("128x128 split-fp16 direct-LDS", "gemm_f16s_big_kernel<false>", "f16x3"),
            7: ("128x128 split-fp16 direct-LDS double-buffered", "gemm_f16s_big_kernel<true>", "f16x3")}
# HBM bytes per launch of the path's GEMM kernel, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# calibrated by a device copy of known size (profiles/gemm_hbm_traffic_cylinder8192_finite_depth.md):
# (workload, path, GEMM tiles, panels per trailing update, batch) -> bytes
TRAFFIC = {("cylinder8192fd", "mixed", "64x64 split-fp16", 8, (0, 0), 16): 8.496e8,
           ("cylinder8192fd", "mixed", "64x64 split-fp16", 8, (0, 0), 48): 2.548e9,
           ("cylinder8192fd", "mixed", "64x64 split-fp16", 16, (0, 0), 48): 2.852e9,
           # the 128 x 128 kernel, averaged over the trailing updates of a solve
           ("cylinder8192fd", "mixed32", "128x128 split-fp16 direct-LDS", 16, (0, 0), 48): 5.637e10,
           # ... and, with sub-blocks, over its trailing and L-shaped updates
           ("cylinder8192fd", "mixed32", "128x128 split-fp16 direct-LDS", 32, (16, 8), 48): 1.320e10}
# HBM bytes per launch (one frequency) of the assembly kernel, from the same calibrated runs:
# (workload, path) -> bytes (profiles/assembly_hbm_traffic_cylinder8192_finite_depth.md)
# (mixed32: the assembly kernel and the copy of A after it, profiles/assemble_fin32_counters_*.md)
ASM_TRAFFIC = {("cylinder8192fd", "fp64"): 8.97e9, ("cylinder8192fd", "mixed"): 8.97e9,
               ("cylinder8192fd", "mixed32"): 5.08e9}
PATH_DESC = {"fp64": "plain FP64: complex128 assembly + LU (MFMA FP64), partial-pivoting check with tournament fallback",
             "mixed": "mixed: FP64 assembly, complex64 LU on the matrix cores (FP32 MFMA, or FP16 MFMA with FP32 "
                      "operands split into hi + lo, autotuned), FP64 iterative refinement to backward error 1e-12",
             "mixed32": "mixed32: FP32 far-field assembly (FP32 tables and statics, self and adjacent pairs in FP64), "
                        "complex64 LU on the matrix cores, FP64 iterative refinement against that matrix to "
                        "backward error 1e-10",
             "mixed_nn": "mixed+NN: finite-depth evanescent remainder from the BF16-MFMA surrogate (FP64 evaluator for "
                         "R/h < 0.1), complex64 LU as the mixed path, FP64 iterative refinement to backward error 1e-12"}
DTYPE = {"fp64": "f64", "mixed": "f32+f64", "mixed32": "f32+f64", "mixed_nn": "bf16+f32+f64"}
# the library precision of each path
PRECISION = {"fp64": "fp64", "mixed": "mixed", "mixed32": "mixed32", "mixed_nn": "mixed"}
PEAK_HBM = 8.0e12    # MI355X HBM3E, B/s


def workload(name):
    from hydrobem_amd import meshes

    if name == "cylinder4096":
        # truncated vertical cylinder R = 1 m, draft 1 m, infinite depth, 4096 panels (side 64 x 32,
        # bottom 64 x 32, cosine spacing), 32 frequencies 0.3..3.0 rad/s, headings 0/45/90/135
        v = meshes.cylinder_panels(1.0, 1.0, 64, 32, 32, quarter=False, cosine=True)
        return dict(name=name, vertices=v, omegas=np.linspace(0.3, 3.0, 32), headings=[0.0, 45.0, 90.0, 135.0],
                    rho=1025.0, g=9.81, desc="truncated cylinder R=1 draft=1, 4096 panels, infinite depth, "
                                              "32 omegas 0.3-3.0 rad/s, 4 headings (10 RHS), complex128")
    if name == "cylinder8192":
        v = meshes.cylinder_panels(1.0, 1.0, 128, 32, 32, quarter=False, cosine=True)
        return dict(name=name, vertices=v, omegas=np.linspace(0.3, 3.0, 8), headings=[0.0],
                    rho=1025.0, g=9.81, desc="truncated cylinder R=1 draft=1, 8192 panels, infinite depth, "
                                              "8 omegas 0.3-3.0 rad/s, 1 heading (7 RHS), complex128")
    if name == "cylinder8192fd":
        # truncated cylinder R = 1 m, draft 1 m, finite depth h = 3 m, 8192 panels (side 128 x 32,
        # bottom 128 x 32, cosine spacing), 48 frequencies nu h log-spaced 0.05..20, heading 0 (7 RHS)
        v = meshes.cylinder_panels(1.0, 1.0, 128, 32, 32, quarter=False, cosine=True)
        nu = np.logspace(np.log10(0.05), np.log10(20.0), 48) / 3.0
        return dict(name=name, vertices=v, omegas=np.sqrt(9.81 * nu), headings=[0.0], rho=1025.0, g=9.81,
                    quadrant=meshes.cylinder_panels(1.0, 1.0, 32, 32, 32, quarter=True, cosine=True),
                    depth=3.0, desc="truncated cylinder R=1 draft=1, finite depth h=3, 8192 panels, "
                                    "48 frequencies nu*h log-spaced 0.05-20, heading 0 (7 RHS), complex128")
    if name == "barge12288":
        # rectangular barge 80 x 20 m, draft 5 m, infinite depth, 12288 quads of about 0.46 m
        # (bottom 168 x 43 = 7224, sides 2 x 168 x 12 = 4032 and 2 x 43 x 12 = 1032), 24
        # frequencies 0.2..2.0 rad/s, 8 headings 0..157.5 deg (14 right-hand sides)
        v = meshes.box_panels(40.0, 10.0, 5.0, (168, 43, 12), quarter=False)
        return dict(name=name, vertices=v, omegas=np.linspace(0.2, 2.0, 24), headings=list(np.arange(8) * 22.5),
                    rho=1025.0, g=9.81, desc="rectangular barge 80 x 20 m, draft 5 m, infinite depth, 12288 panels, "
                                              "24 omegas 0.2-2.0 rad/s, 8 headings (14 RHS), complex128")
    if name == "hemisphere512":
        v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
        return dict(name=name, vertices=v, omegas=np.sqrt(9.81 * np.linspace(0.1, 3.0, 16)), headings=[0.0],
                    rho=1025.0, g=9.81, desc="floating hemisphere R=1, 512 panels, 16 kR 0.1-3.0, heading 0")
    raise SystemExit(f"unknown workload {name}")


def bench_group(a):
    """The batching configuration: truncated cylinders R = 1.0 / 0.75 / 1.5 m (draft 1 m) of 1536 / 2816 /
    6144 panels as separate bodies plus their array at 4 m spacing (10496 panels, three bodies),
    infinite depth, 64 frequencies kR 0.05-4.0 (R = 1 m), heading 0: 256 (N, omega) systems of uneven
    order, solved by one group solve (mixed precision; their systems share the batched LU) and, for
    comparison, by a solve of each geometry on its own. `value`: systems per second of the group
    solve."""
    import hydrobem_amd as hb
    from hydrobem_amd import meshes

    if a.workload == "variants8":
        # a design sweep: eight hull variants (truncated cylinders R = 0.8 .. 1.5 m, draft 1 m, 1536
        # panels each), four frequencies each: separately, every solve is a batch of four systems
        geoms = [(meshes.cylinder_panels(R, 1.0, 48, 16, 16, quarter=False, cosine=True), None, None)
                 for R in np.linspace(0.8, 1.5, 8)]
        omegas = np.sqrt(9.81 * np.array([0.5, 1.0, 1.5, 2.0]))
        desc = ("8 truncated cylinders R=0.8..1.5 draft 1 (1536 panels each), infinite depth, 4 kR 0.5-2.0, "
                "heading 0, mixed precision")
    else:
        bodies = [meshes.cylinder_panels(R, 1.0, *d, quarter=False, cosine=True)
                  for R, d in ((1.0, (48, 16, 16)), (0.75, (64, 22, 22)), (1.5, (96, 32, 32)))]
        shifts = [np.array([4.0 * i, 0.0, 0.0]) for i in range(3)]
        array = np.concatenate([b + sft for b, sft in zip(bodies, shifts)])
        body = np.concatenate([np.full(len(b), i, np.int32) for i, b in enumerate(bodies)])
        geoms = [(b, None, None) for b in bodies] + [(array, body, np.array(shifts))]
        omegas = np.sqrt(9.81 * np.linspace(0.05, 4.0, 64))
        desc = ("cylinders R=1.0/0.75/1.5 draft 1 (1536/2816/6144 panels) + their array (10496 panels, 3 bodies), "
                "infinite depth, 64 kR 0.05-4.0, heading 0, mixed precision")
    heads = [0.0]
    sessions = []
    for v, bd, rc in geoms:
        s = hb.Session(0)
        s.set_environment(1025.0, 9.81)
        s.set_panels(v, bd, rc)
        s.set_batch(a.batch)
        sessions.append(s)
    s0 = sessions[0]
    s0.set_tuning(s0.TUNE_LU_GEMM_MIXED, a.gemm if a.gemm > 0 else 3)
    s0.set_tuning(s0.TUNE_LU_DEPTH_MIXED, a.depth if a.depth > 0 else 8)
    print(f"[bench] batching: panels {[s.n_panels for s in sessions]}, {len(omegas)} frequencies each",
          file=sys.stderr, flush=True)
    # warm-up: statics, buffers and the LU graphs of both forms (--group-form: one form alone, for
    # a profiler's kernel statistics of that form)
    do_group, do_sep = a.group_form in ("both", "group"), a.group_form in ("both", "separate")
    group = sep = None
    if do_group:
        group = hb.solve_group(sessions, omegas, heads, precision="mixed")
    if do_sep:
        sep = [s.solve(omegas, heads, precision="mixed") for s in sessions]
    t_group, t_sep = [float("nan")], [float("nan")]
    for _ in range(max(a.steps, 1)):
        if do_group:
            t0 = time.perf_counter()
            group = hb.solve_group(sessions, omegas, heads, precision="mixed")
            t_group.append(time.perf_counter() - t0)
        if do_sep:
            t0 = time.perf_counter()
            sep = [s.solve(omegas, heads, precision="mixed") for s in sessions]
            t_sep.append(time.perf_counter() - t0)
        print(f"[bench] group {t_group[-1] * 1e3:.1f} ms, separate {t_sep[-1] * 1e3:.1f} ms", file=sys.stderr,
              flush=True)
    t_group = min(t_group[1:]) if do_group else float("nan")
    t_sep = min(t_sep[1:]) if do_sep else float("nan")
    err = max(rel_err(g[k], r[k]) for g, r in zip(group, sep)
              for k in ("added_mass", "radiation_damping", "excitation_diffraction")) if group and sep else None
    systems = len(omegas) * len(sessions)
    line = {
        "metric": f"(N, omega) systems solved/s, {len(sessions)} geometries in one group solve",
        "value": round(systems / (t_group if do_group else t_sep), 3), "unit": "systems/s", "n_gpus": 1,
        "steps": max(a.steps, 1),
        "warmup": 1,
        "ms_per_step": round(1e3 * (t_group if do_group else t_sep), 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16x3+f32+f64", "data": "synthetic: procedural meshes (tests/python/meshes.py layout)",
        "config": {"workload": desc, "timing": "fastest of the steps, group and separate solves alternating",
                   "panels": [s.n_panels for s in sessions], "systems": systems,
                   "group_s": round(t_group, 4) if do_group else None,
                   "separate_s": round(t_sep, 4) if do_sep else None,
                   "group_speedup_vs_separate": round(t_sep / t_group, 3) if do_group and do_sep else None,
                   "rel_err_group_vs_separate": err},
    }
    for s in sessions:
        s.close()
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(w):
    """The FP64 oracle (C++/OpenMP restatement of the reference's algorithm) on this host's cores, on a
    bounded sample: the workload's body meshed with 512 and with 1024 panels (same depth, the middle
    frequency of the sweep), statics, assembly and LU solve timed separately; the 1024-panel times
    scaled to the workload's panel count (statics and assembly by N^2, the LU by N^3, the two sample
    sizes showing the N^2 of the assembly); frequencies/s of the whole sweep with the statics
    amortized over it. Beside it, measured at the workload's full order: the complex64 LU solve of
    the reference's production path as this host's LAPACK (numpy's) runs it on the same cores."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from hydrobem_amd import meshes
    from oracle import oracle as orc

    orc.lib()
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    N = w["vertices"].shape[0]
    depth = w.get("depth", float("inf"))
    mid = float(w["omegas"][len(w["omegas"]) // 2])
    R = 6 + len(w["headings"])
    times = {}
    for nz in (8, 16):
        vs = meshes.cylinder_panels(1.0, 1.0, 32, nz, nz, quarter=False, cosine=True)
        Ns = vs.shape[0]
        P = orc.Problem(vs, rho=w["rho"], grav=w["g"], depth=depth)
        t0 = time.perf_counter()
        P.statics()
        t1 = time.perf_counter()
        S, A = P.assemble(mid)
        t2 = time.perf_counter()
        orc.lu_solve(A, S @ np.ones((Ns, R), complex))
        t3 = time.perf_counter()
        times[Ns] = (t1 - t0, t2 - t1, t3 - t2)
        print(f"[bench] cpu baseline: oracle at {Ns} panels: statics {t1 - t0:.2f} s, assembly {t2 - t1:.2f} s, "
              f"LU {t3 - t2:.2f} s", file=sys.stderr, flush=True)
    (Na, ta), (Ns, (st, asm, lu)) = sorted(times.items())[0], sorted(times.items())[1]
    growth = asm / ta[1]
    s2, s3 = (N / Ns) ** 2, (N / Ns) ** 3
    t_st, t_asm, t_lu = st * s2, asm * s2, lu * s3
    F = len(w["omegas"])
    # the production path's LU at the full order (complex64 LAPACK on the same cores)
    rng = np.random.default_rng(0)
    M = (rng.standard_normal((N, N), dtype=np.float32) + 1j * rng.standard_normal((N, N), dtype=np.float32))
    M = M.astype(np.complex64) / np.float32(np.sqrt(N))
    M[np.diag_indices(N)] += np.complex64(2.0 * np.pi)
    B = np.ones((N, R), np.complex64)
    t4 = time.perf_counter()
    np.linalg.solve(M, B)
    t_lu64 = time.perf_counter() - t4
    del M
    print(f"[bench] cpu baseline: complex64 LAPACK LU solve at {N}: {t_lu64:.2f} s", file=sys.stderr, flush=True)
    return {"value": round(F / (t_st + F * (t_asm + t_lu)), 6), "unit": "freq solves/s", "cores": cores,
            "kind": "port",
            "sample": f"FP64 oracle on {cores} OpenMP threads, the body meshed with {Na} and {Ns} panels (depth "
                      f"{depth}, omega {mid:.3f}): assembly {ta[1]:.2f} s and {asm:.2f} s ({growth:.2f}x for "
                      f"{Ns / Na:.0f}x the panels), at {Ns}: statics {st:.2f} s, LU solve {lu:.2f} s; scaled to "
                      f"{N} panels by N^2 (statics, assembly) and N^3 (LU): {t_st:.0f} s statics + "
                      f"{t_asm + t_lu:.0f} s per frequency. Measured at {N}: the production path's complex64 LU "
                      f"solve by this host's LAPACK on the same cores, {t_lu64:.2f} s. The reference publishes "
                      f"16.4 s (1 thread) / 4.2 s (6 threads) per frequency at 8192 panels for its complex64 "
                      f"production path (docs/theory/13-linear-algebra-and-threads.md:141-144)",
            "assembly_growth_2x_panels": round(growth, 3), "lu_complex64_full_order_s": round(t_lu64, 3)}


def motion_matrices(w):
    """Mass and hydrostatic stiffness of the floating truncated cylinder of the workloads (radius 1 m,
    draft 1 m, centre of gravity at the centre of buoyancy), about the origin on the free surface."""
    rho, g, R, T = w["rho"], w["g"], 1.0, 1.0
    m = rho * np.pi * R ** 2 * T
    zg = -T / 2
    Ixx = m * (R ** 2 / 4 + T ** 2 / 12) + m * zg ** 2
    M = np.diag([m, m, m, Ixx, Ixx, m * R ** 2 / 2])
    M[0, 4] = M[4, 0] = m * zg
    M[1, 3] = M[3, 1] = -m * zg
    C = np.zeros((6, 6))
    C[2, 2] = rho * g * np.pi * R ** 2
    C[3, 3] = C[4, 4] = rho * g * np.pi * R ** 4 / 4
    return M, C


def rel_err(a, b):
    """max |a - b| / scale, the scale of an entry its largest magnitude over the frequencies, at least
    1e-3 of the peak (the reference's regression-test convention)."""
    b = np.asarray(b)
    peak = np.abs(b).max()
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True), 1e-3 * peak)
    return float((np.abs(np.asarray(a) - b) / scale).max())


def run_path(s, w, path, steps, warmup, label):
    """Warm-up steps, `steps` timed steps (stage profiling off, so the LU runs from its captured
    hipGraphs), then one profiled step for the stage times and the GEMM's live roofline."""
    v, omegas, heads = w["vertices"], w["omegas"], w["headings"]
    precision = PRECISION[path]
    surrogate = path == "mixed_nn"

    def step():
        s.set_panels(v)
        return s.solve(omegas, heads, precision=precision, surrogate=surrogate)

    out = None
    for i in range(warmup):
        out = step()
        print(f"[bench] {label}: warmup step {i + 1} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(steps):
        out = step()
        print(f"[bench] {label}: timed step {i + 1} done", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    s.set_profiling(True)
    step()
    prof = s.profile()
    s.set_profiling(False)
    print(f"[bench] {label}: profiled step done", file=sys.stderr, flush=True)
    return out, dt, prof


def summary(prof, dt, steps, F, peak):
    """Throughput of the timed steps; stage times and the GEMM figures of the profiled step."""
    nf = max(prof["frequencies"], 1.0)
    gemm = prof["lu_gemm_flops"] / (prof["lu_gemm_ms"] * 1e-3) if prof["lu_gemm_ms"] > 0 else 0.0
    lu = prof["lu_flops"] / (prof["lu_ms"] * 1e-3) if prof["lu_ms"] > 0 else 0.0
    asm = prof["assembly_bytes"] / (prof["assembly_ms"] * 1e-3) if prof["assembly_ms"] > 0 else 0.0
    return {
        "freq_per_s": round(steps * F / dt, 4),
        "ms_per_step": round(1e3 * dt / steps, 3),
        "stage_ms_per_frequency": {k: round(prof[k] / nf, 4)
                                   for k in ("assembly_ms", "rhs_ms", "lu_ms", "lu_gemm_ms", "mixed_refinement_ms",
                                             "forces_ms")},
        "statics_ms_per_solve": round(prof["statics_ms"] / max(prof["statics_runs"], 1.0), 3),
        "lu_achieved_tflops": round(lu / 1e12, 3),
        "gemm_achieved_tflops": round(gemm / 1e12, 3),
        "gemm_peak_tflops": peak / 1e12,
        "gemm_per_launch": {"flops": prof["lu_gemm_flops"] / max(prof["lu_gemm_launches"], 1.0),
                            "ms": prof["lu_gemm_ms"] / max(prof["lu_gemm_launches"], 1.0)},
        # variants 6 and 7: the trailing updates on the 128 x 128 tiles alone (the panels' strips
        # run on other tile shapes)
        "gemm_big": {"tflops": round(prof["lu_big_gemm_flops"] / (prof["lu_big_gemm_ms"] * 1e-3) / 1e12, 3)
                     if prof["lu_big_gemm_ms"] > 0 else 0.0,
                     "launches": int(prof["lu_big_gemm_launches"]),
                     "per_launch": {"flops": prof["lu_big_gemm_flops"] / max(prof["lu_big_gemm_launches"], 1.0),
                                    "ms": prof["lu_big_gemm_ms"] / max(prof["lu_big_gemm_launches"], 1.0)}},
        "first_solution_backward_error": prof["mixed_max_first_backward_error"],
        "lu_pivoting_fallback_panels": int(prof["lu_fallback_panels"]),
        "assembly_achieved_gbs": round(asm / 1e9, 1),
        "refinement_steps_per_frequency": round(prof["mixed_refinement_steps"] / nf, 3),
        "max_backward_error": prof["mixed_max_backward_error"],
        "fp64_fallback_frequencies": int(prof["mixed_fp64_fallbacks"]),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cylinder8192fd")
    ap.add_argument("--batch", type=int, default=0, help="frequencies in flight (0: automatic)")
    ap.add_argument("--gemm", type=int, default=-1, help="LU GEMM variant (-1: the library's default)")
    ap.add_argument("--depth", type=int, default=-1, help="LU panels per trailing update (-1: the library's default)")
    ap.add_argument("--sub", type=int, default=-1, help="LU panels per sub-block of a block (-1: the library's default)")
    ap.add_argument("--sub2", type=int, default=-1, help="LU panels per sub-block of a sub-block (-1: the default)")
    ap.add_argument("--paths", default="fp64,mixed,mixed_nn,mixed32",
                    help="paths timed; `value` is the fastest whose A, B, X are within 1e-5 of the FP64 path's")
    ap.add_argument("--no-autotune", action="store_true", help="skip the LU autotuner before each path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--group-form", default="both", choices=("both", "group", "separate"),
                    help="batching4 / variants8: time the group solve, the separate solves or both")
    ap.add_argument("--nn-rmin", type=float, default=0.1,
                    help="the surrogate takes the pairs with nn-rmin <= R/h <= 20 (mixed_nn path)")
    ap.add_argument("--nn-sweep", default="",
                    help="comma-separated lower R/h bounds: the mixed_nn path once at each, errors against fp64")
    ap.add_argument("--symmetry", action="store_true",
                    help="solve the body as its quadrant with the symmetry planes x = 0 and y = 0 (not the "
                         "headline workload; the surrogate path is not available there)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only copy this many GiB device to device and exit (the calibration of HBM counters)")
    a = ap.parse_args()
    if a.copy_probe > 0:
        import torch

        n = int(a.copy_probe * (1 << 30)) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}), flush=True)
        return 0
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (the frequency sweep "
              f"would shard across GPUs with no exchange, see the design notes)", file=sys.stderr)
        return 2
    if a.workload in ("batching4", "variants8"):
        return bench_group(a)
    import hydrobem_amd as hb

    w = workload(a.workload)
    wkey = a.workload
    if a.symmetry:
        if "quadrant" not in w:
            raise SystemExit(f"workload {a.workload} has no quadrant mesh")
        nq = w["quadrant"].shape[0]
        w = dict(w, vertices=w["quadrant"], desc=w["desc"] + f"; solved as its {nq}-panel quadrant with the "
                 f"symmetry planes x = 0 and y = 0 (4 character systems of {nq} unknowns per frequency)")
        wkey = a.workload + "_quadrant"
    v, omegas, heads = w["vertices"], w["omegas"], w["headings"]
    N, F = v.shape[0], len(omegas)
    s = hb.Session(0)
    s.set_environment(w["rho"], w["g"], w.get("depth", float("inf")))
    if a.symmetry:
        s.set_symmetry(True, True)
    s.set_batch(a.batch)
    if a.gemm >= 0:
        s.set_tuning(s.TUNE_LU_GEMM, a.gemm)
    if a.depth >= 0:
        s.set_tuning(s.TUNE_LU_DEPTH, a.depth)
    if a.sub >= 0:
        s.set_tuning(s.TUNE_LU_SUB, a.sub)
    if a.sub2 >= 0:
        s.set_tuning(s.TUNE_LU_SUB2, a.sub2)
    s.set_tuning(s.TUNE_NN_RMIN, int(round(100 * a.nn_rmin)))
    print(f"[bench] workload {w['name']}: {N} panels, {F} frequencies", file=sys.stderr, flush=True)
    paths = [p for p in a.paths.split(",") if p and not (a.symmetry and p == "mixed_nn")]
    peaks = {"fp64": PEAK_FP64, "mixed": PEAK_FP32, "mixed32": PEAK_FP32, "mixed_nn": PEAK_FP32}
    res = {}
    if "mixed_nn" in paths:
        from hydrobem_amd import train_surrogate

        if train_surrogate.WEIGHTS.exists():
            config_nn = {"weights": train_surrogate.WEIGHTS.name, "training": train_surrogate.load(s)}
        else:
            print("[bench] no trained surrogate: training one (2e6 samples)", file=sys.stderr, flush=True)
            config_nn = {"weights": "trained in this run",
                         "training": s.surrogate_train(n_train=2_000_000, n_holdout=100_000, epochs=6, batch=4096)}
    tuning = {}
    for p in paths:
        if not a.no_autotune and a.gemm < 0 and a.depth < 0:
            prec = "fp64" if p == "fp64" else "mixed"
            if prec not in tuning:
                rows = s.autotune(N, 6 + len(heads), prec)
                print(f"[bench] autotune {prec} candidates (variant, depth, sub-blocks, batch, ms per system): " +
                      "; ".join(f"{int(r[0])},{int(r[1])},{int(r[2])}/{int(r[3])},{int(r[4])}: {r[5]:.3f}"
                                for r in rows), file=sys.stderr, flush=True)
                best = min(rows, key=lambda r: r[5])
                tuning[prec] = {"gemm_variant": int(best[0]), "gemm_tiles": VARIANTS[int(best[0])][0],
                                "panels_per_update": int(best[1]), "panels_per_sub_block": [int(best[2]), int(best[3])],
                                "batch": int(best[4]), "ms_per_system": round(best[5], 3), "candidates": len(rows)}
                print(f"[bench] autotuned {prec}: {tuning[prec]}", file=sys.stderr, flush=True)
            if a.batch == 0:
                s.set_batch(tuning[prec]["batch"])
        out, dt, prof = run_path(s, w, p, a.steps, a.warmup, p)
        gvar = s.get_tuning(s.TUNE_LU_GEMM if p == "fp64" else s.TUNE_LU_GEMM_MIXED)
        res[p] = (out, summary(prof, dt, a.steps, F, PEAK_FP16 if VARIANTS[gvar][2] else peaks[p]))
        res[p][1]["gemm_variant"] = gvar
    sweep = []
    if a.nn_sweep and "fp64" in res and "mixed_nn" in paths:
        # where the network's pairs end (its lower R/h bound) against what it costs and saves: one
        # timed step at each bound, the errors against the plain FP64 results
        for r in [float(x) for x in a.nn_sweep.split(",") if x]:
            s.set_tuning(s.TUNE_NN_RMIN, int(round(100 * r)))
            o_r, dt_r, prof_r = run_path(s, w, "mixed_nn", 1, 0, f"mixed_nn R/h >= {r}")
            sm = summary(prof_r, dt_r, 1, F, PEAK_FP32)
            sweep.append({"nn_rmin": r, "freq_per_s": sm["freq_per_s"],
                          "assembly_ms": sm["stage_ms_per_frequency"]["assembly_ms"],
                          "rel_err_vs_fp64_gpu": {k: rel_err(o_r[k], res["fp64"][0][k]) for k in
                                                  ("added_mass", "radiation_damping", "excitation_diffraction")}})
            print(f"[bench] nn sweep: {sweep[-1]}", file=sys.stderr, flush=True)
        s.set_tuning(s.TUNE_NN_RMIN, int(round(100 * a.nn_rmin)))
    s.close()
    # `value`: the fastest path whose A, B and X stay within 1e-5 of the plain FP64 results (which
    # match the FP64 oracle to 1e-10); the others are reported beside it with their errors
    ok = [p for p in paths if p == "fp64" or ("fp64" in res and
          max(rel_err(res[p][0][k], res["fp64"][0][k]) for k in ("added_mass", "radiation_damping",
                                                                  "excitation_diffraction")) <= 1e-5)]
    main_path = max(ok, key=lambda p: res[p][1]["freq_per_s"]) if ok else paths[-1]
    out, summ = res[main_path]
    config = {"workload": w["desc"], "panels": N, "frequencies": F, "rhs": 6 + len(heads),
              "path": PATH_DESC[main_path], "value_path": main_path,
              "value_rule": "fastest path with A, B, X within 1e-5 of the plain FP64 GPU results",
              "batch": a.batch or "auto"}
    if a.symmetry:
        config["symmetry_planes"] = "x = 0 and y = 0"
        config["panels_whole_body"] = 4 * N
    for p in paths:
        config[p] = res[p][1]
    if "mixed_nn" in paths:
        config["surrogate"] = config_nn
        config["surrogate"]["nn_rmin"] = a.nn_rmin
    if sweep:
        config["mixed_nn_rmin_sweep"] = sweep
    if tuning:
        config["autotuned_lu"] = tuning
    if "fp64" in res and len(paths) > 1:
        ref = res["fp64"][0]
        M, C = motion_matrices(w)
        s2 = hb.Session(0)
        rao_ref = s2.motions(ref, M, C)
        for p in [q for q in paths if q != "fp64"]:
            config[p]["rel_err_vs_fp64_gpu"] = {k: rel_err(res[p][0][k], ref[k]) for k in
                                                ("added_mass", "radiation_damping", "excitation_diffraction")}
            config[p]["rel_err_vs_fp64_gpu"]["rao"] = rel_err(s2.motions(res[p][0], M, C), rao_ref)
            config[p]["speedup_vs_fp64_gpu"] = round(res[p][1]["freq_per_s"] / res["fp64"][1]["freq_per_s"], 3)
        s2.close()
    gemm_tf, gemm_launch = summ["gemm_achieved_tflops"], summ["gemm_per_launch"]
    if summ.get("gemm_big", {}).get("launches", 0) > 0:
        # variants 6 and 7: the dominant kernel is the 128 x 128 one of the trailing updates
        gemm_tf, gemm_launch = summ["gemm_big"]["tflops"], summ["gemm_big"]["per_launch"]
    # the LU settings the value path ran with: autotuned, or given on the command line, or the defaults
    gv = summ["gemm_variant"]
    tun = tuning.get("fp64" if main_path == "fp64" else "mixed",
                     {"gemm_variant": gv, "gemm_tiles": VARIANTS[gv][0],
                      "panels_per_update": a.depth if a.depth >= 0 else 4,
                      "panels_per_sub_block": [max(a.sub, 0), max(a.sub2, 0)],
                      "batch": a.batch or 48})
    traffic = TRAFFIC.get((wkey, main_path, tun["gemm_tiles"], tun["panels_per_update"],
                           tuple(tun.get("panels_per_sub_block", (0, 0))), tun["batch"]))
    tiles, kernel, split = VARIANTS[tun["gemm_variant"]]
    kernel = kernel.replace("<T,", "<double," if main_path == "fp64" else "<float,")
    if tun["gemm_variant"] >= 6:
        kernel = kernel.replace("gemm_f16s_big_kernel<", "gemm_f16s_big_kernel<128, 128, ")
    # the split-FP16 GEMM is priced against the FP16 matrix peak (each FP32 product is three FP16
    # products, so its FP32-equivalent ceiling is a third of that)
    gpeak = PEAK_FP16 if split else peaks[main_path]
    line = {
        "metric": METRIC,
        "value": summ["freq_per_s"],
        "unit": "freq solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": summ["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f16x3+" if split else "") + DTYPE[main_path],
        "data": "synthetic: procedural mesh of the workload (tests/python/meshes.py layout)",
        "config": config,
        "roofline": {"bound": "mfma", "kernel": kernel + " (LU trailing and sub-block updates, complex GEMM on MFMA"
                                                + (", FP32 operands split into FP16 hi + lo)" if split else ")"),
                     "achieved": gemm_tf, "peak": gpeak / 1e12, "unit": "TFLOP/s",
                     "frac": round(gemm_tf * 1e12 / gpeak, 4), "traffic": traffic,
                     "per_launch": gemm_launch},
        # the other large kernel of the step: the assembly (one launch per frequency), bound by
        # vector issue (FP64, or FP32 for mixed32), not HBM; its algorithmic bytes are N^2 (32 + 2 x 16)
        # (statics read, S and A written), for mixed32 N^2 (16 + 3 x 8) (FP32 statics read, S, A and
        # A's unfactored copy written, the copy by a copy after the kernel), against the stage time
        # (boundary values, tables, matrices, the copy)
        "roofline_assembly": {"bound": "hbm", "kernel": ("assemble_fin32_kernel" if main_path == "mixed32" else
                                                         "assemble_fin_kernel") if np.isfinite(w.get("depth", np.inf))
                              else ("assemble32_kernel" if main_path == "mixed32" else "assemble_kernel"),
                              "achieved": summ["assembly_achieved_gbs"],
                              "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                              "frac": round(summ["assembly_achieved_gbs"] * 1e9 / PEAK_HBM, 4),
                              "traffic": ASM_TRAFFIC.get((wkey, main_path)),
                              "ms_per_launch": summ["stage_ms_per_frequency"]["assembly_ms"]},
    }
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
