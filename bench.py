#!/usr/bin/env python3
"""Throughput of the low-order frequency-domain solve on one MI355X.

A step is one complete solve of the workload: the panels set (the statics recomputed), then every
frequency of the sweep (boundary values, matrices, right-hand sides S U, LU with the right-hand
sides carried along, back substitution, forces). `value` is frequency solves per second over the
timed steps, the inputs resident in HBM. The dominant kernel's roofline is measured live with HIP
events on the library's stream during the timed steps; the CPU baseline is the FP64 oracle (a
C++/OpenMP restatement of the reference's algorithm) timed on this host on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cylinder4096] [--no-cpu-baseline]
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
                    depth=3.0, desc="truncated cylinder R=1 draft=1, finite depth h=3, 8192 panels, "
                                    "48 frequencies nu*h log-spaced 0.05-20, heading 0 (7 RHS), complex128")
    if name == "hemisphere512":
        v = meshes.hemisphere_panels(1.0, 16, 32, quarter=False)
        return dict(name=name, vertices=v, omegas=np.sqrt(9.81 * np.linspace(0.1, 3.0, 16)), headings=[0.0],
                    rho=1025.0, g=9.81, desc="floating hemisphere R=1, 512 panels, 16 kR 0.1-3.0, heading 0")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(w, budget_s=30.0):
    """The FP64 oracle on this host: statics once, then one frequency of the workload; the
    projected frequencies/s of the whole sweep with the statics amortized over it."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as orc

    orc.lib()
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    v = w["vertices"]
    P = orc.Problem(v, rho=w["rho"], grav=w["g"], depth=w.get("depth", float("inf")))
    t0 = time.perf_counter()
    P.statics()
    t_st = time.perf_counter() - t0
    mid = float(w["omegas"][len(w["omegas"]) // 2])
    t0 = time.perf_counter()
    P.solve([mid], w["headings"])
    t_f = time.perf_counter() - t0
    F = len(w["omegas"])
    return {"value": F / (t_st + F * t_f), "unit": "freq solves/s", "cores": cores, "kind": "port",
            "sample": f"oracle (C++/OpenMP FP64 restatement) on {v.shape[0]} panels: statics {t_st:.2f} s once + "
                      f"1 frequency (omega {mid:.3f}) {t_f:.2f} s, projected to the {F}-frequency sweep"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cylinder4096")
    ap.add_argument("--batch", type=int, default=0, help="frequencies in flight (0: automatic)")
    ap.add_argument("--gemm", type=int, default=-1, help="LU GEMM variant (-1: the library's default)")
    ap.add_argument("--depth", type=int, default=-1, help="LU panels per trailing update (-1: the library's default)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (the frequency sweep "
              f"would shard across GPUs with no exchange, see the design notes)", file=sys.stderr)
        return 2
    import hydrobem_amd as hb

    w = workload(a.workload)
    v, omegas, heads = w["vertices"], w["omegas"], w["headings"]
    N, F = v.shape[0], len(omegas)
    s = hb.Session(0)
    s.set_environment(w["rho"], w["g"], w.get("depth", float("inf")))
    s.set_batch(a.batch)
    if a.gemm >= 0:
        s.set_tuning(s.TUNE_LU_GEMM, a.gemm)
    if a.depth >= 0:
        s.set_tuning(s.TUNE_LU_DEPTH, a.depth)

    def step():
        s.set_panels(v)
        return s.solve(omegas, heads, precision=a.precision)

    print(f"[bench] workload {w['name']}: {N} panels, {F} frequencies; warmup {a.warmup}", file=sys.stderr, flush=True)
    for i in range(a.warmup):
        step()
        print(f"[bench] warmup step {i + 1} done", file=sys.stderr, flush=True)
    s.set_profiling(True)
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        print(f"[bench] timed step {i + 1} done", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    prof = s.profile()
    s.set_profiling(False)
    s.close()
    dt = t1 - t0
    value = a.steps * F / dt
    gemm_tf = prof["lu_gemm_flops"] / (prof["lu_gemm_ms"] * 1e-3) if prof["lu_gemm_ms"] > 0 else 0.0
    lu_tf = prof["lu_flops"] / (prof["lu_ms"] * 1e-3) if prof["lu_ms"] > 0 else 0.0
    asm_bw = prof["assembly_bytes"] / (prof["assembly_ms"] * 1e-3) if prof["assembly_ms"] > 0 else 0.0
    launches = max(prof["lu_gemm_launches"], 1.0)
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "freq solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * dt / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: procedural mesh of the workload (tests/python/meshes.py layout)",
        "config": {"workload": w["desc"], "panels": N, "frequencies": F, "rhs": 6 + len(heads),
                   "path": "plain FP64 (complex128 assembly + tournament-pivoted LU)",
                   "stage_ms_per_frequency": {k: round(prof[k] / max(prof["frequencies"], 1.0), 4)
                                              for k in ("assembly_ms", "rhs_ms", "lu_ms", "lu_gemm_ms", "forces_ms")},
                   "statics_ms_per_solve": round(prof["statics_ms"] / max(prof["statics_runs"], 1.0), 3),
                   "lu_achieved_tflops": round(lu_tf / 1e12, 3),
                   "lu_pivoting_fallback_panels": int(prof["lu_fallback_panels"]),
                   "assembly_achieved_gbs": round(asm_bw / 1e9, 1)},
        "roofline": {"bound": "mfma", "kernel": "gemm_mfma_kernel (LU trailing updates, complex FP64 on MFMA)",
                     "achieved": round(gemm_tf / 1e12, 3), "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": round(gemm_tf / PEAK_FP64, 4), "traffic": None,
                     "per_launch": {"flops": prof["lu_gemm_flops"] / launches,
                                    "ms": prof["lu_gemm_ms"] / launches}},
    }
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
