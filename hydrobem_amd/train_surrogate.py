#!/usr/bin/env python3
"""Train the finite-depth surrogate on the GPU and write its parameters.

    python -m hydrobem_amd.train_surrogate [--samples 20000000] [--epochs 8] [--batch 8192]
                                           [--lr 2e-3] [--seed 1] [--out hydrobem_amd/surrogate_weights.npz]

Sampling, the FP64 evaluator's targets, the forward and backward passes and Adam all run on the
GPU (hbg_surrogate_train). The output holds the FP32 parameters (`params`, the C ABI's layout)
and the training report; `load()` gives a session the trained network.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

WEIGHTS = Path(__file__).resolve().parent / "surrogate_weights.npz"


def load(session, path=WEIGHTS) -> dict:
    """Give the session the trained network; returns the training report stored with it."""
    with np.load(path, allow_pickle=False) as f:
        session.surrogate_set(f["params"])
        return json.loads(str(f["report"]))


def main(argv=None) -> int:
    import hydrobem_amd as hb

    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20_000_000)
    ap.add_argument("--holdout", type=int, default=200_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=str(WEIGHTS))
    a = ap.parse_args(argv)
    s = hb.Session(0)
    t0 = time.perf_counter()
    rep = s.surrogate_train(n_train=a.samples, n_holdout=a.holdout, epochs=a.epochs, batch=a.batch,
                            learning_rate=a.lr, seed=a.seed)
    rep.update({"samples": a.samples, "holdout": a.holdout, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
                "seed": a.seed, "wall_seconds": time.perf_counter() - t0})
    params = s.surrogate_get()
    s.close()
    np.savez(a.out, params=params, report=json.dumps(rep))
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
