"""Writes tests/golden/meshes_layout.npz: the panel layout of the reference's procedural test
meshes (tests/python/meshes.py of the reference: box_panels, hemisphere_panels, cylinder_panels)
for the parameter sets the solve's tests and workloads use.

Run in the build container only (it imports the reference's meshes.py from /root/reference):
    python tests/golden/write_meshes.py
Small meshes are stored as their vertex arrays [P][4][3]; the large workload meshes as their
shape and the SHA-256 of their float64 bytes (C order), which a bitwise reproduction must match.
"""

from __future__ import annotations

import hashlib
import importlib.util
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/tests/python/meshes.py")
OUT = Path(__file__).resolve().parent / "meshes_layout.npz"

# name -> (generator, positional args, keyword args); "full": vertex arrays, "hash": shape + SHA-256
CASES = {
    "hemi_16x32": ("hemisphere_panels", (1.0, 16, 32), {"quarter": False}, "full"),
    "hemi_6x24": ("hemisphere_panels", (1.0, 6, 24), {"quarter": False}, "full"),
    "hemi_4x8": ("hemisphere_panels", (1.0, 4, 8), {"quarter": False}, "full"),
    "hemi_q_6x7": ("hemisphere_panels", (1.0, 6, 7), {"quarter": True}, "full"),
    "hemi_q_3x3": ("hemisphere_panels", (1.0, 3, 3), {"quarter": True}, "full"),
    "hemi_q_graded": ("hemisphere_panels", (0.8, 5, 6), {"quarter": True, "graded": True}, "full"),
    "cyl_drift": ("cylinder_panels", (1.0, 0.5), {"n_phi": 32, "n_z": 4, "n_r": 4, "quarter": False}, "full"),
    "cyl_quarter": ("cylinder_panels", (1.0, 0.5), {"n_phi": 8, "n_z": 4, "n_r": 4, "quarter": True}, "full"),
    "cyl_q_odd": ("cylinder_panels", (1.0, 0.8, 9, 4, 3), {"quarter": True}, "full"),
    "cyl_cos_16": ("cylinder_panels", (1.0, 1.0, 16, 8, 8), {"quarter": False, "cosine": True}, "full"),
    "cyl_cos_q": ("cylinder_panels", (1.0, 1.0, 16, 12, 8), {"quarter": True, "cosine": True}, "full"),
    "box_q": ("box_panels", (1.0, 0.5, 0.5), {"n": (6, 3, 3), "quarter": True}, "full"),
    "box_whole": ("box_panels", (1.0, 0.6, 0.5, (6, 4, 3)), {"quarter": False}, "full"),
    "box_lid": ("box_panels", (1.0, 0.5, 0.5, (4, 2, 2)), {"quarter": False, "lid": True}, "full"),
    "cyl_1536": ("cylinder_panels", (1.0, 1.0, 48, 16, 16), {"quarter": False, "cosine": True}, "hash"),
    "cyl_2816": ("cylinder_panels", (0.75, 1.0, 64, 22, 22), {"quarter": False, "cosine": True}, "hash"),
    "cyl_4096": ("cylinder_panels", (1.0, 1.0, 64, 32, 32), {"quarter": False, "cosine": True}, "hash"),
    "cyl_6144": ("cylinder_panels", (1.5, 1.0, 96, 32, 32), {"quarter": False, "cosine": True}, "hash"),
    "cyl_8192": ("cylinder_panels", (1.0, 1.0, 128, 32, 32), {"quarter": False, "cosine": True}, "hash"),
    "box_12288": ("box_panels", (40, 10, 5), {"n": (168, 43, 12)}, "hash"),
}


def digest(v: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(v, dtype=np.float64).tobytes()).hexdigest()


def main() -> None:
    spec = importlib.util.spec_from_file_location("reference_meshes", REF)
    ref = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ref)
    out = {}
    for name, (gen, args, kw, kind) in CASES.items():
        v = np.asarray(getattr(ref, gen)(*args, **kw), dtype=np.float64)
        if kind == "full":
            out[name] = v
        else:
            out[name + "__shape"] = np.array(v.shape, dtype=np.int64)
            out[name + "__sha256"] = np.array(digest(v))
        print(f"{name}: {v.shape}", file=sys.stderr)
    np.savez_compressed(OUT, **out)


if __name__ == "__main__":
    main()
