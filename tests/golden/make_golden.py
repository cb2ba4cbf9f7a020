"""Regenerates tests/golden/ref_seq.npz from the reference's own sequential
OPS library.

Builds libops_seq from the reference sources (oracle/Makefile.ref, outputs in
oracle/_ref/), runs this project's fixture drivers (oracle/refdrv/ref_2d.cpp,
ref_3d.cpp) through it, and packs their raw outputs with the problem sizes.
Needs /root/reference, so it runs in the build container only; the committed
.npz is what the tests read.

    python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "ref_seq.npz")

# problem sizes baked into the drivers (kept here so tests can read them)
META = {
    "seed": 20240611,
    "laplace_top": {"nx": 62, "ny": 46, "iters": 60},
    "laplace_app": {"nx": 38, "ny": 30, "iters": 25},
    "laplace_f32": {"nx": 50, "ny": 40, "iters": 30},
    "laplace_kat": {"nx": 4094, "ny": 4094, "iters": 100, "expected_app": 2.421354960840227e-03},
    "cg2d": {"nx": 30, "ny": 26, "iters": 40, "rx": 1.0, "ry": 1.0},
    "poisson2d": {"nx": 40, "ny": 33, "iters": 40},
    "red": {"nx": 17, "ny": 13},
    "heat3d": {"nx": 24, "ny": 20, "nz": 18, "iters": 50, "rx": 1.0, "ry": 1.0, "rz": 1.0},
    "poisson3d": {"nx": 20, "ny": 18, "nz": 16, "iters": 40},
    "adi": {"nx": 16, "ny": 12, "nz": 10, "lambda": 1.0},
    # the reference's multi-block Poisson app (apps/c/poisson/poisson.cpp),
    # unchanged: -sizex -sizey -iters; its "Total error" line
    "mb_poisson": {"configs": [[20, 20, 10], [97, 61, 40], [64, 50, 200], [33, 47, 25], [200, 130, 100]],
                   "expected_app": 16.5173570176708},
}
F32 = {"laplace_f32_err", "laplace_f32_A", "laplace_f32_init", "poisson3d_f32_w"}
SKIP = {"heat3d_Kz"}


def main():
    subprocess.run(["make", "-s", "-f", "oracle/Makefile.ref", "-j8"], cwd=ROOT, check=True)
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.run([os.path.join(ROOT, "oracle/_ref/ref_2d"), tmp, "kat"], check=True)
        subprocess.run([os.path.join(ROOT, "oracle/_ref/ref_3d"), tmp], check=True)
        arrays = {}
        for fn in sorted(os.listdir(tmp)):
            if not fn.endswith(".bin"):
                continue
            key = fn[:-4]
            if key in SKIP:
                continue
            dt = np.float32 if key in F32 else np.float64
            arrays[key] = np.fromfile(os.path.join(tmp, fn), dtype=dt)
        errs = []
        for nx, ny, it in META["mb_poisson"]["configs"]:
            out = subprocess.run([os.path.join(ROOT, "oracle/_ref/ref_mb_poisson"), f"-sizex={nx}", f"-sizey={ny}",
                                  f"-iters={it}"], cwd=tmp, check=True, capture_output=True, text=True).stdout
            line = [ln for ln in out.splitlines() if ln.startswith("Total error:")][0]
            errs.append(float(line.split(":")[1]))
        arrays["mb_poisson_err"] = np.array(errs)
    arrays["meta_json"] = np.frombuffer(json.dumps(META).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, "with", len(arrays), "arrays")


if __name__ == "__main__":
    sys.exit(main())
