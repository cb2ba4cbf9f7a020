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
# ops_halo_transfer cases for oracle/refdrv/ref_halo.cpp.  dats: name, type,
# components, size, d_m, d_p; groups: lists of halos (from, to, iter_size,
# from_base, to_base, from_dir, to_dir), transferred in order.  "mblock" is
# the five halo groups of the reference's multi-block app
# (apps/c/mblock/mblock.cpp:60-157) on its 20 x 20 dats with depth-2 halos.
_MB = (1, [20, 20], [-2, -2], [2, 2])
HALO_CASES = [
    {"name": "mblock", "ndim": 2,
     "dats": [["data0", "double", *_MB], ["data1", "double", *_MB]],
     "groups": [
         [["data0", "data1", [2, 20], [18, 0], [-2, 0], [1, 2], [1, 2]],
          ["data1", "data0", [2, 20], [0, 0], [20, 0], [1, 2], [1, 2]]],
         [["data0", "data1", [20, 2], [0, 18], [0, -2], [1, 2], [1, 2]],
          ["data1", "data0", [20, 2], [0, 0], [0, 20], [1, 2], [1, 2]]],
         [["data0", "data1", [2, 20], [0, 0], [20, 0], [1, 2], [1, -2]],
          ["data1", "data0", [2, 20], [18, 0], [-2, 0], [1, -2], [1, 2]]],
         [["data0", "data1", [20, 2], [0, 0], [0, 20], [1, 2], [-1, 2]],
          ["data1", "data0", [20, 2], [0, 18], [0, -2], [-1, 2], [1, 2]]],
         [["data0", "data1", [2, 20], [18, 0], [0, -2], [1, 2], [2, 1]],
          ["data1", "data0", [2, 20], [0, 0], [20, 0], [2, 1], [1, 2]]],
     ]},
    # 3D, unequal halos, permuted and flipped axes; then a self-transfer
    # whose source and target rows overlap (all values are packed first)
    {"name": "perm3d", "ndim": 3,
     "dats": [["A", "double", 1, [7, 5, 4], [-1, -2, -1], [2, 1, 1]],
              ["B", "double", 1, [5, 4, 7], [-2, -1, -1], [1, 1, 2]]],
     "groups": [
         [["A", "B", [3, 2, 4], [2, 1, 0], [-2, 0, 4], [1, 2, 3], [2, 3, -1]],
          ["B", "A", [2, 3, 2], [0, -1, 5], [-1, -2, 2], [-1, 2, -3], [3, 1, 2]]],
         [["A", "A", [7, 1, 1], [0, 0, 0], [0, -1, 0], [1, 2, 3], [-1, 2, 3]],
          ["A", "A", [7, 2, 4], [0, 0, 0], [0, 1, 0], [1, 2, 3], [1, 2, 3]]],
     ]},
    # two components per point
    {"name": "multi", "ndim": 2,
     "dats": [["P", "double", 2, [6, 5], [-1, -1], [1, 1]], ["Q", "double", 2, [6, 5], [-1, -1], [1, 1]]],
     "groups": [
         [["P", "Q", [1, 5], [5, 0], [-1, 0], [1, 2], [1, -2]],
          ["Q", "P", [6, 1], [0, 0], [0, 5], [1, 2], [-1, 2]]],
     ]},
    # float and double dats meeting: values pass through float
    {"name": "mixed", "ndim": 2,
     "dats": [["D", "double", 1, [8, 6], [-1, -1], [1, 1]], ["F", "float", 1, [8, 6], [-1, -1], [1, 1]]],
     "groups": [
         [["D", "F", [8, 1], [0, 5], [0, -1], [1, 2], [1, 2]],
          ["F", "D", [1, 6], [7, 0], [8, 0], [1, 2], [1, -2]]],
         [["D", "F", [2, 3], [1, 1], [3, 2], [1, 2], [2, 1]]],
     ]},
    # rows of more than 1 KiB (x-padded layouts), x of one dat against y of
    # the other, flipped on the way back
    {"name": "wide", "ndim": 2,
     "dats": [["W1", "double", 1, [140, 3], [-1, -1], [1, 1]], ["W2", "double", 1, [3, 140], [-1, -1], [1, 1]]],
     "groups": [
         [["W1", "W2", [140, 1], [0, 2], [-1, 0], [1, 2], [2, 1]],
          ["W2", "W1", [1, 140], [2, 0], [0, 3], [1, -2], [2, 1]]],
     ]},
]
META["halo"] = HALO_CASES


def halo_spec_text(cases):
    """The cases in the line format oracle/refdrv/ref_halo.cpp reads."""
    pad3 = lambda v, fill: list(v) + [fill] * (3 - len(v))  # noqa: E731
    out = []
    for c in cases:
        out.append(f"case {c['name']} {c['ndim']}")
        for name, typ, dim, size, dm, dp in c["dats"]:
            out.append(" ".join(map(str, ["dat", name, typ, dim, *pad3(size, 1), *pad3(dm, 0), *pad3(dp, 0)])))
        for grp in c["groups"]:
            out.append("group")
            for f, t, it, fb, tb, fd, td in grp:
                out.append(" ".join(map(str, ["halo", f, t, *pad3(it, 1), *pad3(fb, 0), *pad3(tb, 0),
                                              *pad3(fd, 3), *pad3(td, 3)])))
        out.append("end")
    return "\n".join(out) + "\n"


def halo_arrays(tmp):
    """Runs ref_halo on HALO_CASES; returns {halo_<case>_<dat>_g<k>: array}."""
    spec = os.path.join(tmp, "halo_spec.txt")
    with open(spec, "w") as f:
        f.write(halo_spec_text(HALO_CASES))
    hdir = os.path.join(tmp, "halo")
    os.makedirs(hdir)
    subprocess.run([os.path.join(ROOT, "oracle/_ref/ref_halo"), spec, hdir], check=True)
    types = {(c["name"], d[0]): d[1] for c in HALO_CASES for d in c["dats"]}
    arrays = {}
    for fn in sorted(os.listdir(hdir)):
        key = fn[:-4]
        case, dat = key.split("_")[1:3]
        dt = np.float32 if types[(case, dat)] == "float" else np.float64
        arrays[key] = np.fromfile(os.path.join(hdir, fn), dtype=dt)
    return arrays


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
        arrays.update(halo_arrays(tmp))
    arrays["meta_json"] = np.frombuffer(json.dumps(META).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, "with", len(arrays), "arrays")


if __name__ == "__main__":
    sys.exit(main())
