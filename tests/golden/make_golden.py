"""Regenerates the bcsr golden fixtures in this folder.

Text edge lists (*.el) are written here, converted to bcsr by the
reference's own converter (tools/convert_from_txt.cc built into
oracle/_ref/convert_from_txt by `make -C oracle ref`), and the resulting
*.bcsr files are kept as fixtures.  The R-MAT edge lists come from the
oracle's sampler (ltor_rmat_edges) so the fixtures also pin the sampler's
output through the converter.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CONVERT = os.path.join(ROOT, "oracle", "_ref", "convert_from_txt")
ORACLE = os.path.join(ROOT, "oracle", "build", "liblt_oracle.so")

# (name, scale, edge_factor, seed) — a,b,c = .57/.19/.19
RMAT_CASES = [("rmat_s8_ef4", 8, 4, 0x5EED0008), ("rmat_s10_ef8", 10, 8, 0x5EED000A)]

HAND_EDGES = """# hand-made: self loops, duplicates, reversed duplicates, isolated ids 1,4,6
0 2
2 0
0 2
3 3
2 5
5 7
7 7
0 7
9 5
% comment line
9 2
"""


def rmat_edges(scale, ef, seed):
    lib = ctypes.CDLL(ORACLE)
    n = ef << scale
    src = np.zeros(n, np.uint32)
    dst = np.zeros(n, np.uint32)
    P = ctypes.c_void_p
    lib.ltor_rmat_edges.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_uint64, P, P, ctypes.c_int]
    rc = lib.ltor_rmat_edges(scale, n, 0.57, 0.19, 0.19, seed, src.ctypes.data, dst.ctypes.data, 1)
    assert rc == 0
    return src, dst


def convert(el_path):
    subprocess.run([CONVERT, el_path], check=True, stdout=subprocess.DEVNULL)
    return el_path[:-2] + "bcsr"


def main():
    if not os.path.exists(CONVERT):
        sys.exit("build the reference converter first: make -C oracle ref")
    p = os.path.join(HERE, "hand.el")
    with open(p, "w") as f:
        f.write(HAND_EDGES)
    convert(p)
    for name, scale, ef, seed in RMAT_CASES:
        src, dst = rmat_edges(scale, ef, seed)
        p = os.path.join(HERE, name + ".el")
        with open(p, "w") as f:
            for s, d in zip(src.tolist(), dst.tolist()):
                f.write(f"{s} {d}\n")
        convert(p)
    print("ok")


if __name__ == "__main__":
    main()
