"""Write the golden fixtures under tests/golden/ (test infrastructure).

Runs oracle/_ref/ref_gen -- a driver over the reference's own generator
headers (include/dr/detail/generate_csr.hpp, include/dr/detail/matrix_io.hpp),
compiled from the reference sources by `make -C oracle ref` -- and packs every
matrix it writes into one .npz (no pickles) per case, plus manifest.json.

Usage:  make -C oracle ref && python tests/golden/make_fixtures.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_GEN = os.path.join(REPO, "oracle", "_ref", "ref_gen")

DT = {"f32": np.float32, "f64": np.float64, "u64": np.uint64,
      "i64": np.int64, "i32": np.int32}


def fnv1a64(b: bytes) -> str:
    h = 0xcbf29ce484222325
    for byte in b:
        h ^= byte
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


def main():
    if not os.path.exists(REF_GEN):
        sys.exit("build oracle/_ref/ref_gen first: make -C oracle ref")
    manifest = {}
    with tempfile.TemporaryDirectory() as td:
        subprocess.check_call([REF_GEN, td])
        for line in open(os.path.join(td, "manifest.txt")):
            name, m, n, nnz, vt, it = line.split()
            m, n, nnz = int(m), int(n), int(nnz)
            rd = lambda suf, t: np.fromfile(os.path.join(td, name + suf),
                                            dtype=DT[t])
            vals, rowptr, colind = rd(".vals", vt), rd(".rowptr", it), rd(".colind", it)
            assert vals.size == nnz and colind.size == nnz and rowptr.size == m + 1
            np.savez(os.path.join(HERE, name + ".npz"), vals=vals,
                     rowptr=rowptr, colind=colind,
                     shape=np.array([m, n], dtype=np.int64))
            manifest[name] = {
                "m": m, "n": n, "nnz": nnz, "vtype": vt, "itype": it,
                "fnv1a64_rowptr": fnv1a64(rowptr.tobytes()),
                "fnv1a64_colind": fnv1a64(colind.tobytes()),
                "fnv1a64_vals": fnv1a64(vals.tobytes()),
            }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote %d fixtures" % len(manifest))


if __name__ == "__main__":
    main()
