"""Writes the golden fixtures under tests/golden/ from the reference's own
host-side code (Matrix Market reader, Tile_create, blockspmv_cpu), compiled
from the reference sources by `make -C oracle ref` into oracle/_ref/ref_dump.

Runs only in the build container (it needs the reference tree to build the
dumper); the resulting .npz files are data and are committed.

Inputs:
  add20          test/add20.mtx from the reference (the only fixture it ships),
                 truncated to a multiple of 16 rows as the reference driver does.
  lap2d64        2D 5-point Laplacian 64x64 (the CPU-runnable workload).
  lap3d_sym_notrunc  3D 7-point Laplacian 10^3 stored as a *symmetric*
                 lower triangle (exercises mirroring), tiled without
                 truncation (1000 rows: a short last tile-row).
  int_pattern    a small general integer-field file and a pattern-field file
                 with unsorted and duplicate-free entries in file order.
Usage: python tests/golden/gen_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "mille-feuille_amd"))
from millefeuille import gen  # noqa: E402

DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")

INT_ARRAYS = {"csr_rowptr": np.int32, "csr_colidx": np.int32, "csr_val": np.float64,
              "tile_ptr": np.int32, "tile_columnidx": np.int32, "tile_nnz": np.int32,
              "blknnz": np.int32, "blknnznnz": np.uint8, "csr_offset": np.int32,
              "csrptr_offset": np.int32, "Blockcsr_Val": np.float64,
              "Blockcsr_Val_Low": np.float32, "Tile_csr_Col": np.uint8,
              "Blockcsr_Ptr": np.uint8, "csr_compressedIdx": np.uint8,
              "ptroffset1": np.int32, "ptroffset2": np.int32,
              "blkcoostylerowidx": np.uint32, "blkcoostylerowidx_colstart": np.int32,
              "blkcoostylerowidx_colstop": np.int32, "spmv_x": np.float64,
              "spmv_y_nibble": np.float64}


def run_dump(mtx, name, notrunc=False):
    with tempfile.TemporaryDirectory() as d:
        cmd = [DUMP, mtx, d] + (["notrunc"] if notrunc else [])
        subprocess.run(cmd, check=True)
        out = {k: np.fromfile(os.path.join(d, k + ".bin"), dtype=t) for k, t in INT_ARRAYS.items()}
        with open(os.path.join(d, "meta.txt")) as f:
            for line in f:
                k, v = line.split()
                out["meta_" + k] = np.array(int(v))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, {k[5:]: int(v) for k, v in out.items() if k.startswith("meta_")})


def write_sym_lower(path, rowptr, colidx, val):
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    keep = colidx <= rows
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% lower triangle\n")
        f.write(f"{n} {n} {int(keep.sum())}\n")
        for r, c, v in zip(rows[keep], colidx[keep], val[keep]):
            f.write(f"{r + 1} {c + 1} {v:.17g}\n")


def main():
    if not os.path.exists(DUMP):
        sys.exit("build the dumper first: make -C oracle ref")
    run_dump(os.path.join(HERE, "add20.mtx"), "add20")
    with tempfile.TemporaryDirectory() as d:
        rp, ci, va = gen.laplace2d(64)
        p = os.path.join(d, "lap2d64.mtx")
        gen.write_mtx(p, rp, ci, va, comment="2D 5-point Laplacian 64x64, seed 1")
        run_dump(p, "lap2d64")
        rp, ci, va = gen.laplace3d(10)
        p = os.path.join(d, "lap3d10_sym.mtx")
        write_sym_lower(p, rp, ci, va)
        # keep the symmetric file as an input fixture for the reader tests
        with open(p) as src, open(os.path.join(HERE, "lap3d10_sym.mtx"), "w") as dst:
            dst.write(src.read())
        run_dump(p, "lap3d_sym_notrunc", notrunc=True)
    # integer and pattern fields, entries deliberately out of order
    ent = [(3, 1, 5), (1, 1, 2), (2, 3, -4), (1, 2, 7), (3, 3, 1), (2, 2, 9), (1, 3, 3)]
    for field, name in (("integer", "int3x3"), ("pattern", "pat3x3")):
        p = os.path.join(HERE, name + ".mtx")
        with open(p, "w") as f:
            f.write(f"%%MatrixMarket matrix coordinate {field} general\n%\n3 3 {len(ent)}\n")
            for i, j, v in ent:
                f.write(f"{i} {j} {v}\n" if field == "integer" else f"{i} {j}\n")
        run_dump(p, name, notrunc=True)


if __name__ == "__main__":
    main()
