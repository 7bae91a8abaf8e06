#!/usr/bin/env python3
"""Writes the committed fixtures of tests/golden/ (run in a checkout that has
the reference at /root/reference; GPU runs and the tests only read the
outputs).

* sptrans_golden.npz: small seeded CSR matrices (edge cases included) and
  their transposes computed by the reference's own sptrans_scanTrans_specialized
  (sptrans.h:397-545), compiled from the reference source by `make -C oracle ref`
  into oracle/_ref/.  Every matrix is transposed with 1, 2, 3 and 8 OpenMP
  threads and the outputs must agree before they are written.
* tiny_3x4.csr4 + tiny_3x4.json: a csr4 file assembled byte by byte from the
  format (csr4matrix.cpp:69-148: 28-byte header, 60-byte scanner block with
  1 layer, blocksize 1, 1 ring, u64 inclusive row ends, AoS {u32 col, f32 val})
  and the arrays it must decode to.
* sino_5.i32: a raw int32 vector file (vector.hpp:33-53) of 5 known values.
"""
import json
import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def csr_from_lengths(lens, cols, rng, dup=False):
    lens = np.asarray(lens, np.int64)
    rowptr = np.zeros(len(lens) + 1, np.uint64)
    rowptr[1:] = np.cumsum(lens)
    nnz = int(rowptr[-1])
    col = np.empty(nnz, np.uint32)
    for r in range(len(lens)):
        b, e = int(rowptr[r]), int(rowptr[r + 1])
        if e == b:
            continue
        if dup or e - b > cols:
            c = rng.integers(0, cols, e - b)
        else:
            c = rng.choice(cols, e - b, replace=False)
        col[b:e] = np.sort(c)
    val = (rng.random(nnz) * 4.0 - 1.0).astype(np.float32)
    return len(lens), cols, rowptr, col, val


def matrices():
    rng = np.random.default_rng(20260)
    out = []
    out.append(("ragged", csr_from_lengths(rng.integers(0, 30, 300), 200, rng)))
    out.append(("empty_rows", csr_from_lengths(rng.integers(0, 12, 400) * (rng.random(400) < 0.3), 90, rng)))
    out.append(("empty_cols", csr_from_lengths(rng.integers(1, 6, 120), 1000, rng)))
    out.append(("dups", csr_from_lengths(rng.integers(1, 25, 60), 11, rng, dup=True)))
    out.append(("single_row", csr_from_lengths([77], 100, rng)))
    out.append(("single_col", csr_from_lengths(rng.integers(0, 2, 50), 1, rng, dup=True)))
    out.append(("all_empty", csr_from_lengths(np.zeros(13, np.int64), 7, rng)))
    out.append(("tiny", csr_from_lengths([2, 0, 3], 4, rng)))
    out.append(("nnz_prime", csr_from_lengths([7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73],
                                               2500, rng)))
    out.append(("wide_3000cols", csr_from_lengths(rng.integers(0, 40, 200), 3000, rng)))
    return out


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True)
    import oracle as O
    assert O.ref_available(), "oracle/_ref was not built (no reference checkout?)"
    arrays = {}
    names = []
    for name, (rows, cols, rowptr, col, val) in matrices():
        outs = [O.ref_sptrans(rows, cols, rowptr, col, val, threads=t) for t in (1, 2, 3, 8)]
        for o in outs[1:]:
            for a, b in zip(outs[0], o):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), name
        colptr, rowT, valT = outs[0]
        names.append(name)
        arrays[f"{name}/shape"] = np.array([rows, cols], np.uint32)
        arrays[f"{name}/rowptr"] = rowptr
        arrays[f"{name}/col"] = col
        arrays[f"{name}/val"] = val
        arrays[f"{name}/colptrT"] = colptr
        arrays[f"{name}/rowT"] = rowT
        arrays[f"{name}/valT"] = valT
    arrays["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "sptrans_golden.npz"), **arrays)

    # tiny 3 x 4 csr4, 5 entries: row 0 = {(1, 0.5), (3, -2.0)}, row 1 empty,
    # row 2 = {(0, 1.25), (2, 3.0), (3, 0.125)}.
    rows, cols = 3, 4
    entries = [[(1, 0.5), (3, -2.0)], [], [(0, 1.25), (2, 3.0), (3, 0.125)]]
    nnz = sum(len(r) for r in entries)
    scan = b"".join([
        struct.pack("<I", 1),                               # nLayers
        struct.pack("<5f", 2.0, 4.0, 20.0, 40.0, 0.087),    # width, height, depth, radius, mu
        struct.pack("<I", 8), struct.pack("<I", 1),         # nBlocks, blocksize (no gaps)
        struct.pack("<I", 1),                               # nRings (no gaps)
        struct.pack("<3f", 1.0, 1.0, 2.0),                  # voxel size
        struct.pack("<3I", 4, 1, 1),                        # nx, ny, nz
    ])
    assert len(scan) == 60
    header = b"PCSR" + struct.pack("<BBxx", 4, 0) + struct.pack("<IIQI", rows, cols, nnz, len(scan))
    assert len(header) == 28
    ends, run = [], 0
    for r in entries:
        run += len(r)
        ends.append(run)
    body = struct.pack(f"<{rows}Q", *ends) + b"".join(struct.pack("<If", c, v) for r in entries for c, v in r)
    with open(os.path.join(HERE, "tiny_3x4.csr4"), "wb") as f:
        f.write(header + scan + body)
    rowptr = [0] + ends
    expect = {"rows": rows, "cols": cols, "nnz": nnz, "flags": 0, "scan_config_bytes": 60, "row_ends": ends,
              "rowptr": rowptr, "col": [c for r in entries for c, _ in r], "val": [v for r in entries for _, v in r],
              "scanner": {"n_layers": 1, "layer0": [2.0, 4.0, 20.0, 40.0, 0.087], "n_blocks": 8, "blocksize": 1,
                          "n_rings": 1, "voxel_size": [1.0, 1.0, 2.0], "grid": [4, 1, 1]},
              "dense": [[0.0, 0.5, 0.0, -2.0], [0.0, 0.0, 0.0, 0.0], [1.25, 0.0, 3.0, 0.125]]}
    with open(os.path.join(HERE, "tiny_3x4.json"), "w") as f:
        json.dump(expect, f, indent=1)

    with open(os.path.join(HERE, "sino_5.i32"), "wb") as f:
        f.write(struct.pack("<5i", 3, 0, -7, 2147483647, 12))
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
