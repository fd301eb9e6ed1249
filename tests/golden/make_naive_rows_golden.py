"""Write tests/golden/naive_rows.npz: the reference's `-f` (naive counter) matrices on mixed-case
alignments, from both of its matrix paths.

Run in the build container only (needs /root/reference and ``make -C oracle ref``):

    python tests/golden/make_naive_rows_golden.py

With fastdist off the reference counts with simpleDistanceCalculator::calculateTsTv
(simpleDistanceCalculator.cpp:58-98), which classifies a mismatch only when the row taxon's
character is uppercase A, C, G or T, so its counts are not symmetric in the two taxa. The
in-memory path computes j > i only and mirrors (KimuraDistance.cpp:308-320 with offset i + 1,
postProcessDistanceMatrix :366-395); the disk-matrix path computes every (i, j) with i the row
(offset 0) and takes the maximum over the whole square. oracle/_ref/ref_harness runs both:
`matrix ... naive` and `dmatrix ... naive` (a diskMatrix in a temporary directory, rows read back
with readArray).

Cases: "mixed" (40 x 300: uppercase and lowercase ACGT, U, u, N, '-', with related rows, one row
of gaps only and two unrelated rows) and "sim" (24 x 1000 from the seeded generator, a third of
its cells lowercased and some T turned into U).
npz keys per case <c>: <c>_aln (n x L uint8), <c>_mem_jc, <c>_mem_k2p, <c>_disk_jc, <c>_disk_k2p
(n x n float32).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from rapidnj_amd.workloads import simulate_alignment, write_fasta  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def read_matrix(path: str) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.uint32)
    n = int(raw[0])
    return raw[1:1 + n * n].view(np.float32).reshape(n, n).copy()


def mixed(n: int, L: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    base = np.frombuffer(b"ACGT", dtype=np.uint8)
    root = base[rng.integers(0, 4, L)]
    rows = []
    for i in range(n):
        row = root.copy()
        m = rng.random(L) < 0.02 * i
        row[m] = base[rng.integers(0, 4, int(m.sum()))]
        low = rng.random(L) < 0.3
        row[low] += 32                                   # lowercase
        t = (row == ord("T")) & (rng.random(L) < 0.3)
        row[t] = ord("U")
        row[(row == ord("t")) & (rng.random(L) < 0.3)] = ord("u")
        row[rng.random(L) < 0.03] = ord("N")
        row[rng.random(L) < 0.05] = ord("-")
        rows.append(row)
    out = np.stack(rows)
    out[0] = np.where(np.isin(out[0], np.frombuffer(b"ACGTUNacgtun-", dtype=np.uint8)), out[0],
                      ord("A"))
    out[7] = ord("-")                                    # no valid site with anyone
    out[11] = base[rng.integers(0, 4, L)]                # unrelated rows: saturated pairs
    out[12] = base[rng.integers(0, 4, L)] + 32
    return np.ascontiguousarray(out)


def sim_mixed(n: int, L: int, seed: int) -> np.ndarray:
    aln = simulate_alignment(n, L, 0.04, seed).copy()
    rng = np.random.default_rng(seed + 1)
    letters = aln != ord("-")
    aln[letters & (rng.random(aln.shape) < 0.33)] += 32
    aln[(aln == ord("T")) & (rng.random(aln.shape) < 0.2)] = ord("U")
    return np.ascontiguousarray(aln)


def main() -> None:
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, aln in (("mixed", mixed(40, 300, 5)), ("sim", sim_mixed(24, 1000, 9))):
            fa = os.path.join(tmp, "in.fa")
            write_fasta(fa, aln)
            res[f"{name}_aln"] = aln
            for model, key in (("jc", "jc"), ("kim", "k2p")):
                out = os.path.join(tmp, "m.bin")
                subprocess.run([HARNESS, "matrix", fa, model, "4", out, "naive"], check=True,
                               capture_output=True)
                res[f"{name}_mem_{key}"] = read_matrix(out)
                cache = os.path.join(tmp, "cache")
                os.makedirs(cache, exist_ok=True)
                subprocess.run([HARNESS, "dmatrix", fa, model, "4", cache, out, "naive"],
                               check=True, capture_output=True)
                res[f"{name}_disk_{key}"] = read_matrix(out)
    np.savez_compressed(os.path.join(HERE, "naive_rows.npz"), **res)
    print("naive row vectors written:", sorted(res))


if __name__ == "__main__":
    main()
