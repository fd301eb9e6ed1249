"""Write tests/golden/nj.npz: the reference's own neighbour-joining trees.

Run in the build container only (needs /root/reference and ``make -C oracle ref``):

    python tests/golden/make_nj_golden.py

Every case is a distance matrix the reference computed itself (``ref_harness matrix``, its
KimuraDistance / JCdistance on the FASTA written here) and the Newick text its rapidNJ builds on
that matrix (``ref_harness nj``, rapidNJ.cpp + polytree.cpp, leaf names t0..t{n-1}), with
negative_branches off and on. The cases:

* ``dup``: 60 taxa with copies (pairs and a group of four identical sequences), so rapidNJ's
  redundancy handling (rapidNJ.cpp:118-129,155-164,251-272,418-430) joins copies;
* ``short``: 150 taxa x 40 sites, so that many distances are equal;
* ``sat``: 50 taxa x 150 sites evolved far enough that pairs saturate (2 * max entries);
* ``sim300``: 300 taxa x 800 sites, K2P;
* ``sim300jc``: the same alignment, JC;
* ``sim800``: 800 taxa x 400 sites, K2P; stored as the alignment only (the matrix is the
  reference's K2P matrix of it, which the oracle reproduces bit for bit).

npz keys per case <c>: <c>_matrix (n x n float32; absent for sim800), <c>_aln (uint8, sim800
only), <c>_newick and <c>_newick_neg (uint8 text, trailing newline kept).
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


def run(*args: str) -> None:
    subprocess.run([HARNESS, *args], check=True, capture_output=True)


def ref_matrix(aln: np.ndarray, model: str, tmp: str) -> np.ndarray:
    fa, out = os.path.join(tmp, "a.fa"), os.path.join(tmp, "m.bin")
    write_fasta(fa, aln)
    run("matrix", fa, model, "4", out)
    raw = np.fromfile(out, dtype=np.uint8)
    n = int(raw[:4].view(np.uint32)[0])
    return raw[4:].view(np.float32).reshape(n, n).copy()


def ref_newick(matrix: np.ndarray, tmp: str, neg: bool) -> bytes:
    mf, out = os.path.join(tmp, "m.bin"), os.path.join(tmp, "t.nwk")
    with open(mf, "wb") as fh:
        np.array([matrix.shape[0]], dtype=np.uint32).tofile(fh)
        np.ascontiguousarray(matrix, dtype=np.float32).tofile(fh)
    run("nj", mf, out, *(["neg"] if neg else []))
    with open(out, "rb") as fh:
        return fh.read()


def dup_alignment() -> np.ndarray:
    base = simulate_alignment(52, 300, 0.02, 41)
    rows = list(base)
    for src in (3, 3, 3, 10, 20, 31):   # taxon 3 four times in all, three more pairs
        rows.insert(int(np.random.default_rng(src + len(rows)).integers(0, len(rows))), base[src])
    rows.append(base[45])
    rows.append(base[0])
    return np.stack(rows)


def main() -> None:
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        sim300 = simulate_alignment(300, 800, 0.05, 43)
        cases = {
            "dup": ref_matrix(dup_alignment(), "kim", tmp),
            "short": ref_matrix(simulate_alignment(150, 40, 0.0, 42), "kim", tmp),
            "sat": ref_matrix(simulate_alignment(50, 150, 0.05, 44, mean_branch=0.6), "kim", tmp),
            "sim300": ref_matrix(sim300, "kim", tmp),
            "sim300jc": ref_matrix(sim300, "jc", tmp),
        }
        for name, m in cases.items():
            res[f"{name}_matrix"] = m
            for neg in (False, True):
                key = f"{name}_newick_neg" if neg else f"{name}_newick"
                res[key] = np.frombuffer(ref_newick(m, tmp, neg), dtype=np.uint8)
        sim800 = simulate_alignment(800, 400, 0.05, 45)
        m = ref_matrix(sim800, "kim", tmp)
        res["sim800_aln"] = sim800
        for neg in (False, True):
            key = "sim800_newick_neg" if neg else "sim800_newick"
            res[key] = np.frombuffer(ref_newick(m, tmp, neg), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "nj.npz"), **res)
    print("NJ fixtures written:", sorted({k.split("_")[0] for k in res}))


if __name__ == "__main__":
    main()
