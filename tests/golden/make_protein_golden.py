"""Write tests/golden/protein.npz by running the reference's protein distance engine.

Run in the build container only (needs /root/reference and ``make -C oracle ref``):

    python tests/golden/make_protein_golden.py

For each case the alignment is written as FASTA and read by the reference's own reader and
estimator through oracle/_ref/ref_harness: `pcounts` (bitDistanceProtein::computeDistance,
bitDistanceProtein.cpp:31-61: data[0], data[1], data[2] per pair) and `pmatrix` (KimuraDistance /
JCdistance on PROTEIN input, KimuraDistance.cpp:322-338, JCdistance.cpp:155-183, post-processed).
Cases: the reference's own protein test file test_data/PF15224_full.fa (stored as its parsed
character matrix), and seeded alignments from a random root mutated at rates from 0 to 0.97,
over amino acids, lowercase letters and every character the reader resolves to a gap
(- . X x z Z b B J j ?), at lengths 1, 128 and 256 (where the estimator's valid count comes out
128 short, wrapping modulo 2^64 for pairs with few common residues), 300 and 640, and one at
length 200 with a row of gaps only (valid count 0) and a row with two residues.

The same cases are also read with fastdist off, the reference's `-f` path: `pcounts ... naive`
and `pmatrix ... naive` run simpleDistanceCalculator::calculateHammingDistance
(simpleDistanceCalculator.cpp:40-55) on the loader's resolved characters, whose valid count is
the exact number of residue pairs (no 128-position shortfall).

npz keys per case <c>: <c>_aln (n x L uint8), <c>_d0, <c>_d1, <c>_d2 (upper triangle, row-major,
uint64), <c>_kim and <c>_jc (n x n float32); the -f path's <c>_nd0, <c>_nd1, <c>_nd2, <c>_nkim and
<c>_njc.
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

from oracle import oracle as O  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_PF = "/root/reference/test_data/PF15224_full.fa"
AMINO = b"ACDEFGHIKLMNPQRSTVWY"
NOISE = b"acdefghiklmnpqrstvwy-.XxzZbBJj?*"


def run(*args: str) -> None:
    subprocess.run([HARNESS, *args], check=True, capture_output=True)


def write_fasta(path: str, aln: np.ndarray) -> None:
    with open(path, "wb") as fh:
        for i, row in enumerate(aln):
            fh.write(b">p%d\n" % i + row.tobytes() + b"\n")


def simulate(n: int, L: int, seed: int, gap: float) -> np.ndarray:
    rng = np.random.default_rng(seed)
    amino = np.frombuffer(AMINO, dtype=np.uint8)
    noise = np.frombuffer(NOISE, dtype=np.uint8)
    root = amino[rng.integers(0, len(amino), L)]
    rates = np.linspace(0.0, 0.97, n)
    rng.shuffle(rates)
    out = np.empty((n, L), dtype=np.uint8)
    for i in range(n):
        row = root.copy()
        mut = rng.random(L) < rates[i]
        row[mut] = amino[rng.integers(0, len(amino), int(mut.sum()))]
        odd = rng.random(L) < gap
        row[odd] = noise[rng.integers(0, len(noise), int(odd.sum()))]
        out[i] = row
    out[0, : min(L, 4)] = np.frombuffer(b"EFLW"[: min(L, 4)], dtype=np.uint8)  # protein type
    return out


def with_gap_rows(aln: np.ndarray) -> np.ndarray:
    """Row 3 all gap characters (no common residue with any row: valid count 0), row 5 all gaps
    but its last two sites."""
    gaps = np.frombuffer(b"-.XxzZbBJj?", dtype=np.uint8)
    aln = aln.copy()
    aln[3] = gaps[np.arange(aln.shape[1]) % len(gaps)]
    aln[5, :-2] = gaps[np.arange(aln.shape[1] - 2) % len(gaps)]
    return aln


def case(aln: np.ndarray, tmp: str) -> dict:
    fa = os.path.join(tmp, "p.fa")
    write_fasta(fa, aln)
    names, rows = O.parse_fasta(open(fa, "rb").read())
    assert np.array_equal(rows, aln)
    out = {"aln": aln}
    for pre, extra in (("", []), ("n", ["naive"])):
        cnt = os.path.join(tmp, "c.bin")
        run("pcounts", fa, cnt, *extra)
        raw = np.fromfile(cnt, dtype=np.uint8)
        d = np.frombuffer(raw[8:].tobytes(), dtype=np.uint64).reshape(-1, 3)
        for k in range(3):
            out[f"{pre}d{k}"] = d[:, k].copy()
        for model in ("kim", "jc"):
            mat = os.path.join(tmp, "m.bin")
            run("pmatrix", fa, model, mat, *extra)
            raw = np.fromfile(mat, dtype=np.uint8)
            n = int(np.frombuffer(raw[:4].tobytes(), dtype=np.uint32)[0])
            out[pre + model] = np.frombuffer(raw[4:].tobytes(), dtype=np.float32).reshape(n, n).copy()
    return out


def main() -> None:
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        _, pf = O.parse_fasta(open(REF_PF, "rb").read())
        cases = {"pf15224": pf,
                 "p12x1": simulate(12, 1, 1, 0.2),
                 "p20x128": simulate(20, 128, 2, 0.5),
                 "p30x256": simulate(30, 256, 3, 0.1),
                 "p25x300": simulate(25, 300, 4, 0.05),
                 "p16x640": simulate(16, 640, 5, 0.02),
                 "p10x200gap": with_gap_rows(simulate(10, 200, 6, 0.1))}
        for name, aln in cases.items():
            for k, v in case(aln, tmp).items():
                res[f"{name}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "protein.npz"), **res)
    print("protein golden vectors written:", sorted({k.split("_")[0] for k in res}))


if __name__ == "__main__":
    main()
