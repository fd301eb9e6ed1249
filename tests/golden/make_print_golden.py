"""Write tests/golden/matrix_text.npz: the distance matrices the reference's own program prints.

Run in the build container only (needs /root/reference and ``make -C oracle ref``, which also
links the reference's command-line program from its sources into oracle/_ref/rapidnj):

    python tests/golden/make_print_golden.py

For each case the alignment is written as FASTA (names t0, t1, ..) and run through
`rapidnj FILE -i fa -o m -x OUT -a kim|jc -c 4`, the reference's `-o m` output
(printDistanceMatrix, src/main.cpp:212-221, on the in-memory matrix of
computeDistanceMatrix, :237-286). Cases: the small JC configuration (100 x 1000, 2 % gaps,
seed 1) in both models, the saturated fixture (sentinel entries 2 * max), LGT_70 from the
reference's test data and the protein PF15224 from its test data (Kimura protein).
npz keys: <case>_<model>_text (uint8 bytes of the printed text), <case>_aln (n x L uint8).
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

from rapidnj_amd.workloads import read_fasta, simulate_alignment, write_fasta  # noqa: E402

CLI = os.path.join(ROOT, "oracle", "_ref", "rapidnj")


def printed(aln: np.ndarray, model: str, tmp: str) -> bytes:
    fa = os.path.join(tmp, "in.fa")
    write_fasta(fa, aln)
    out = os.path.join(tmp, "m.txt")
    subprocess.run([CLI, fa, "-i", "fa", "-o", "m", "-x", out, "-a", model, "-c", "4"],
                   check=True, capture_output=True)
    return open(out, "rb").read()


def main() -> None:
    if not os.path.exists(CLI):
        raise SystemExit("build the reference programs first: make -C oracle ref")
    sat = dict(np.load(os.path.join(HERE, "saturated.npz"), allow_pickle=False))["aln"]
    _, lgt = read_fasta("/root/reference/test_data/LGT_70.fa")
    prot = dict(np.load(os.path.join(HERE, "protein.npz"), allow_pickle=False))["pf15224_aln"]
    cases = {"c1": (simulate_alignment(100, 1000, 0.02, 1), ("jc", "kim")),
             "saturated": (sat, ("kim",)),
             "lgt70": (lgt, ("kim",)),
             "pf15224": (prot, ("kim",))}
    res = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (aln, models) in cases.items():
            res[f"{name}_aln"] = np.ascontiguousarray(aln, dtype=np.uint8)
            for model in models:
                res[f"{name}_{model}_text"] = np.frombuffer(printed(aln, model, tmp), np.uint8)
    np.savez_compressed(os.path.join(HERE, "matrix_text.npz"), **res)
    print("printed matrices written:", sorted(k for k in res if k.endswith("_text")))


if __name__ == "__main__":
    main()
