"""The FASTA reader restatement (oracle/rnj_oracle.c) against the reference reader's own output.

tests/golden/fasta/*.fa exercise one reader behaviour each (comments, CRLF, tabs ending names,
empty records, data before the first header, bytes >= 128, control bytes, no final newline,
rejections); tests/golden/fasta_golden.npz holds what the reference's dataloaderFasta produced for
them (tests/golden/make_fasta_golden.py). The restatement must give the same names, the same
sequences (compared through the reference's packing) and the same rejections.
"""
import glob
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(GOLDEN, "fasta", "*.fa")))


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "fasta_golden.npz"), allow_pickle=False))


def golden_names(g, case):
    blob = bytes(g[f"{case}_names"])
    out, o = [], 0
    for k in g[f"{case}_name_len"]:
        out.append(blob[o:o + int(k)])
        o += int(k)
    return out


def test_every_case_has_golden(golden):
    assert CASES and all(f"{c}_status" in golden for c in CASES)


@pytest.mark.parametrize("case", CASES)
def test_oracle_reader_matches_reference(golden, case):
    data = open(os.path.join(GOLDEN, "fasta", case + ".fa"), "rb").read()
    status = int(golden[f"{case}_status"])
    if status:
        msg = bytes(golden[f"{case}_message"]).decode(errors="replace")
        with pytest.raises(O.FastaError) as e:
            O.parse_fasta(data)
        m = re.search(r"Sequence (\d+) has length (\d+)", msg)
        if m:
            assert e.value.kind == 1 and (e.value.seq, e.value.length) == (int(m[1]), int(m[2]))
        else:
            assert e.value.kind == 2 and "At least 2 sequences" in msg
        return
    names, rows = O.parse_fasta(data)
    assert names == golden_names(golden, case)
    assert rows.shape == (int(golden[f"{case}_n"]), int(golden[f"{case}_L"]))
    assert O.is_dna(rows[0].tobytes())
    bits, gaps, _ = O.pack(rows)
    np.testing.assert_array_equal(bits, golden[f"{case}_bits"])
    np.testing.assert_array_equal(gaps, golden[f"{case}_gaps"])


def test_oracle_reader_on_generated_alignment():
    """write_fasta output (the repo's own writer) reads back as the alignment it wrote."""
    from rapidnj_amd.workloads import simulate_alignment
    aln = simulate_alignment(30, 257, 0.1, 5)
    lines = []
    for i, row in enumerate(aln):
        lines.append(b">t%d" % i)
        s = row.tobytes()
        lines += [s[k:k + 61] for k in range(0, len(s), 61)]
    names, rows = O.parse_fasta(b"\n".join(lines) + b"\n")
    assert names == [b"t%d" % i for i in range(30)]
    np.testing.assert_array_equal(rows, aln)
