"""The Stockholm reader restatement (oracle/rnj_oracle.c) against the reference reader's own output.

tests/golden/stockholm/*.sth exercise one reader behaviour each (markup and blank lines, tabs and
CRLF, a name alone on its line, an empty name, a leading '/', bytes >= 128, control bytes ending a
name, text after "//", interleaved blocks, rejections); tests/golden/stockholm_golden.npz holds what
the reference's dataloaderStockholm produced for them (tests/golden/make_stockholm_golden.py). The
restatement must give the same names, the same sequences (compared through the reference's
packing) and the same rejections.
"""
import glob
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "stockholm", "*.sth")))


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "stockholm_golden.npz"), allow_pickle=False))


def golden_names(g, case):
    blob = bytes(g[f"{case}_names"])
    out, o = [], 0
    for k in g[f"{case}_name_len"]:
        out.append(blob[o:o + int(k)])
        o += int(k)
    return out


def expected_error(msg: str):
    """(kind, seq, length) of the oracle's StockholmError for the reference's message."""
    m = re.search(r"Sequence (\d+) has length (\d+)", msg)
    if m:
        return 1, int(m[1]), int(m[2])
    if "At least 2 sequences" in msg:
        return 2, None, None
    if "has zero length" in msg:
        return 3, None, None
    assert "Stream ended unexpectedly" in msg, msg
    return 4, None, None


def test_every_case_has_golden(golden):
    assert len(CASES) >= 20 and all(f"{c}_status" in golden for c in CASES)


@pytest.mark.parametrize("case", CASES)
def test_oracle_reader_matches_reference(golden, case):
    data = open(os.path.join(GOLDEN, "stockholm", case + ".sth"), "rb").read()
    status = int(golden[f"{case}_status"])
    if status:
        kind, seq, length = expected_error(bytes(golden[f"{case}_message"]).decode(errors="replace"))
        with pytest.raises(O.StockholmError) as e:
            O.parse_stockholm(data)
        assert e.value.kind == kind
        if kind == 1:
            assert (e.value.seq, e.value.length) == (seq, length)
        return
    names, rows = O.parse_stockholm(data)
    assert names == golden_names(golden, case)
    assert rows.shape == (int(golden[f"{case}_n"]), int(golden[f"{case}_L"]))
    assert O.is_dna(rows[0].tobytes())
    bits, gaps, _ = O.pack(rows)
    np.testing.assert_array_equal(bits, golden[f"{case}_bits"])
    np.testing.assert_array_equal(gaps, golden[f"{case}_gaps"])


def test_zero_length_reports_the_sequence():
    """The reference names the sequence whose line is empty; the oracle reports its index."""
    with pytest.raises(O.StockholmError) as e:
        O.parse_stockholm(b"a ACGT\nb ACGT\nname\n\n//\n")
    assert (e.value.kind, e.value.seq) == (3, 2)


def test_oracle_reader_on_generated_alignment():
    """A generated alignment written as Stockholm reads back as the alignment it wrote."""
    from rapidnj_amd.workloads import simulate_alignment
    aln = simulate_alignment(30, 257, 0.1, 5)
    text = b"# STOCKHOLM 1.0\n" + b"".join(b"t%d  " % i + row.tobytes() + b"\n"
                                         for i, row in enumerate(aln)) + b"//\n"
    names, rows = O.parse_stockholm(text)
    assert names == [b"t%d" % i for i in range(30)]
    np.testing.assert_array_equal(rows, aln)
