"""Device FASTA reader (rnj_alignment_from_fasta / _file) against the reference reader.

Fixtures: tests/golden/fasta/*.fa with the reference dataloaderFasta's names, packed sequences
and rejections (tests/golden/fasta_golden.npz); larger inputs are checked against the oracle's
restatement of the same reader (oracle/rnj_oracle.c, itself pinned to the fixtures by
tests/test_fasta_oracle.py).
"""
import glob
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from rapidnj_amd.workloads import simulate_alignment

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(GOLDEN, "fasta", "*.fa")))


@pytest.fixture(scope="module")
def N():
    from rapidnj_amd import native
    return native


@pytest.fixture(scope="module")
def ctx(N):
    c = N.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def golden():
    return dict(np.load(os.path.join(GOLDEN, "fasta_golden.npz"), allow_pickle=False))


def _golden_names(g, case):
    blob = bytes(g[f"{case}_names"])
    out, o = [], 0
    for k in g[f"{case}_name_len"]:
        out.append(blob[o:o + int(k)])
        o += int(k)
    return out


def _read(case):
    with open(os.path.join(GOLDEN, "fasta", case + ".fa"), "rb") as fh:
        return fh.read()


def _check_case(ctx, N, golden, case, make):
    status = int(golden[f"{case}_status"])
    if status:
        msg = bytes(golden[f"{case}_message"]).decode(errors="replace")
        with pytest.raises(N.RnjError) as e:
            make()
        assert e.value.status == N.RNJ_E_INVALID
        m = re.search(r"Sequence (\d+) has length (\d+)", msg)
        if m:
            assert f"sequence {m[1]} has length {m[2]}" in str(e.value)
        else:
            assert "at least 2 are needed" in str(e.value)
        return
    A = make()
    assert A.names == _golden_names(golden, case)
    assert (A.n, A.L) == (int(golden[f"{case}_n"]), int(golden[f"{case}_L"]))
    ref = ctx.alignment_from_bitstrings(golden[f"{case}_bits"], golden[f"{case}_gaps"])
    np.testing.assert_array_equal(A.planes()[:, :, :], ref.planes()[:, : A.planes().shape[1], :])


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_reference_from_host_bytes(ctx, N, golden, case):
    _check_case(ctx, N, golden, case, lambda: ctx.alignment_from_fasta(_read(case)))


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_reference_from_file(ctx, N, golden, case):
    path = os.path.join(GOLDEN, "fasta", case + ".fa")
    _check_case(ctx, N, golden, case, lambda: ctx.alignment_from_fasta_file(path))


@pytest.mark.parametrize("case", ["wrapped_lf", "header_tab", "empty_records", "high_bit"])
def test_reader_from_device_memory_aligned_and_not(ctx, N, golden, case):
    import torch
    raw = np.frombuffer(_read(case), dtype=np.uint8)
    dev = torch.from_numpy(np.concatenate([np.zeros(1, np.uint8), raw])).to("cuda:0")
    for t in (dev[1:].clone(), dev[1:]):              # 16-byte aligned copy, then an offset view
        _check_case(ctx, N, golden, case, lambda: ctx.alignment_from_fasta(t))


def _fasta(aln, width, eol=b"\n", comment_every=0):
    out = []
    for i, row in enumerate(aln):
        if comment_every and i % comment_every == 0:
            out.append(b";block %d" % i)
        out.append(b">taxon_%d  description %d" % (i, i * 7))
        s = row.tobytes()
        out += [s[k:k + width] for k in range(0, len(s), width)]
    return eol.join(out) + eol


@pytest.mark.parametrize("width,eol", [(60, b"\n"), (71, b"\r\n"), (10 ** 6, b"\n")])
def test_large_input_against_oracle_and_char_path(ctx, width, eol):
    aln = simulate_alignment(1500, 3001, 0.05, 41)
    data = _fasta(aln, width, eol, comment_every=97)
    names, rows = O.parse_fasta(data)
    A = ctx.alignment_from_fasta(data)
    assert A.names == names
    np.testing.assert_array_equal(rows, aln)
    B = ctx.alignment(aln)
    np.testing.assert_array_equal(A.planes(), B.planes())
    da = ctx.distances(A, "k2p").cpu().numpy()
    db = ctx.distances(B, "k2p").cpu().numpy()
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))


def test_rejections(ctx, N):
    with pytest.raises(N.RnjError) as e:
        ctx.alignment_from_fasta(b"")
    assert e.value.status == N.RNJ_E_INVALID
    # a first sequence the reference classifies as protein gives a protein alignment
    assert ctx.alignment_from_fasta(b">p1\nMKVLAAGGWW\n>p2\nMKVLAAGGWW\n").is_protein
    with pytest.raises(N.RnjError) as e:
        ctx.alignment_from_fasta_file("/nonexistent/input.fa")
    assert e.value.status == N.RNJ_E_INVALID
