"""Device Stockholm reader (rnj_alignment_from_stockholm / _file) against the reference reader.

Fixtures: tests/golden/stockholm/*.sth with the reference dataloaderStockholm's names, packed
sequences and rejections (tests/golden/stockholm_golden.npz); larger and randomised inputs are
checked against the oracle's restatement of the same reader (oracle/rnj_oracle.c, itself pinned
to the fixtures by tests/test_stockholm_oracle.py).
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
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "stockholm", "*.sth")))


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
    return dict(np.load(os.path.join(GOLDEN, "stockholm_golden.npz"), allow_pickle=False))


def _golden_names(g, case):
    blob = bytes(g[f"{case}_names"])
    out, o = [], 0
    for k in g[f"{case}_name_len"]:
        out.append(blob[o:o + int(k)])
        o += int(k)
    return out


def _read(case):
    with open(os.path.join(GOLDEN, "stockholm", case + ".sth"), "rb") as fh:
        return fh.read()


def _assert_same_error(err, kind, seq=None, length=None):
    """The device's message for the oracle's / reference's error kind."""
    text = str(err)
    if kind == 1:
        assert f"sequence {seq} has length {length}" in text, text
    elif kind == 2:
        assert "at least 2 are needed" in text, text
    elif kind == 3:
        assert "has zero length" in text and (seq is None or f"sequence {seq} " in text), text
    else:
        assert "ends before the \"//\"" in text, text


def _check_case(ctx, N, golden, case, make):
    status = int(golden[f"{case}_status"])
    if status:
        msg = bytes(golden[f"{case}_message"]).decode(errors="replace")
        with pytest.raises(N.RnjError) as e:
            make()
        assert e.value.status == N.RNJ_E_INVALID
        m = re.search(r"Sequence (\d+) has length (\d+)", msg)
        if m:
            _assert_same_error(e.value, 1, m[1], m[2])
        elif "At least 2 sequences" in msg:
            _assert_same_error(e.value, 2)
        elif "has zero length" in msg:
            name = re.search(r'name "(.*)" has', msg)[1]
            assert f'("{name}") has zero length' in str(e.value)
        else:
            _assert_same_error(e.value, 4)
        return
    A = make()
    assert A.names == _golden_names(golden, case)
    assert (A.n, A.L) == (int(golden[f"{case}_n"]), int(golden[f"{case}_L"]))
    ref = ctx.alignment_from_bitstrings(golden[f"{case}_bits"], golden[f"{case}_gaps"])
    np.testing.assert_array_equal(A.planes(), ref.planes()[:, : A.planes().shape[1], :])


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_reference_from_host_bytes(ctx, N, golden, case):
    _check_case(ctx, N, golden, case, lambda: ctx.alignment_from_stockholm(_read(case)))


@pytest.mark.parametrize("case", CASES)
def test_reader_matches_reference_from_file(ctx, N, golden, case):
    path = os.path.join(GOLDEN, "stockholm", case + ".sth")
    _check_case(ctx, N, golden, case, lambda: ctx.alignment_from_stockholm_file(path))


@pytest.mark.parametrize("case", ["basic", "name_own_line", "high_bit", "zero_length"])
def test_reader_from_device_memory_aligned_and_not(ctx, N, golden, case):
    import torch
    raw = np.frombuffer(_read(case), dtype=np.uint8)
    dev = torch.from_numpy(np.concatenate([np.zeros(1, np.uint8), raw])).to("cuda:0")
    for t in (dev[1:].clone(), dev[1:]):              # 16-byte aligned copy, then an offset view
        _check_case(ctx, N, golden, case, lambda: ctx.alignment_from_stockholm(t))


def _stockholm(aln, eol=b"\n", markup_every=0, own_line_every=0):
    out = [b"# STOCKHOLM 1.0", b"#=GF ID large"]
    for i, row in enumerate(aln):
        if markup_every and i % markup_every == 0:
            out += [b"#=GS taxon_%d DE block %d" % (i, i), b""]
        if own_line_every and i % own_line_every == 1:
            out += [b"taxon_%d" % i, b"\t" + row.tobytes()]
        else:
            out.append(b"taxon_%d\t " % i + row.tobytes())
    return eol.join(out + [b"//"]) + eol


# (a name alone on its line needs a bare '\n': after "name\r" the reference reads the '\n' while
# skipping control bytes, which is its "zero length" error)
@pytest.mark.parametrize("eol,markup,own", [(b"\n", 97, 0), (b"\r\n", 0, 0), (b"\n", 13, 3)])
def test_large_input_against_oracle_and_char_path(ctx, eol, markup, own):
    aln = simulate_alignment(1500, 3001, 0.05, 43)
    data = _stockholm(aln, eol, markup, own)
    names, rows = O.parse_stockholm(data)
    A = ctx.alignment_from_stockholm(data)
    assert A.names == names == [b"taxon_%d" % i for i in range(1500)]
    np.testing.assert_array_equal(rows, aln)
    B = ctx.alignment(aln)
    np.testing.assert_array_equal(A.planes(), B.planes())
    da = ctx.distances(A, "k2p").cpu().numpy()
    db = ctx.distances(B, "k2p").cpu().numpy()
    np.testing.assert_array_equal(da.view(np.uint32), db.view(np.uint32))


def _random_stockholm(rng, n, L):
    """Lines of every kind the reader distinguishes, in random order: markup, blank and
    control-byte lines, sequences with a space, tab or control byte after the name, names alone on
    their line (data on the next), '/'-prefixed and empty names, bytes >= 128; sometimes a
    sequence of another length, an empty data line or no "//"."""
    alpha = np.frombuffer(b"ACGTacgtN-.", dtype=np.uint8)
    lines = []
    for i in range(n):
        if rng.random() < 0.3:
            lines.append(rng.choice([b"#=GC SS_cons ...", b"", b"\t\r", b"\x01\x02", b"\xff",
                                     b"# comment / // x"]))
        data = bytes(alpha[rng.integers(0, len(alpha), L)])
        if rng.random() < 0.004:
            data = data[:-1]                                   # a length mismatch
        pre = rng.choice([b"", b"\t", b"\r", b"\x80"]) if rng.random() < 0.2 else b""
        name = rng.choice([b"s%d" % i, b"/s%d" % i, b"s%d\xc3\xa9" % i])
        k = rng.integers(0, 6)
        if k == 0:
            lines.append(pre + name + b" " + data)
        elif k == 1:
            lines.append(pre + name + b"\t\t" + data)
        elif k == 2:
            lines.append(pre + name + b"\x03" + data)
        elif k == 3:
            lines += [pre + name, rng.choice([b"", b"\t"]) + data]    # data on the next line
        elif k == 4:
            lines.append(b" " + data)                                 # an empty name
        else:
            lines.append(pre + name + b"   " + data[:L // 2] + b" \t " + data[L // 2:])
        if rng.random() < 0.003:
            lines += [name, b""]                                       # zero length
    if rng.random() < 0.85:
        lines.append(b"//")
        if rng.random() < 0.5:
            lines.append(b"junk after the end\n\nx")
    text = b"\n".join(lines)
    return text + (b"\n" if rng.random() < 0.8 else b"")


def test_randomised_inputs_against_oracle(ctx, N):
    """200 random inputs: the device gives the oracle's names and sequences, or its error."""
    rng = np.random.default_rng(77)
    outcomes = set()
    for case in range(200):
        data = _random_stockholm(rng, int(rng.integers(2, 40)), int(rng.integers(1, 90)))
        try:
            names, rows = O.parse_stockholm(data)
        except O.StockholmError as err:
            outcomes.add(err.kind)
            with pytest.raises(N.RnjError) as e:
                ctx.alignment_from_stockholm(data)
            _assert_same_error(e.value, err.kind, err.seq, err.length)
            continue
        outcomes.add(0)
        A = ctx.alignment_from_stockholm(data)
        assert A.names == names, case
        if A.is_protein:
            continue
        np.testing.assert_array_equal(A.planes(), ctx.alignment(rows, skip_type_check=True).planes(),
                                      err_msg=f"case {case}")
    assert {0, 1, 3, 4} <= outcomes  # parsed, a length mismatch, a zero length, no "//"


def test_protein_and_naive_type_discovery(ctx, N):
    """Protein input gives a protein alignment; under the naive rules the type is discovered after
    resolveChar, so an X in the first sequence no longer makes it protein (the -f path)."""
    prot = b"p1 MKVLAAGGWWKL\np2 MKVLAQGGWWKL\n//\n"
    A = ctx.alignment_from_stockholm(prot)
    assert A.is_protein and A.names == [b"p1", b"p2"]
    d = ctx.distances(A, "k2p").cpu().numpy()
    rows = np.frombuffer(b"MKVLAAGGWWKLMKVLAQGGWWKL", dtype=np.uint8).reshape(2, 12)
    np.testing.assert_array_equal(d.view(np.uint32),
                                  ctx.distances(ctx.alignment(rows), "k2p").cpu().numpy().view(np.uint32))
    dna_x = b"a ACGTXXACGT\nb ACGTACACGT\n//\n"
    assert ctx.alignment_from_stockholm(dna_x).is_protein
    ctx.set_char_rules(True)
    try:
        B = ctx.alignment_from_stockholm(dna_x)
        assert not B.is_protein
    finally:
        ctx.set_char_rules(False)


def test_crlf_name_alone_on_its_line_is_zero_length(ctx, N):
    data = b"a ACGT\r\nb\r\nACGT\r\n//\r\n"
    with pytest.raises(O.StockholmError) as err:
        O.parse_stockholm(data)
    assert (err.value.kind, err.value.seq) == (3, 1)
    with pytest.raises(N.RnjError) as e:
        ctx.alignment_from_stockholm(data)
    _assert_same_error(e.value, 3, 1)


def test_rejections(ctx, N):
    for data in (b"", b"# STOCKHOLM 1.0\n", b"a ACGT\nb ACGT\n"):
        with pytest.raises(N.RnjError) as e:
            ctx.alignment_from_stockholm(data)
        assert e.value.status == N.RNJ_E_INVALID
    with pytest.raises(N.RnjError) as e:
        ctx.alignment_from_stockholm(b"a \nb \n//\n")                # empty sequences
    assert "empty" in str(e.value)
    with pytest.raises(N.RnjError) as e:
        ctx.alignment_from_stockholm_file("/nonexistent/input.sth")
    assert e.value.status == N.RNJ_E_INVALID
