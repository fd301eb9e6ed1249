"""Write the FASTA reader fixtures: tests/golden/fasta/<case>.fa and tests/golden/fasta_golden.npz.

Run in the build container only (needs /root/reference and ``make -C oracle ref``):

    python tests/golden/make_fasta_golden.py

Each input file is generated here (seeded) to exercise one behaviour of the reference reader
dataloaderFasta::parseData / storeSequence (dataloaderFasta.hpp:78-133, 201-276); the expected
outputs come from oracle/_ref/ref_harness `fasta` mode, i.e. the reference's own reader compiled
from its sources: names of every '>' record, the sequence count and length, and the reference's
packed bit strings / gap filters. Cases the reference rejects record its exit status and
message instead.

npz keys per case <c>: <c>_status (0 = read, else the reference's exit status), and for status
0: <c>_n, <c>_L, <c>_bits, <c>_gaps (n x 4*bsc uint32), <c>_names (uint8 blob of the names back
to back), <c>_name_len (uint32 per name); otherwise <c>_message (uint8, the reference's stderr).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT_DIR = os.path.join(HERE, "fasta")

RNG = np.random.default_rng(8)
ALPHA = b"ACGTACGTACGTACGTacgtNnUu-.?RYKMSWBDHV*"
DNA_FIRST = b"ACGTACGTACGTacgtNnUu-.?*"  # the first sequence decides the type: bytes < 'A' or ACGTUN


def seqs(n: int, L: int, alphabet: bytes = ALPHA) -> list[bytes]:
    """n sequences over `alphabet`; the first one over DNA_FIRST (within `alphabet` if it is
    narrower), so that the reference detects DNA."""
    first = bytes(sorted(set(alphabet) & set(DNA_FIRST))) or alphabet
    out = []
    for k in range(n):
        a = np.frombuffer(first if k == 0 else alphabet, dtype=np.uint8)
        out.append(bytes(a[RNG.integers(0, len(a), L)]))
    return out


def wrap(s: bytes, w: int, eol: bytes = b"\n") -> bytes:
    return eol.join(s[i:i + w] for i in range(0, len(s), w)) + eol


def fasta(names, ss, w: int = 60, eol: bytes = b"\n") -> bytes:
    return b"".join(b">" + nm + eol + wrap(s, w, eol) for nm, s in zip(names, ss))


def cases() -> dict[str, bytes]:
    c = {}
    s7 = seqs(7, 150)
    names7 = [b"taxon_%d" % i for i in range(7)]
    c["wrapped_lf"] = fasta(names7, s7)
    c["crlf"] = fasta(names7, s7, eol=b"\r\n")
    c["unwrapped"] = fasta([b"u%d some description" % i for i in range(5)], seqs(5, 3000), w=10 ** 6)
    s6 = seqs(6, 90)
    body = b";comment before the first record\n"
    for i, s in enumerate(s6):
        body += b">r%d\n" % i + b"\n" + s[:30] + b" \t " + s[30:60] + b"\n\n;between lines\n" + s[60:] + b"\n"
        body += b"; comment after record %d\n" % i
    c["comments_blank_spaces"] = body
    s4 = seqs(4, 64)
    # names ended by a tab: the rest of the header line is parsed again from the main loop, so
    # "\tAC" adds AC to the record, "\t>x" opens another record and "\t;..." is a comment
    c["header_tab"] = (b">a\tAC\n" + s4[0][2:] + b"\n>b\tAC\n" + s4[1][2:] + b"\n"
                       b">c\t>d\tAC\n" + s4[2][2:] + b"\n>e\t;remark\n" + s4[3] + b"\n")
    s5 = seqs(5, 40)
    c["empty_records"] = (b">first\n" + s5[0] + b"\n>empty1\n>empty2\n\n>second\n" + s5[1] +
                          b"\n>third\n" + s5[2] + b"\n>empty3\n")
    c["data_before_header"] = s5[0] + b"\n>named\n" + s5[4] + b"\n>also\n" + s5[3] + b"\n"
    c["no_final_newline"] = fasta([b"x", b"y", b"z"], seqs(3, 77))[:-1]
    hb = seqs(3, 50)
    # bytes >= 128 are negative chars: dropped from sequence lines, and they end a name
    c["high_bit"] = (b">cafe\xc3\n" + hb[0][:25] + b"\xc3\xa9" + hb[0][25:] + b"\n"
                     b">naive\xef\n" + hb[1] + b"\n>plain\n" + hb[2][:10] + b"\xff\x80" + hb[2][10:] + b"\n")
    c["high_bit_mid_name"] = (b">caf\xc3\xa9 one\n" + hb[0] + b"\n>na\xefve\n" + hb[1] + b"\n")
    gt = seqs(3, 45)
    c["gt_semicolon_inside_lines"] = (b">p\n" + gt[0][:20] + b">" + gt[0][21:] + b"\n>q\n" +
                                      gt[1][:20] + b";" + gt[1][21:] + b"\n>r\n" + gt[2] + b"\n")
    # a control byte ends a name and the rest of the line is data; NUL is dropped, DEL kept
    c["controls"] = (b">a\x01G\n" + b"ACGTACGTAC\x00GTA\n>c\x1fd\nACGTACGTACGT\x7f\n"
                     b">e\x02f\nACGTACGTACGTA\n>g\x7fh\nbACGTACGTACGTA\n")
    c["cr_only_lines"] = b">a\rACGT\r>b\rACGT\r"
    c["lowercase_only"] = fasta([b"l1", b"l2", b"l3"], [s.lower() for s in seqs(3, 70, b"ACGT")])
    c["length_mismatch"] = b">a\nACGTACGT\n>b\nACGTACGT\n>c\nACGTAC\n"
    c["one_sequence"] = b">only\nACGTACGT\n>empty\n"
    big = seqs(40, 1000, b"ACGTACGTACGTN-")
    c["forty_by_1000"] = fasta([b"s%02d" % i for i in range(40)], big, w=70)
    return c


def run_case(path: str, tmp: str):
    out = os.path.join(tmp, "f.bin")
    r = subprocess.run([HARNESS, "fasta", path, out], capture_output=True)
    if r.returncode != 0:
        return r.returncode, (r.stderr + r.stdout)
    raw = np.fromfile(out, dtype=np.uint8)
    u32 = lambda o: int(np.frombuffer(raw[o:o + 4].tobytes(), dtype=np.uint32)[0])  # noqa: E731
    n, L, bsc, H = u32(0), u32(4), u32(8), u32(12)
    o = 16
    names, lens = [], []
    for _ in range(H):
        k = u32(o)
        names.append(raw[o + 4:o + 4 + k].tobytes())
        lens.append(k)
        o += 4 + k
    body = np.frombuffer(raw[o:].tobytes(), dtype=np.uint32)
    w = bsc * 4
    bits = body[: n * w].reshape(n, w)
    gaps = body[n * w: 2 * n * w].reshape(n, w)
    return 0, (n, L, names, np.array(lens, dtype=np.uint32), bits, gaps)


def main() -> None:
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    os.makedirs(OUT_DIR, exist_ok=True)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, data in cases().items():
            path = os.path.join(OUT_DIR, name + ".fa")
            with open(path, "wb") as fh:
                fh.write(data)
            status, res = run_case(path, tmp)
            out[f"{name}_status"] = np.array(status, dtype=np.int32)
            if status:
                out[f"{name}_message"] = np.frombuffer(res, dtype=np.uint8)
                continue
            n, L, names, lens, bits, gaps = res
            out[f"{name}_n"] = np.array(n, dtype=np.uint32)
            out[f"{name}_L"] = np.array(L, dtype=np.uint32)
            out[f"{name}_names"] = np.frombuffer(b"".join(names), dtype=np.uint8)
            out[f"{name}_name_len"] = lens
            out[f"{name}_bits"] = bits
            out[f"{name}_gaps"] = gaps
    np.savez_compressed(os.path.join(HERE, "fasta_golden.npz"), **out)
    print("FASTA fixtures written:", sorted(k[:-7] for k in out if k.endswith("_status")))


if __name__ == "__main__":
    sys.exit(main())
