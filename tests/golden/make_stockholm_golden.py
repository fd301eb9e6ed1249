"""Write the Stockholm reader fixtures: tests/golden/stockholm/<case>.sth and
tests/golden/stockholm_golden.npz.

Run in the build container only (needs /root/reference and ``make -C oracle ref``):

    python tests/golden/make_stockholm_golden.py

Each input file is generated here (seeded) to exercise one behaviour of the reference reader
dataloaderStockholm::parseData / parseSequenceName / parseSequenceLineFastDist
(dataloaderStockholm.hpp:83-111, 118-132, 183-236, 294-314); the expected outputs come from
oracle/_ref/ref_harness `sth` mode, i.e. the reference's own reader compiled from its sources: the
name of every sequence, the sequence count and length, and the reference's packed bit strings /
gap filters. Cases the reference rejects record its exit status and message instead.

npz keys per case <c>: as tests/golden/fasta_golden.npz (make_fasta_golden.py).
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_fasta_golden import HARNESS, seqs  # noqa: E402

OUT_DIR = os.path.join(HERE, "stockholm")
RNG = np.random.default_rng(9)


def sth(names, ss, sep: bytes = b" ", eol: bytes = b"\n", header: bool = True,
        end: bytes = b"//\n") -> bytes:
    out = b"# STOCKHOLM 1.0" + eol + b"#=GF ID test" + eol if header else b""
    width = max(len(n) for n in names) + 2
    for nm, s in zip(names, ss):
        pad = sep * (width - len(nm)) if sep == b" " else sep
        out += nm + pad + s + eol
    return out + end


def cases() -> dict[str, bytes]:
    c = {}
    s8 = seqs(8, 120)
    n8 = [b"taxon_%d" % i for i in range(8)]
    c["basic"] = sth(n8, s8)
    c["tabs"] = sth(n8, s8, sep=b"\t\t")
    c["crlf"] = sth(n8, s8, eol=b"\r\n", end=b"//\r\n")
    c["no_header"] = sth(n8, s8, header=False)
    # markup of every kind, blank lines, control bytes before lines, '#' and '/' inside data
    s5 = seqs(5, 64)
    body = b"# STOCKHOLM 1.0\n\n#=GF AC PF00001\n#=GS a DE first\n"
    for i, s in enumerate(s5):
        body += b"\n\t\r" + (b"s%d   " % i) + s[:20] + b"#" + s[21:40] + b"/" + s[41:] + b"\n"
        body += b"#=GR s%d SS ....\n" % i
    c["markup_blank_controls"] = body + b"#=GC SS_cons ....\n//\n"
    # a name alone on its line: the sequence is the next line
    s4 = seqs(4, 50)
    c["name_own_line"] = (b"# STOCKHOLM 1.0\nalpha\n" + s4[0] + b"\nbeta " + s4[1] +
                          b"\ngamma\n\t" + s4[2] + b"\ndelta\t" + s4[3] + b"\n//\n")
    # a line starting with a space has an empty name; '/x' starts a name with '/'
    s3 = seqs(3, 40)
    c["empty_name_and_slash"] = b"#\n " + s3[0] + b"\n/slash " + s3[1] + b"\n/ " + s3[2] + b"\n//\n"
    # bytes >= 128: skipped like control bytes before a line, end a name, dropped from data
    hb = seqs(3, 48)
    c["high_bit"] = (b"\xc3\xa9first " + hb[0][:20] + b"\xff\x80" + hb[0][20:] + b"\n"
                     b"sec\xc3\xa9 " + hb[1] + b"\nthird\xef" + hb[2] + b"\n//\n")
    # a byte >= 128 inside a name ends it, and the name's rest becomes data (here: too long)
    c["high_bit_mid_name"] = (b"first " + hb[0] + b"\nsec\xc3\xa9ond " + hb[1] + b"\n//\n")
    # a control byte ends a name: the data continues on the same line
    c["control_ends_name"] = b"a\x01ACGTACGTAC\nb\x1fACGTACGTAC\nc\x7fd ACGTACGTAC\n//\n"
    # everything after "//" is ignored, including malformed lines
    c["after_end_ignored"] = sth(n8[:3], s8[:3]) + b"x ACG\nno data\n\n"
    # the same names in two blocks (interleaved Stockholm) read as four sequences
    b2 = seqs(2, 60)
    c["two_blocks"] = (b"# STOCKHOLM 1.0\nu " + b2[0][:30] + b"\nv " + b2[1][:30] + b"\n\n"
                       b"u " + b2[0][30:] + b"\nv " + b2[1][30:] + b"\n//\n")
    c["lowercase_n_gaps"] = sth([b"l%d" % i for i in range(4)],
                                [s.lower() for s in seqs(4, 70, b"ACGTN-.")])
    big = seqs(40, 1000, b"ACGTACGTACGTN-")
    c["forty_by_1000"] = sth([b"s%02d" % i for i in range(40)], big)
    # rejections
    c["length_mismatch"] = b"a ACGTACGT\nb ACGTACGT\nc ACGTAC\n//\n"
    c["one_sequence"] = b"# STOCKHOLM 1.0\nonly ACGTACGT\n//\n"
    c["zero_length"] = b"a ACGTACGT\nname\n\nb ACGTACGT\n//\n"
    c["zero_length_tab"] = b"a ACGTACGT\nname\t\nb ACGTACGT\n//\n"
    c["no_end_marker"] = sth(n8[:3], s8[:3], end=b"")
    c["ends_in_data"] = b"a ACGTACGT\nb ACGTACGT"
    return c


def run_case(path: str, tmp: str):
    import subprocess
    out = os.path.join(tmp, "s.bin")
    r = subprocess.run([HARNESS, "sth", path, out], capture_output=True)
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
    return 0, (n, L, names, np.array(lens, dtype=np.uint32), body[: n * w].reshape(n, w),
               body[n * w: 2 * n * w].reshape(n, w))


def main() -> None:
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    os.makedirs(OUT_DIR, exist_ok=True)
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, data in cases().items():
            path = os.path.join(OUT_DIR, name + ".sth")
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
    np.savez_compressed(os.path.join(HERE, "stockholm_golden.npz"), **out)
    print("Stockholm fixtures written:", sorted(k[:-7] for k in out if k.endswith("_status")))


if __name__ == "__main__":
    sys.exit(main())
