"""ctypes binding of the CPU oracle (oracle/rnj_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this module, and
only to check or time the reference algorithm; the product (``rapidnj_amd``) never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librnj_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

JC, K2P = 0, 1
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u64_3 = ctypes.c_uint64 * 3

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.rnj_oracle_bitstrings_count.restype = ctypes.c_uint
        L.rnj_oracle_bitstrings_count.argtypes = [ctypes.c_uint]
        L.rnj_oracle_is_dna.restype = ctypes.c_int
        L.rnj_oracle_is_dna.argtypes = [ctypes.c_char_p, ctypes.c_uint]
        L.rnj_oracle_pack.restype = None
        L.rnj_oracle_pack.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                      _u32p, _u32p]
        for name in ("rnj_oracle_count_chars", "rnj_oracle_count_naive"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint, _u64_3]
        for name in ("rnj_oracle_count_words", "rnj_oracle_count_swar"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint, _u64_3]
        for name in ("rnj_oracle_k2p", "rnj_oracle_jc"):
            f = getattr(L, name)
            f.restype = ctypes.c_float
            f.argtypes = [ctypes.c_uint64] * 3
        L.rnj_oracle_counts_triangle.restype = ctypes.c_int
        L.rnj_oracle_counts_triangle.argtypes = [_u32p, _u32p, ctypes.c_uint, ctypes.c_uint,
                                                 ctypes.c_int, _u32p, _u32p, _u32p]
        L.rnj_oracle_distance_matrix.restype = ctypes.c_int
        L.rnj_oracle_distance_matrix.argtypes = [_u32p, _u32p, ctypes.c_uint, ctypes.c_uint,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        for name in ("rnj_oracle_count_protein", "rnj_oracle_count_protein_naive"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint, _u64_3]
        L.rnj_oracle_protein_gpu_counts.restype = None
        L.rnj_oracle_protein_gpu_counts.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                                    ctypes.c_uint, _u64_3]
        for f in (L.rnj_oracle_kimura_protein, L.rnj_oracle_jc_protein):
            f.restype = ctypes.c_float
            f.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.rnj_oracle_protein_matrix.restype = ctypes.c_int
        L.rnj_oracle_protein_matrix.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_uint,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.POINTER(ctypes.c_float)]
        L.rnj_oracle_parse_fasta.restype = ctypes.c_int
        L.rnj_oracle_parse_fasta.argtypes = [ctypes.c_char_p, ctypes.c_size_t] + \
            [ctypes.c_void_p] * 9
        L.rnj_oracle_parse_stockholm.restype = ctypes.c_int
        L.rnj_oracle_parse_stockholm.argtypes = [ctypes.c_char_p, ctypes.c_size_t] + \
            [ctypes.c_void_p] * 8
        L.rnj_oracle_correct_triangle.restype = ctypes.c_int
        L.rnj_oracle_correct_triangle.argtypes = [_u32p, _u32p, _u32p, ctypes.c_uint, ctypes.c_int,
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.POINTER(ctypes.c_float)]
        L.rnj_oracle_nj_skip_audit.restype = ctypes.c_int
        L.rnj_oracle_nj_skip_audit.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_void_p]
        L.rnj_oracle_nj.restype = ctypes.c_int
        L.rnj_oracle_nj.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_int] + \
            [ctypes.c_void_p] * 5 + [ctypes.POINTER(ctypes.c_float)]
        L.rnj_oracle_newick.restype = ctypes.c_size_t
        L.rnj_oracle_newick.argtypes = [ctypes.c_uint] + [ctypes.c_void_p] * 5 + \
            [ctypes.c_float, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        _lib = L
    return _lib


def model_id(model: str) -> int:
    return {"jc": JC, "k2p": K2P, "kim": K2P}[model]


def bitstrings_count(L: int) -> int:
    return lib().rnj_oracle_bitstrings_count(L)


def is_dna(seq: bytes) -> bool:
    return bool(lib().rnj_oracle_is_dna(seq, len(seq)))


def pack(aln: np.ndarray) -> tuple[np.ndarray, np.ndarray, int]:
    """Reference-layout packing of an (n, L) uint8 alignment -> (bits, gaps, bsc)."""
    aln = np.ascontiguousarray(aln, dtype=np.uint8)
    n, L = aln.shape
    bsc = bitstrings_count(L)
    bits = np.zeros((n, bsc * 4), dtype=np.uint32)
    gaps = np.zeros((n, bsc * 4), dtype=np.uint32)
    lib().rnj_oracle_pack(aln.tobytes(), n, L, bsc, bits, gaps)
    return bits, gaps, bsc


def count_chars(a: bytes, b: bytes) -> tuple[int, int, int]:
    out = _u64_3()
    lib().rnj_oracle_count_chars(a, b, len(a), out)
    return tuple(out)


def count_naive(a: bytes, b: bytes) -> tuple[int, int, int]:
    out = _u64_3()
    lib().rnj_oracle_count_naive(a, b, len(a), out)
    return tuple(out)


def _row_ptr(arr: np.ndarray, i: int) -> int:
    return int(arr.ctypes.data) + int(i) * int(arr.strides[0])


def count_words(bits, gaps, i, j) -> tuple[int, int, int]:
    out = _u64_3()
    lib().rnj_oracle_count_words(_row_ptr(bits, i), _row_ptr(gaps, i), _row_ptr(bits, j),
                                 _row_ptr(gaps, j), bits.shape[1], out)
    return tuple(out)


def count_swar(bits, gaps, i, j) -> tuple[int, int, int]:
    out = _u64_3()
    lib().rnj_oracle_count_swar(_row_ptr(bits, i), _row_ptr(gaps, i), _row_ptr(bits, j),
                                _row_ptr(gaps, j), bits.shape[1] // 4, out)
    return tuple(out)


def k2p(ts: int, tv: int, valid: int) -> float:
    return lib().rnj_oracle_k2p(ts, tv, valid)


def jc(ts: int, tv: int, valid: int) -> float:
    return lib().rnj_oracle_jc(ts, tv, valid)


def counts_triangle(aln: np.ndarray, threads: int = 8):
    bits, gaps, bsc = pack(aln)
    n = aln.shape[0]
    m = n * (n - 1) // 2
    ts = np.zeros(max(m, 1), dtype=np.uint32)
    tv = np.zeros_like(ts)
    va = np.zeros_like(ts)
    rc = lib().rnj_oracle_counts_triangle(bits, gaps, n, bsc, threads, ts, tv, va)
    if rc != 0:
        raise RuntimeError("oracle counts failed")
    return ts[:m], tv[:m], va[:m]


def distance_matrix(aln: np.ndarray, model: str, threads: int = 8):
    """-> (triangle before post-processing, full post-processed n x n matrix, max)."""
    bits, gaps, bsc = pack(aln)
    n = aln.shape[0]
    m = n * (n - 1) // 2
    tri = np.zeros(max(m, 1), dtype=np.float32)
    full = np.zeros((n, n), dtype=np.float32)
    mx = ctypes.c_float()
    rc = lib().rnj_oracle_distance_matrix(bits, gaps, n, bsc, model_id(model), threads,
                                          tri.ctypes.data, full.ctypes.data, ctypes.byref(mx))
    if rc != 0:
        raise RuntimeError("oracle distance matrix failed")
    return tri[:m], full, mx.value


def correct_triangle(ts, tv, valid, n: int, model: str, with_full: bool = True):
    """-> (triangle before post-processing, full post-processed n x n matrix or None, max)."""
    m = n * (n - 1) // 2
    tri = np.zeros(max(m, 1), dtype=np.float32)
    full = np.zeros((n, n), dtype=np.float32) if with_full else None
    mx = ctypes.c_float()
    ts = np.ascontiguousarray(ts, dtype=np.uint32)
    tv = np.ascontiguousarray(tv, dtype=np.uint32)
    valid = np.ascontiguousarray(valid, dtype=np.uint32)
    if m == 0:
        ts = tv = valid = np.zeros(1, dtype=np.uint32)
    lib().rnj_oracle_correct_triangle(ts, tv, valid, n, model_id(model), tri.ctypes.data,
                                      full.ctypes.data if full is not None else None,
                                      ctypes.byref(mx))
    return tri[:m], full, mx.value


def triangle_to_full(tri: np.ndarray, n: int, maxd: float | None = None) -> np.ndarray:
    """Post-process a raw upper triangle the way the reference does (symmetric, diag 0, -1 -> 2max)."""
    full = np.zeros((n, n), dtype=np.float32)
    iu = np.triu_indices(n, 1)
    t = tri.astype(np.float32).copy()
    if maxd is None:
        maxd = float(np.max(np.concatenate([[np.float32(0)], t[t > 0]])))
    t[t == -1] = np.float32(maxd) * np.float32(2)
    full[iu] = t
    full[(iu[1], iu[0])] = t
    return full


def count_protein(a: bytes, b: bytes) -> tuple[int, int, int]:
    out = _u64_3()
    lib().rnj_oracle_count_protein(a, b, len(a), out)
    return int(out[0]), int(out[1]), int(out[2])


def count_protein_naive(a: bytes, b: bytes) -> tuple[int, int, int]:
    """The -f path's Hamming counter (simpleDistanceCalculator::calculateHammingDistance)."""
    out = _u64_3()
    lib().rnj_oracle_count_protein_naive(a, b, len(a), out)
    return int(out[0]), int(out[1]), int(out[2])


def protein_gpu_counts(a: bytes, b: bytes) -> tuple[int, int]:
    """The reference's dormant protein kernel on two staged rows: (mismatches, gap positions)."""
    out = _u64_3()
    lib().rnj_oracle_protein_gpu_counts(a, b, len(a), out)
    return int(out[0]), int(out[1])


def protein_staging(aln: np.ndarray) -> tuple[np.ndarray, int, int]:
    """The host array dataloader::initGPUData stages for a protein alignment (dataloader.hpp
    :38-70): the loader's resolved rows padded with '-' to bitStringsCount = L/16 + 8 blocks of
    16 characters (dataloaderFasta.hpp storeSequenceFastDist), copied into zeroed rows of bsStride
    = roundup(ceil(L/4), 128) words, of which min(4 bitStringsCount, bsStride) words are copied.
    -> (n x bsStride uint32, bsStride, paddedLength = roundup(L, 128))."""
    aln = np.ascontiguousarray(aln, dtype=np.uint8)
    n, L = aln.shape
    res = aln.copy()
    res[np.isin(res, np.frombuffer(b"-.XxzZbBJj?", dtype=np.uint8))] = ord("-")
    bsc = L // 16 + 8
    loader = np.full((n, bsc * 16), ord("-"), dtype=np.uint8)
    loader[:, :L] = res
    stride = ((L + 3) // 4 + 127) & ~127
    staged = np.zeros((n, stride * 4), dtype=np.uint8)
    words = min(bsc * 4, stride)
    staged[:, :words * 4] = loader[:, :words * 4]
    return staged.view("<u4").reshape(n, stride), stride, (L + 127) & ~127


def kimura_protein(dist: int, valid: int) -> float:
    return lib().rnj_oracle_kimura_protein(dist, valid)


def jc_protein(dist: int, valid: int) -> float:
    return lib().rnj_oracle_jc_protein(dist, valid)


def protein_counts_triangle(aln: np.ndarray, naive: bool = False):
    """(dist, valid) upper triangles (uint64) of an (n, L) protein character alignment, counted
    by bitDistanceProtein or (naive) by the -f path's Hamming counter."""
    aln = np.ascontiguousarray(aln, dtype=np.uint8)
    n = aln.shape[0]
    rows = [aln[i].tobytes() for i in range(n)]
    count = count_protein_naive if naive else count_protein
    dist, valid = [], []
    for i in range(n):
        for j in range(i + 1, n):
            d, _, v = count(rows[i], rows[j])
            dist.append(d)
            valid.append(v)
    return np.array(dist, dtype=np.uint64), np.array(valid, dtype=np.uint64)


def protein_matrix(aln: np.ndarray, model: str, naive: bool = False):
    """(full post-processed n x n matrix, max) for an (n, L) protein character alignment
    (naive: counted by the -f path's Hamming counter)."""
    aln = np.ascontiguousarray(aln, dtype=np.uint8)
    n, L = aln.shape
    full = np.zeros((n, n), dtype=np.float32)
    mx = ctypes.c_float()
    rc = lib().rnj_oracle_protein_matrix(aln.tobytes(), n, L, model_id(model), int(naive),
                                         full.ctypes.data, ctypes.byref(mx))
    if rc != 0:
        raise RuntimeError("oracle protein matrix failed")
    return full, mx.value


def nj(matrix: np.ndarray, negative_branches: bool = False):
    """rapidNJ's joins on an n x n float32 matrix: (left, right, llen, rlen, last (2,), last_dist)."""
    m = np.ascontiguousarray(matrix, dtype=np.float32)
    n = m.shape[0]
    k = max(n - 2, 1)
    left, right = np.zeros(k, dtype=np.int32), np.zeros(k, dtype=np.int32)
    llen, rlen = np.zeros(k, dtype=np.float64), np.zeros(k, dtype=np.float64)
    last = np.zeros(2, dtype=np.int32)
    ld = ctypes.c_float()
    rc = lib().rnj_oracle_nj(m.ctypes.data, n, int(negative_branches), left.ctypes.data,
                             right.ctypes.data, llen.ctypes.data, rlen.ctypes.data,
                             last.ctypes.data, ctypes.byref(ld))
    if rc != 0:
        raise RuntimeError(f"oracle nj failed ({rc})")
    return left[: n - 2], right[: n - 2], llen[: n - 2], rlen[: n - 2], last, ld.value


def nj_skip_audit(matrix: np.ndarray, negative_branches: bool = False,
                  drop_drift: bool = False) -> dict:
    """The device NJ's row-skipping rule checked along rapidNJ's own run (O(n^3));
    drop_drift plants a mistake (the bound without the separations' drift)."""
    m = np.ascontiguousarray(matrix, dtype=np.float32)
    out = np.zeros(4, dtype=np.int64)
    rc = lib().rnj_oracle_nj_skip_audit(m.ctypes.data, m.shape[0], int(negative_branches),
                                        1 if drop_drift else 0, out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle nj audit failed ({rc})")
    return {"bound_violations": int(out[0]), "record_violations": int(out[1]),
            "skipped": int(out[2]), "scanned": int(out[3])}


def newick(n: int, tree, names=None) -> bytes:
    """polytree::serialize_tree text of a tree from :func:`nj` (names default to t0..t{n-1})."""
    left, right, llen, rlen, last, ld = tree
    names = names or [b"t%d" % i for i in range(n)]
    arr = (ctypes.c_char_p * n)(*names)
    args = [n, np.ascontiguousarray(left, dtype=np.int32).ctypes.data,
            np.ascontiguousarray(right, dtype=np.int32).ctypes.data,
            np.ascontiguousarray(llen, dtype=np.float64).ctypes.data,
            np.ascontiguousarray(rlen, dtype=np.float64).ctypes.data,
            np.ascontiguousarray(last, dtype=np.int32).ctypes.data, ctypes.c_float(ld), arr]
    need = lib().rnj_oracle_newick(*args, None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    lib().rnj_oracle_newick(*args, buf, need + 1)
    return buf.raw[:need]


class FastaError(ValueError):
    """The reference's exits while reading FASTA: kind 1 = a stored record's length differs from
    the first one's (seq, length), kind 2 = fewer than 2 sequences."""

    def __init__(self, kind: int, seq: int = 0, length: int = 0):
        super().__init__(f"fasta error {kind}: sequence {seq} length {length}")
        self.kind, self.seq, self.length = kind, seq, length


def parse_fasta(data: bytes) -> tuple[list[bytes], np.ndarray]:
    """(names of every '>' record in file order, (n, L) uint8 kept bytes of the stored records)."""
    raw = bytes(data)
    u = ctypes.c_uint
    n, L, H, bad, bad_len = u(), u(), u(), u(), u()
    nb = ctypes.c_size_t()
    rc = lib().rnj_oracle_parse_fasta(raw, len(raw), ctypes.byref(n), ctypes.byref(L),
                                      ctypes.byref(H), ctypes.byref(nb), ctypes.byref(bad),
                                      ctypes.byref(bad_len), None, None, None)
    if rc == 1:
        raise FastaError(1, bad.value, bad_len.value)
    if rc == 2:
        raise FastaError(2, n.value)
    rows = np.zeros((n.value, L.value), dtype=np.uint8)
    names = ctypes.create_string_buffer(max(nb.value, 1))
    lens = (ctypes.c_uint * max(H.value, 1))()
    lib().rnj_oracle_parse_fasta(raw, len(raw), ctypes.byref(n), ctypes.byref(L), ctypes.byref(H),
                                 ctypes.byref(nb), ctypes.byref(bad), ctypes.byref(bad_len),
                                 rows.ctypes.data_as(ctypes.c_void_p), names, lens)
    out, o = [], 0
    for k in range(H.value):
        out.append(names.raw[o:o + lens[k]])
        o += lens[k]
    return out, rows


class StockholmError(ValueError):
    """The reference's exits while reading Stockholm: kind 1 = a sequence's length differs from
    the first one's (seq, length), 2 = fewer than 2 sequences, 3 = a sequence of zero length
    (seq), 4 = the input ends before "//"."""

    def __init__(self, kind: int, seq: int = 0, length: int = 0):
        super().__init__(f"stockholm error {kind}: sequence {seq} length {length}")
        self.kind, self.seq, self.length = kind, seq, length


def parse_stockholm(data: bytes) -> tuple[list[bytes], np.ndarray]:
    """(the name of every sequence, (n, L) uint8 kept bytes of every sequence)."""
    raw = bytes(data)
    u = ctypes.c_uint
    n, L, bad, bad_len = u(), u(), u(), u()
    nb = ctypes.c_size_t()
    args = (raw, len(raw), ctypes.byref(n), ctypes.byref(L), ctypes.byref(nb), ctypes.byref(bad),
            ctypes.byref(bad_len))
    rc = lib().rnj_oracle_parse_stockholm(*args, None, None, None)
    if rc:
        raise StockholmError(rc, bad.value, bad_len.value)
    rows = np.zeros((n.value, L.value), dtype=np.uint8)
    names = ctypes.create_string_buffer(max(nb.value, 1))
    lens = (ctypes.c_uint * max(n.value, 1))()
    lib().rnj_oracle_parse_stockholm(*args, rows.ctypes.data_as(ctypes.c_void_p), names, lens)
    out, o = [], 0
    for k in range(n.value):
        out.append(names.raw[o:o + lens[k]])
        o += lens[k]
    return out, rows


def ref_harness_available() -> bool:
    return os.path.exists(REF_HARNESS)
