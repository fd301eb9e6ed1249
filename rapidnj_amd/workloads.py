"""Seeded synthetic DNA and protein alignments and bootstrap weights for the benchmark
configurations.

The reference ships no generator for its benchmark inputs (its ``sim_seq.cpp`` is a time-seeded
star model, ``src/distanceCalculation/sim_seq.cpp:12-167``), so this module fixes one:

* topology: random joining (coalescent shape) -- start from ``n`` leaf lineages and repeatedly
  join two lineages chosen uniformly at random until one root remains;
* branch lengths: i.i.d. Exponential(mean 0.05), in expected substitutions per site;
* root: i.i.d. uniform over A, C, G, T;
* evolution: Kimura 2-parameter with kappa = alpha/beta = 2 (the rate ratio), rates normalised so
  that alpha + 2*beta = 1 substitution per unit branch length;
* gaps: every leaf cell independently becomes ``'-'`` with probability ``gap``.

Bases are held in the reference's 2-bit code (A=0, G=1, T=2, C=3, ``src/stdinclude.h:25``): a
transition flips the low bit, a transversion flips the high bit. Sequences are returned as an
``(n, L)`` uint8 array of ASCII letters (uppercase ``ACGT`` and ``-``).

Protein alignments (:func:`simulate_protein_alignment`) use the same topology and branch lengths
with a uniform 20-state model: root residues uniform over the 20 amino acids, and along a branch
of length t a site changes with probability (19/20)(1 - exp(-20t/19)) to one of the other 19
residues, uniformly; gaps as above.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

CODE_TO_ASCII = np.frombuffer(b"AGTC", dtype=np.uint8)
GAP = ord("-")


def _random_joining_tree(n: int, rng: np.random.Generator) -> np.ndarray:
    """Parent index of every node; leaves are 0..n-1, the root is 2n-2 (parent -1)."""
    parent = np.full(2 * n - 1, -1, dtype=np.int64)
    lineages = list(range(n))
    nxt = n
    while len(lineages) > 1:
        k = len(lineages)
        a = int(rng.integers(k))
        b = int(rng.integers(k - 1))
        if b >= a:
            b += 1
        na, nb = lineages[a], lineages[b]
        parent[na] = nxt
        parent[nb] = nxt
        # remove a and b (swap-remove, larger index first)
        for idx in sorted((a, b), reverse=True):
            lineages[idx] = lineages[-1]
            lineages.pop()
        lineages.append(nxt)
        nxt += 1
    return parent


def simulate_alignment(n: int, L: int, gap: float, seed: int, kappa: float = 2.0,
                       mean_branch: float = 0.05) -> np.ndarray:
    """Return an ``(n, L)`` uint8 ASCII alignment drawn from the generator described above."""
    if n < 2 or L < 1:
        raise ValueError("need n >= 2 and L >= 1")
    rng = np.random.default_rng(seed)
    parent = _random_joining_tree(n, rng)
    nodes = 2 * n - 1
    blen = rng.exponential(mean_branch, size=nodes)
    beta = 1.0 / (kappa + 2.0)
    alpha = kappa / (kappa + 2.0)
    # depth of each node (root = 0); children have larger ids than parents except leaves,
    # so walk internal nodes from the root down (ids descending), then the leaves.
    depth = np.zeros(nodes, dtype=np.int64)
    for v in range(nodes - 2, -1, -1):
        depth[v] = depth[parent[v]] + 1
    order = np.argsort(depth, kind="stable")
    levels = np.split(order, np.flatnonzero(np.diff(depth[order])) + 1)
    codes = np.empty((nodes, L), dtype=np.uint8)
    codes[nodes - 1] = rng.integers(0, 4, size=L, dtype=np.uint8)
    for lev in levels[1:]:
        t = blen[lev][:, None]
        e4 = np.exp(-4.0 * beta * t)
        e2 = np.exp(-2.0 * (alpha + beta) * t)
        p_ts = 0.25 + 0.25 * e4 - 0.5 * e2
        p_tv = 0.5 - 0.5 * e4
        u = rng.random((lev.size, L))
        flip = np.where(u < p_ts, 1, 0).astype(np.uint8)
        is_tv = (u >= p_ts) & (u < p_ts + p_tv)
        tv_low = rng.integers(0, 2, size=(lev.size, L), dtype=np.uint8)
        flip = np.where(is_tv, 2 | tv_low, flip).astype(np.uint8)
        codes[lev] = codes[parent[lev]] ^ flip
    leaves = CODE_TO_ASCII[codes[:n]]
    if gap > 0:
        mask = rng.random((n, L)) < gap
        leaves[mask] = GAP
    return np.ascontiguousarray(leaves)


def _simulate_args(args) -> np.ndarray:
    return simulate_alignment(*args)


def simulate_many(specs, procs: int = 8) -> list:
    """:func:`simulate_alignment` for every ``(n, L, gap, seed)`` of ``specs``, in a pool of
    ``procs`` spawned interpreter processes (fresh processes: safe after the caller has
    initialised a GPU, unlike forked ones)."""
    import multiprocessing as mp
    specs = list(specs)
    if procs <= 1 or len(specs) <= 1:
        return [simulate_alignment(*s) for s in specs]
    with mp.get_context("spawn").Pool(min(procs, len(specs))) as pool:
        return pool.map(_simulate_args, specs, chunksize=1)


AMINO_ACIDS = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)


def simulate_protein_alignment(n: int, L: int, gap: float, seed: int,
                               mean_branch: float = 0.015) -> np.ndarray:
    """Return an ``(n, L)`` uint8 ASCII protein alignment (20 amino acids and ``-``)."""
    if n < 2 or L < 1:
        raise ValueError("need n >= 2 and L >= 1")
    rng = np.random.default_rng(seed)
    parent = _random_joining_tree(n, rng)
    nodes = 2 * n - 1
    blen = rng.exponential(mean_branch, size=nodes)
    depth = np.zeros(nodes, dtype=np.int64)
    for v in range(nodes - 2, -1, -1):
        depth[v] = depth[parent[v]] + 1
    order = np.argsort(depth, kind="stable")
    levels = np.split(order, np.flatnonzero(np.diff(depth[order])) + 1)
    codes = np.empty((nodes, L), dtype=np.uint8)
    codes[nodes - 1] = rng.integers(0, 20, size=L, dtype=np.uint8)
    for lev in levels[1:]:
        p_change = 0.95 * (1.0 - np.exp(-blen[lev][:, None] / 0.95))
        change = rng.random((lev.size, L)) < p_change
        step = rng.integers(1, 20, size=(lev.size, L), dtype=np.uint8)  # to another residue
        codes[lev] = (codes[parent[lev]] + np.where(change, step, 0)) % 20
    leaves = AMINO_ACIDS[codes[:n]]
    if gap > 0:
        leaves[rng.random((n, L)) < gap] = GAP
    return np.ascontiguousarray(leaves)


def bootstrap_weights(L: int, replicates: int, seed0: int = 0) -> np.ndarray:
    """``(replicates, L)`` uint16 column weights; replicate r ~ Multinomial(L; uniform), seed r+seed0."""
    out = np.empty((replicates, L), dtype=np.uint16)
    for r in range(replicates):
        rng = np.random.default_rng(seed0 + r)
        out[r] = rng.multinomial(L, np.full(L, 1.0 / L)).astype(np.uint16)
    return out


def expand_columns(aln: np.ndarray, weights: np.ndarray) -> np.ndarray:
    """Alignment whose columns are column c repeated weights[c] times, in column order."""
    idx = np.repeat(np.arange(aln.shape[1]), weights.astype(np.int64))
    return np.ascontiguousarray(aln[:, idx])


def write_fasta(path: str, aln: np.ndarray, width: int = 80) -> None:
    with open(path, "wb") as f:
        for i in range(aln.shape[0]):
            f.write(b">t%d\n" % i)
            row = aln[i].tobytes()
            for s in range(0, len(row), width):
                f.write(row[s:s + width])
                f.write(b"\n")


def read_fasta(path: str) -> tuple[list[str], np.ndarray]:
    """Minimal FASTA reader for files written by :func:`write_fasta` and the reference fixtures.

    Follows the reference's line rules (``dataloaderFasta.hpp:78-99,201-217``): ``;`` lines are
    comments, bytes <= 32 inside sequence lines are dropped, names run to the first byte < 32.
    """
    names: list[str] = []
    seqs: list[bytes] = []
    cur: list[bytes] = []
    with open(path, "rb") as f:
        for line in f:
            if line.startswith(b";"):
                continue
            if line.startswith(b">"):
                if cur:
                    seqs.append(b"".join(cur))
                    cur = []
                name = bytearray()
                for ch in line[1:]:
                    if ch <= 31:
                        break
                    name.append(ch)
                names.append(name.decode("latin-1"))
            else:
                cur.append(bytes(ch for ch in line if ch > 32))
    if cur:
        seqs.append(b"".join(cur))
    L = len(seqs[0])
    if any(len(s) != L for s in seqs):
        raise ValueError("sequences of unequal length")
    return names[: len(seqs)], np.frombuffer(b"".join(seqs), dtype=np.uint8).reshape(len(seqs), L).copy()


@dataclass(frozen=True)
class Workload:
    name: str
    n: int
    L: int
    gap: float
    model: str  # "jc" or "k2p" (Kimura protein for protein workloads)
    seed: int
    replicates: int = 0
    protein: bool = False

    @property
    def site_pairs(self) -> int:
        return self.n * (self.n - 1) // 2 * self.L * max(1, self.replicates)


WORKLOADS = {
    "small_jc": Workload("100x1000 JC, 2% gaps", 100, 1000, 0.02, "jc", 1),
    "plain": Workload("2000x10000 K2P, 5% gaps", 2000, 10000, 0.05, "k2p", 2),
    "large": Workload("20000x5000 K2P, 5% gaps", 20000, 5000, 0.05, "k2p", 3),
    "bootstrap": Workload("1000x5000 K2P, 100 bootstrap replicates", 1000, 5000, 0.05, "k2p", 4, 100),
    "protein": Workload("20000x1250 Kimura protein, 10% gaps", 20000, 1250, 0.10, "k2p", 6,
                        protein=True),
}


def make_alignment(wl: Workload) -> np.ndarray:
    """The workload's alignment (DNA or protein generator)."""
    if wl.protein:
        return simulate_protein_alignment(wl.n, wl.L, wl.gap, wl.seed)
    return simulate_alignment(wl.n, wl.L, wl.gap, wl.seed)


@dataclass(frozen=True)
class ServedJob:
    job_id: int
    arrival_s: float
    n: int
    L: int
    gap: float
    model: str
    seed: int


def served_stream(jobs: int = 512, rate: float = 20.0, seed: int = 5) -> list[ServedJob]:
    """Poisson arrivals; n log-uniform in [50, 5000]; L uniform in [500, 30000]; gaps U(0, 8%)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(jobs):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(math.exp(rng.uniform(math.log(50), math.log(5000)))))
        L = int(rng.integers(500, 30001))
        g = float(rng.uniform(0.0, 0.08))
        model = "jc" if rng.random() < 0.5 else "k2p"
        out.append(ServedJob(k, t, n, L, g, model, 1000 + k))
    return out
