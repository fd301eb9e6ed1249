"""rapidnj_amd -- MI355X (gfx950) pairwise evolutionary-distance engine for DNA alignments.

Drop-in for rapidNJ's distance engine (src/distanceCalculation): exact transition /
transversion / valid-site counts per taxon pair from bit-packed alignments, the Jukes-Cantor
and Kimura 2-parameter corrections, the reference post-process, batched bootstrap and served
jobs. Compute runs in HIP kernels in librapidnj_amd.so (C ABI: include/rapidnj_amd.h); this
package is its Python binding (``rapidnj_amd.native``, which loads the library on first use
and raises if it is missing) plus the seeded workload generator.
"""
from . import native  # noqa: F401
from .workloads import (WORKLOADS, Workload, bootstrap_weights, expand_columns,  # noqa: F401
                        read_fasta, served_stream, simulate_alignment, write_fasta)

__all__ = ["WORKLOADS", "Workload", "bootstrap_weights", "expand_columns", "native",
           "read_fasta", "served_stream", "simulate_alignment", "write_fasta"]
