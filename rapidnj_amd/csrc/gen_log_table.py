#!/usr/bin/env python3
"""Writes log_table.inc: the 128-entry table of the FP64 logarithm the tile kernels' epilogues
use (log_table in device_common.hpp).

Entry i covers mantissas m in [0.5 + i/256, 0.5 + (i+1)/256) and is anchored at its right end
c_i = 0.5 + (i+1)/256 (exact in binary; the last is 1, so that near z = 1 the reduced argument
is z - 1 itself). It holds 1/c_i and ln(c_i), each rounded once to the nearest double from a
50-digit decimal evaluation (Python's decimal module: ln() is correctly rounded at the working
precision, and float(Decimal) rounds correctly).

    python3 rapidnj_amd/csrc/gen_log_table.py > rapidnj_amd/csrc/log_table.inc
"""
from decimal import Decimal, getcontext

getcontext().prec = 50


def entries():
    for i in range(128):
        c = Decimal(1) / 2 + Decimal(i + 1) / 256
        yield float(Decimal(1) / c), float(c.ln())


def main() -> None:
    print("// Generated by gen_log_table.py: {1/c_i, ln(c_i)}, c_i = 0.5 + (i+1)/256, i = 0..127.")
    for inv, ln in entries():
        print(f"    {{{inv.hex()}, {ln.hex()}}},")


if __name__ == "__main__":
    main()
