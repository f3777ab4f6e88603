"""Generates exp_table.inc: the 128-entry table of 2^(k/128) used by the
device restatement of glibc's exp (lbm_exp.h).

Each entry pair is (T[k], H[k] - (k << 52)/128) as IEEE-754 bit patterns,
where H[k] is 2^(k/128) rounded to the nearest double and T[k] the relative
tail (2^(k/128) - H[k]) / H[k] rounded to the nearest double -- the table
definition of the exp algorithm in ARM optimized-routines (math/exp_data.c),
which glibc >= 2.28 uses.  Values are recomputed here from that definition
with 80-digit decimal arithmetic.

    python mesobasic_amd/csrc/gen_exp_table.py
"""
import os
import struct
from decimal import Decimal, getcontext


def u64(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main() -> None:
    getcontext().prec = 80
    n = 128
    ln2 = Decimal(2).ln()
    rows = []
    for k in range(n):
        exact = (ln2 * k / n).exp()
        h = float(exact)                       # correctly rounded (CPython)
        t = float((exact - Decimal(h)) / Decimal(h))
        rows.append(f"    0x{u64(t):016x}ull, 0x{(u64(h) - ((k << 52) // n)) & (2**64 - 1):016x}ull,")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "exp_table.inc")
    with open(out, "w") as fh:
        fh.write("/* Generated by gen_exp_table.py: 2^(k/128) = H*(1+T), pairs (T, H-(k<<45)). */\n")
        fh.write("\n".join(rows) + "\n")


if __name__ == "__main__":
    main()
