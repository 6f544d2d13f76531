"""Writes pow5_table.h: 128-bit normalised approximations of 5^q for
q in [-342, 308], the table Eisel-Lemire decimal->binary conversion needs
(D. Lemire, "Number Parsing at a Gigabyte per Second", SPE 2021, sec. 5).

Entry q is a 128-bit integer in [2^127, 2^128): for q >= 0 it is 5^q
shifted so its top bit is bit 127 and truncated; for q < 0 it is
floor(2^b / 5^-q) + 1 for a b that places the top bit at 127 (a rounded-up
reciprocal), so the product with a 64-bit significand never underestimates.
"""
import os

LO, HI = -342, 308


def entry(q):
    if q >= 0:
        v = 5 ** q
        while v < (1 << 127):
            v <<= 1
        while v >= (1 << 128):
            v >>= 1
        return v
    p = 5 ** (-q)
    z = p.bit_length()
    b = z + 127 if q >= -27 else 2 * z + 128
    c = (1 << b) // p + 1
    while c >= (1 << 128):
        c >>= 1
    return c


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pow5_table.h")
    lines = ["// Generated by gen_pow5.py; do not edit.",
             "#pragma once", "#include <cstdint>",
             f"#define WJQ_POW5_MIN_Q ({LO})", f"#define WJQ_POW5_MAX_Q ({HI})",
             "WJQ_TABLE_QUAL uint64_t wjq_pow5_128[%d] = {" % (2 * (HI - LO + 1))]
    for q in range(LO, HI + 1):
        v = entry(q)
        assert (1 << 127) <= v < (1 << 128), q
        lines.append("  0x%016xULL, 0x%016xULL, // 5^%d" % (v >> 64, v & ((1 << 64) - 1), q))
    lines.append("};")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
