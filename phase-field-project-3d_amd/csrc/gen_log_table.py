"""Writes pf_log_table.h: the table of the FP64 logarithm log_tab() in pf_math.h.

Entry j (j = -37 .. 53) belongs to c_j = 1 + j/128.  It holds
  inv = 1/c_j rounded to the nearest double, and
  L   = -ln(inv), computed in 60-digit decimal arithmetic from that double
        and rounded to the nearest double,
so that ln(m) = L + ln(m * inv) holds for every m, whatever the rounding of
inv.  Run from any directory: python3 gen_log_table.py [output path]
"""
import decimal
import os
import sys

LO, HI = -37, 53  # trunc(128 (m - 1)) for m in [sqrt(1/2), sqrt(2))


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    decimal.getcontext().prec = 60
    rows = []
    for j in range(LO, HI + 1):
        inv = 1.0 / (1.0 + j / 128.0)  # 1 + j/128 is exact; one rounding
        L = -(decimal.Decimal(inv).ln())
        rows.append((j, inv, float(L)))
    out = argv[0] if argv else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                            "pf_log_table.h")
    with open(out, "w") as f:
        f.write("// pf_log_table.h -- written by gen_log_table.py; do not edit.\n")
        f.write("// {1/c_j, -ln(1/c_j)} for c_j = 1 + j/128, j = %d .. %d (see log_tab).\n" % (LO, HI))
        f.write("#pragma once\n\nnamespace pf {\n\n")
        f.write("constexpr int kLogTabLo = %d, kLogTabN = %d;\n\n" % (LO, HI - LO + 1))
        f.write("constexpr double kLogTab[kLogTabN][2] = {\n")
        for j, inv, L in rows:
            f.write("    {%s, %s},  // j = %d\n" % (inv.hex(), L.hex(), j))
        f.write("};\n\n}  // namespace pf\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
