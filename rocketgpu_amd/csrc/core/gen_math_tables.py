"""Generates rg_math_tables.inc: the lookup tables of the FP32 trigonometric / power kernels in rg_math.hpp.

Every entry is a mathematical constant computed here with 80-digit decimal arithmetic and rounded exactly to a
double through fractions.Fraction: to nearest, except the atan table, which the UCRT kernel stores rounded
toward zero:

  atan_tab[k]  = atan((k + 16) / 256)          k = 0..240   (the atan2f table lookup; rounded toward zero)
  exp2_tab[k]  = 2 ** (k / 64)                 k = 0..63    (the powf exp step)
  inv_c[k]     = 512 / (256 + k)               k = 0..256   (1 / c,  c = 1/2 + k/512, the powf log step)
  log_c[k]     = ln((256 + k) / 256)           k = 0..256   (ln(2 c))
  two_over_pi[k] = bits 8k - 23 .. 8k + 8 of the binary fraction of 2 / pi (32-bit windows)  k = 0..23
                 (the large-argument reduction of sinf / cosf; pi from Machin's formula)

Run:  python rocketgpu_amd/csrc/core/gen_math_tables.py > rocketgpu_amd/csrc/core/rg_math_tables.inc
tests/test_math_tables.py re-derives the tables and checks the committed file is what this script prints.
"""
from decimal import Decimal, getcontext
from fractions import Fraction
import math

getcontext().prec = 80


def d_atan(x: Decimal) -> Decimal:
    # Argument halving: atan(x) = 2 atan(x / (1 + sqrt(1 + x^2))), then the Taylor series.
    halvings = 0
    while x > Decimal("0.05"):
        x = x / (1 + (1 + x * x).sqrt())
        halvings += 1
    term, total, x2, n = x, x, x * x, 1
    while True:
        term = -term * x2
        add = term / (2 * n + 1)
        if abs(add) < Decimal(10) ** -78:
            break
        total += add
        n += 1
    return total * (2 ** halvings)


def to_double(d: Decimal) -> float:
    return float(Fraction(d))


def to_double_toward_zero(d: Decimal) -> float:
    f = float(Fraction(d))
    if abs(Fraction(f)) > abs(Fraction(d)):
        f = math.nextafter(f, 0.0)
    return f


def fmt(values, per_line=3):
    out = []
    for i in range(0, len(values), per_line):
        out.append("    " + ", ".join(float.hex(v) for v in values[i:i + per_line]) + ",")
    return "\n".join(out)


def d_pi() -> Decimal:
    # Machin: pi = 16 atan(1/5) - 4 atan(1/239).
    return 16 * d_atan(Decimal(1) / 5) - 4 * d_atan(Decimal(1) / 239)


def two_over_pi_windows():
    frac = Fraction(2) / Fraction(d_pi())  # 2 / pi to ~260 bits
    return [int(frac * 2 ** (8 * k + 8)) % 2 ** 32 for k in range(24)]


def tables():
    atan_tab = [to_double_toward_zero(d_atan(Decimal(k + 16) / 256)) for k in range(241)]
    exp2_tab = [to_double((Decimal(k) / 64 * Decimal(2).ln()).exp()) for k in range(64)]
    inv_c = [to_double(Decimal(512) / (256 + k)) for k in range(257)]
    log_c = [to_double((Decimal(256 + k) / 256).ln()) for k in range(257)]
    return atan_tab, exp2_tab, inv_c, log_c


def main():
    atan_tab, exp2_tab, inv_c, log_c = tables()
    print("// Generated by gen_math_tables.py: doubles of atan((k+16)/256) (rounded toward zero), and correctly rounded")
    print("// 2^(k/64), 512/(256+k) and ln((256+k)/256).  Do not edit.")
    print(f"RG_MATH_TABLE(double, kAtanTab, 241) = {{\n{fmt(atan_tab)}\n}};")
    print(f"RG_MATH_TABLE(double, kExp2Tab, 64) = {{\n{fmt(exp2_tab)}\n}};")
    print(f"RG_MATH_TABLE(double, kInvCTab, 257) = {{\n{fmt(inv_c)}\n}};")
    print(f"RG_MATH_TABLE(double, kLogCTab, 257) = {{\n{fmt(log_c)}\n}};")
    w = two_over_pi_windows()
    rows = ["    " + ", ".join(f"0x{v:08x}u" for v in w[i:i + 6]) + "," for i in range(0, 24, 6)]
    print("RG_MATH_TABLE(uint32_t, kTwoOverPiBits, 24) = {\n" + "\n".join(rows) + "\n};")


if __name__ == "__main__":
    main()
