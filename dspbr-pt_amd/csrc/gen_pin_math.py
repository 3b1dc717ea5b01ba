"""Derives every constant of pin_math.hpp from exact arithmetic.

pi comes from Machin's formula (pi/4 = 4 atan(1/5) - atan(1/239)) in Python
integers scaled by 2^PREC; atan(k/8) and ln 2 from their Taylor / atanh series
in the same fixed point; the Taylor coefficients 1/n! and 1/n are exact
fractions. Every double is the correctly rounded value (Fraction -> float).

Run `python gen_pin_math.py` to print the constant block; the test suite
checks that pin_math.hpp holds exactly this block.
"""
from fractions import Fraction
from math import factorial, isqrt

PREC = 640  # bits of the fixed-point evaluations


def _atan_inv(n, prec):
    """atan(1/n) * 2^prec (floor of the series, |error| < few units)."""
    one = 1 << prec
    x = one // n
    total, k, sign, n2 = x, 1, -1, n * n
    while x:
        x //= n2
        k += 2
        total += sign * (x // k)
        sign = -sign
    return total


def pi_fixed(prec=PREC):
    g = 32  # guard bits
    v = 4 * (4 * _atan_inv(5, prec + g) - _atan_inv(239, prec + g))
    return v >> g


def atan_frac_fixed(p, q, prec=PREC):
    """atan(p/q) * 2^prec for 0 <= p/q <= 1 by the Euler series
    atan(x) = sum_k 2^(2k) (k!)^2 / (2k+1)! * x^(2k+1) / (1+x^2)^(k+1)."""
    g = 32
    P = prec + g
    one = 1 << P
    y = Fraction(p * p, p * p + q * q)  # x^2/(1+x^2)
    term = Fraction(p * q, p * p + q * q)  # x/(1+x^2)
    s = Fraction(0)
    k = 0
    while True:
        s += term
        k += 1
        term = term * y * Fraction(2 * k, 2 * k + 1)
        if term * one < 1:
            break
    return int(s * one) >> g


def ln2_fixed(prec=PREC):
    """ln 2 = 2 atanh(1/3)."""
    g = 32
    P = prec + g
    one = 1 << P
    x = one // 3
    s, k = x, 1
    while x:
        x //= 9
        k += 2
        s += x // k
    return (2 * s) >> g


def rd(fr):
    return float(fr)


def hexd(x):
    return float.hex(x)


def split_bits(value_fixed, prec, bits):
    """Leading `bits` significant bits of value (value_fixed / 2^prec), truncated."""
    top = value_fixed.bit_length()
    keep = value_fixed >> (top - bits) << (top - bits)
    return Fraction(keep, 1 << prec)


def block():
    pi = pi_fixed()
    PI = Fraction(pi, 1 << PREC)
    pio2 = PI / 2
    pio2_fixed = pi >> 1  # pi/2 * 2^PREC (exact shift of the same fixed point)
    P1 = split_bits(pio2_fixed, PREC, 28)
    rest = pio2 - P1
    P2 = split_bits(int(rest * (1 << PREC)), PREC, 28)
    P3 = rest - P2
    two_over_pi = Fraction(2) / PI
    words = []
    frac = two_over_pi
    for _ in range(8):
        frac *= 1 << 32
        w = int(frac)
        words.append(w)
        frac -= w
    ln2 = Fraction(ln2_fixed(), 1 << PREC)
    LN2_HI = split_bits(ln2_fixed(), PREC, 42)
    LN2_LO = ln2 - LN2_HI
    sqrt2 = Fraction(isqrt(2 << (2 * PREC)), 1 << PREC)
    atans = [Fraction(atan_frac_fixed(k, 8), 1 << PREC) for k in range(9)]

    out = []
    w = out.append
    w("// ---- BEGIN GENERATED CONSTANTS (gen_pin_math.py) ----")
    w(f"constexpr double PM_PI = {hexd(rd(PI))};")
    w(f"constexpr double PM_PIO2 = {hexd(rd(pio2))};")
    w(f"constexpr double PM_2_PI = {hexd(rd(two_over_pi))};")
    w(f"constexpr double PM_PIO2_1 = {hexd(rd(P1))};  // 28 leading bits of pi/2")
    w(f"constexpr double PM_PIO2_2 = {hexd(rd(P2))};  // next 28 bits")
    w(f"constexpr double PM_PIO2_3 = {hexd(rd(P3))};  // the rest, rounded")
    w(f"constexpr double PM_LN2 = {hexd(rd(ln2))};")
    w(f"constexpr double PM_LN2_HI = {hexd(rd(LN2_HI))};  // 42 leading bits of ln 2")
    w(f"constexpr double PM_LN2_LO = {hexd(rd(LN2_LO))};")
    w(f"constexpr double PM_INV_LN2 = {hexd(rd(1 / ln2))};")
    w(f"constexpr double PM_SQRT2 = {hexd(rd(sqrt2))};  // split point of the log reduction")
    w("// fraction bits of 2/pi, 32 per word, most significant first")
    for i, x in enumerate(words):
        w(f"constexpr uint32_t PM_TOP{i} = 0x{x:08x}u;")
    w("// atan(k/8)")
    for k, a in enumerate(atans):
        w(f"constexpr double PM_ATAN8_{k} = {hexd(rd(a))};")
    # Taylor coefficients (exact fractions, rounded once)
    w("// sin r = r + r^3 (S3 + r^2 (S5 + ...)), S_n = (-1)^((n-1)/2) / n!")
    for n in range(3, 16, 2):
        w(f"constexpr double PM_S{n} = {hexd(rd(Fraction((-1) ** ((n - 1) // 2), factorial(n))))};")
    w("// cos r = 1 + r^2 (C2 + r^2 (C4 + ...)), C_n = (-1)^(n/2) / n!")
    for n in range(2, 17, 2):
        w(f"constexpr double PM_C{n} = {hexd(rd(Fraction((-1) ** (n // 2), factorial(n))))};")
    w("// exp r = 1 + r + r^2 (E2 + r (E3 + ...)), E_n = 1 / n!")
    for n in range(2, 13):
        w(f"constexpr double PM_E{n} = {hexd(rd(Fraction(1, factorial(n))))};")
    w("// atanh s = s + s^3 (L3 + s^2 (L5 + ...)), L_n = 1 / n")
    for n in range(3, 18, 2):
        w(f"constexpr double PM_L{n} = {hexd(rd(Fraction(1, n)))};")
    w("// atan u = u + u^3 (A3 + u^2 (A5 + ...)), A_n = (-1)^((n-1)/2) / n")
    for n in range(3, 14, 2):
        w(f"constexpr double PM_A{n} = {hexd(rd(Fraction((-1) ** ((n - 1) // 2), n)))};")
    w("// ---- END GENERATED CONSTANTS ----")
    return "\n".join(out)


if __name__ == "__main__":
    print(block())
