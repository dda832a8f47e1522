"""Computes the constants of mdy_xmath.h from the mathematics (100-digit
decimal arithmetic): atan(k/8) for k = 0..8 and pi/2 as double-double pairs,
and ln 2 split for Cody-Waite reduction. Prints C initializers."""
from decimal import Decimal as D, getcontext

getcontext().prec = 110


def atan_series(x):
    # atan(x) = sum (-1)^n x^(2n+1)/(2n+1), after halving the argument three
    # times with atan(x) = 2 atan(x / (1 + sqrt(1 + x^2)))
    halvings = 0
    while abs(x) > D("0.05"):
        x = x / (1 + (1 + x * x).sqrt())
        halvings += 1
    s, term, n = D(0), x, 0
    x2 = x * x
    while abs(term) > D(10) ** -105:
        s += term / (2 * n + 1) if n % 2 == 0 else -term / (2 * n + 1)
        term *= x2
        n += 1
    return s * (2 ** halvings)


def split(v):
    hi = float(v)
    lo = float(v - D(hi))
    return hi, lo


pi = 4 * atan_series(D(1))
print("/* atan(k/8), k = 0..8: high and low parts */")
his, los = [], []
for k in range(9):
    hi, lo = split(atan_series(D(k) / 8))
    his.append(hi.hex()); los.append(lo.hex())
print("hi:", ", ".join(his))
print("lo:", ", ".join(los))
hi, lo = split(pi / 2)
print("pi/2:", hi.hex(), lo.hex())
ln2 = D(2).ln()
# ln2_hi with 32 significant bits so k*ln2_hi is exact for |k| < 2^21
import math
m, e = math.frexp(float(ln2))
hi32 = math.ldexp(math.floor(m * 2 ** 32), e - 32)
lo = float(ln2 - D(hi32))
print("ln2_hi:", hi32.hex(), "ln2_lo:", lo.hex(), "1/ln2:", float(1 / ln2).hex())
