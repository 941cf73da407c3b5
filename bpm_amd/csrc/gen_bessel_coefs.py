"""Print the float64 coefficient tables of bessel.hpp.

Series (A&S 9.1.10): J_n(x) = (x/2)^n sum_k c_k (-(x/2)^2)^k with
c_k = 1/(k! (k+n)!), k < 32.
Hankel asymptotic (A&S 9.2.5, 9.2.9, 9.2.10), with y = 1/x:
P = sum_m (-1)^m a_{2m} y^{2m},  Q = y sum_m (-1)^m a_{2m+1} y^{2m},  m < 12,
a_k = prod_{j=1..k} (4n^2 - (2j-1)^2) / (k! 8^k).
Each coefficient is an exact rational rounded once to float64.
Usage: python gen_bessel_coefs.py  (prints the C++ arrays)
"""
from fractions import Fraction
from math import factorial

KS, KA = 32, 12


def series(n):
    return [float(Fraction(1, factorial(k) * factorial(k + n))) for k in range(KS)]


def asymptotic(n):
    mu = 4 * n * n
    a = [Fraction(1)]
    for k in range(1, 2 * KA + 2):
        odd = 2 * k - 1
        a.append(a[-1] * Fraction(mu - odd * odd, 8 * k))
    P = [float((-1) ** m * a[2 * m]) for m in range(KA)]
    Q = [float((-1) ** m * a[2 * m + 1]) for m in range(KA)]
    return P, Q


def array(name, values):
    rows = [", ".join(repr(x) for x in values[i:i + 3]) for i in range(0, len(values), 3)]
    return "  constexpr double %s[%d] = {\n      %s};\n" % (name, len(values), ",\n      ".join(rows))


def main():
    out = []
    for n in range(3):
        out.append(array("S%d" % n, series(n)))
    for n in range(3):
        P, Q = asymptotic(n)
        out.append(array("P%d" % n, P))
        out.append(array("Q%d" % n, Q))
    print("".join(out), end="")


if __name__ == "__main__":
    main()
