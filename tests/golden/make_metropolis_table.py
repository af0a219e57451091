"""Writes metropolis_table.json: the Metropolis acceptance table the strict
sweep uses, as exact double bits, for the temperatures the tests and the
single-lattice workloads run at.

tab[d] = exp((-d) * (1.0/kT)) for d = 1..26 is the value the reference
compares a uniform draw against (src/app_potts.cpp:130,
src/app_ising.cpp:120, with t_inverse = 1.0/temperature from
src/app_lattice.cpp:893-898).  The values come from this machine's libm
through math.exp (glibc, the same exp the reference calls when built here).
Each is also checked against a 60-digit decimal evaluation of exp: every
entry is within 0.51 ulp of the exact value, and the entries where glibc is
not correctly rounded are listed, so a build that swaps in another exp (a
device exp, or a correctly rounded one) shows up against this file.

thr[d] is the integer form of the same comparison for Park-Miller states
s in [1, 2^31-2]: a move is rejected when (1.0/2147483647)*s > tab[d], i.e.
when s >= thr[d].  thr[0] splits the Ising proposal: u < 0.5 iff s < thr[0].
Both are found here by bisection over s on the double products, not from
the library's own formula.

Run: python tests/golden/make_metropolis_table.py
"""
import decimal
import json
import math
import os
import platform

IM = 2147483647
AM = 1.0 / IM
MAXD = 26
TEMPERATURES = [0.5, 2.269, 0.1, 0.9, 1.0, 3.5, 1e-3, 100.0]


def first_above(pred):
    """Smallest s in [1, IM] with pred(s) true (pred monotone); IM if none."""
    lo, hi = 1, IM
    while lo < hi:
        mid = (lo + hi) // 2
        if pred(mid):
            hi = mid
        else:
            lo = mid + 1
    return lo


def ulp_error(value, x):
    decimal.getcontext().prec = 60
    exact = decimal.Decimal(x).exp()
    return float(abs(decimal.Decimal(value) - exact) / decimal.Decimal(math.ulp(value)))


def table(kt):
    tab, thr, errs = [0.0], [first_above(lambda s: AM * s >= 0.5)], [0.0]
    for d in range(1, MAXD + 1):
        x = (-d) * (1.0 / kt)
        e = math.exp(x)
        tab.append(e)
        errs.append(ulp_error(e, x) if e > 0.0 else 0.0)
        thr.append(first_above(lambda s, e=e: AM * s > e))
    return tab, thr, errs


def main():
    out = {"libm": "glibc " + platform.libc_ver()[1], "maxd": MAXD, "tables": {}}
    for kt in TEMPERATURES:
        tab, thr, errs = table(kt)
        assert max(errs) < 0.51, (kt, max(errs))
        out["tables"][repr(kt)] = {
            "exp_bits": [e.hex() for e in tab],
            "thr": thr,
            "not_correctly_rounded": [d for d in range(1, MAXD + 1) if errs[d] > 0.5],
        }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "metropolis_table.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
