"""Golden Wigner 6j / 9j values from sympy's exact rational arithmetic (sympy.physics.wigner), for
every argument set the SU(2) couplings of the contraction can request: bond / operator spins
2j in {0,1,2}, sector spins 2j <= 8.  Writes wigner_golden.json next to this script.

The 9j arguments follow SU2::set_coupling / mod_coupling (gsl_coupling.h:38-82):
  {a b c; d e f; g h i} with a, c, g, i sector spins, b, h site spins (0..2), d, e, f in {0,1,2}.
"""
import itertools
import json
import os

from sympy import Rational
from sympy.physics.wigner import wigner_6j, wigner_9j


def tri(a, b, c):
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def main():
    out = {"9j": [], "6j": []}
    seen = set()
    for a, c in itertools.product(range(0, 9), repeat=2):
        for b in (0, 1, 2):
            if not tri(a, b, c):
                continue
            for d, e, f in itertools.product((0, 1, 2), repeat=3):
                for g, i in itertools.product(range(max(0, a - 2), a + 3), range(max(0, c - 2), c + 3)):
                    for h in (0, 1, 2):
                        key = (a, b, c, d, e, f, g, h, i)
                        if key in seen:
                            continue
                        seen.add(key)
                        ok = tri(a, b, c) and tri(d, e, f) and tri(g, h, i) and tri(a, d, g) and tri(b, e, h) and tri(c, f, i)
                        if not ok:
                            continue
                        v = wigner_9j(*[Rational(x, 2) for x in key])
                        out["9j"].append(list(key) + [float(v)])
    for key in itertools.product(range(0, 7), repeat=6):
        a, b, c, d, e, f = key
        if not (tri(a, b, c) and tri(a, e, f) and tri(d, b, f) and tri(d, e, c)):
            continue
        if len(out["6j"]) > 4000:
            break
        v = wigner_6j(*[Rational(x, 2) for x in key])
        out["6j"].append(list(key) + [float(v)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wigner_golden.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print(len(out["9j"]), "9j values,", len(out["6j"]), "6j values ->", path)


if __name__ == "__main__":
    main()
