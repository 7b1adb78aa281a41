"""SU2U1 two-site tensors: the recoupling between the sequential and the pair-coupled two-site basis
(reference: mp_tensors/ts_reduction.h:36-260, used by TwoSiteTensor::make_mps and operator<<,
twositetensor.hpp:144-151, 247-304).

A two-site tensor formed from two site tensors (TwoSiteTensor ctor) has its product physical
states (s1, s2) coupled in sequence: left spin jl, then jm = jl + s1, then jr = jm + s2.  The
two-site MPO acts on the pair-coupled basis, where (s1, s2) first couple to the pair spin j and
then j couples with jl to jr.  The change of basis is the 6j recoupling

    <(jl, S1) jm, S2; jr | jl, (S1, S2) j; jr> = (-1)^{(jl+jr+S1+S2)/2} sqrt((j+1)(jm+1)) {jl jr j; S2 S1 jm}

(all spins as 2S).  When both sites are singly occupied and jl == jr > 0 the two paths jm = jl +- 1
mix into the pair singlet (stored at the position of the state (1,1)x(1,-1)) and the pair
triplet (at (1,-1)x(1,1)); otherwise every state maps onto itself with a scalar factor.  The map
is orthogonal, so the inverse (unreduce) is its transpose.
"""
import numpy as np

from . import abi
from .model import qadd


def _sixj(a, b, c, d, e, f):
    return abi.lib().mq_wigner_6j(int(a), int(b), int(c), int(d), int(e), int(f))


class TwoSiteReduction:
    """The recoupling of one two-site layout (MPSLayout(left, phys x phys, right), right-paired):
    a list of (left block, input column, output column, column count, coefficient)."""

    def __init__(self, lay2, phys, off12):
        self.lay2 = lay2
        moves = []
        cache = {}
        for lb, lc in enumerate(lay2.left.charges):
            jl = lc[1]
            for s1 in phys.charges:
                for s2 in phys.charges:
                    c, o12 = off12[(s1, s2)]
                    r = qadd(lc, c)
                    rb = lay2.right.position(r)
                    if rb >= len(lay2.right) or (c, r) not in lay2.rpb:
                        continue
                    jr, S1, S2, jm = r[1], abs(s1[1]), abs(s2[1]), jl + s1[1]
                    if jm < 0:
                        continue
                    rs = lay2.right.sizes[rb]
                    base_col = lay2.rpb[(c, r)]
                    sign = -1.0 if ((jl + jr + S1 + S2) // 2) % 2 else 1.0
                    if jl == jr and jl > 0 and S1 == 1 and S2 == 1:
                        base = o12 if s1[1] == 1 else o12 - 1
                        targets = [(base + j // 2, j) for j in (0, 2)]
                    else:
                        targets = [(o12, abs(s1[1] + s2[1]))]
                    for out, j in targets:
                        key = (jl, jr, j, S2, S1, jm)
                        if key not in cache:
                            cache[key] = sign * np.sqrt((j + 1.0) * (jm + 1.0)) * _sixj(jl, jr, j, S2, S1, jm)
                        coef = cache[key]
                        if coef != 0.0:
                            moves.append((lb, base_col + o12 * rs, base_col + out * rs, rs, coef))
        self.moves = moves

    def _apply(self, x, transpose):
        lay = self.lay2
        y = np.zeros_like(x)
        for lb, cin, cout, rs, coef in self.moves:
            ls = lay.left.sizes[lb]
            src = x[lay.rp_off[lb]:lay.rp_off[lb + 1]].reshape(lay.rp_cols[lb], ls)
            dst = y[lay.rp_off[lb]:lay.rp_off[lb + 1]].reshape(lay.rp_cols[lb], ls)
            if transpose:
                cin, cout = cout, cin
            dst[cout:cout + rs] += coef * src[cin:cin + rs]
        return y

    def reduce(self, x):
        """sequentially coupled -> pair coupled (ts_reduction::reduce_right)"""
        return self._apply(x, False)

    def unreduce(self, y):
        """pair coupled -> sequentially coupled (ts_reduction::unreduce_left / unreduce_right)"""
        return self._apply(y, True)
