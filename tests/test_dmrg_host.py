"""Host-side pieces of the sweep driver: two-site merge / SVD split round trip, the reference
truncation rule, canonical forms and the allowed bond sectors."""
import numpy as np

from qcmaquis_amd.dmrg import (allowed_bonds, estimate_truncation, merge_two, right_normalize, rp_block, split_two,
                               twosite_offsets)
from qcmaquis_amd.model import MPSLayout
from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index


def chain(L=5, total=(2, 3), m=4, seed=0):
    rng = np.random.default_rng(seed)
    bonds = allowed_bonds([TWOU1_SITE] * L, total, m)
    lay = [MPSLayout(bonds[k], TWOU1_SITE, bonds[k + 1]) for k in range(L)]
    return lay, [rng.normal(size=l.n) for l in lay]


def test_allowed_bonds_counts():
    b = allowed_bonds([TWOU1_SITE] * 4, (2, 2), 10 ** 9)
    assert b[0].charges == [(0, 0)] and b[4].charges == [(2, 2)]
    # bond 2 of 4 sites at half filling: sector (1,1) has C(2,1)^2 = 4 states from both sides
    assert b[2].size_of((1, 1)) == 4 and b[2].size_of((2, 2)) == 1


def test_merge_split_roundtrip():
    lay, data = chain()
    phys2 = fuse_index(TWOU1_SITE, TWOU1_SITE)
    off = twosite_offsets(TWOU1_SITE)
    lay2, psi2 = merge_two(lay[1], data[1], lay[2], data[2], phys2, off)
    for to_right in (True, False):
        la, a, lb, b, tw = split_two(lay2, psi2, TWOU1_SITE, off, 10 ** 6, 0.0, to_right)
        assert tw < 1e-28
        _, back = merge_two(la, a, lb, b, phys2, off)
        assert np.allclose(back, psi2, atol=1e-12)
        # the normalised factor is an isometry
        x = (la, a) if to_right else (lb, b)
        if to_right:
            lp = x[0].to_left(x[1])
            for rb in range(len(x[0].right)):
                blk = lp[x[0].lp_off[rb]:x[0].lp_off[rb + 1]].reshape(x[0].right.sizes[rb], x[0].lp_rows[rb]).T
                assert np.allclose(blk.T @ blk, np.eye(blk.shape[1]), atol=1e-12)
        else:
            for lb_ in range(len(x[0].left)):
                blk = rp_block(x[0], x[1], lb_)
                assert np.allclose(blk @ blk.T, np.eye(blk.shape[0]), atol=1e-12)


def test_truncation_rule():
    s = {"a": np.array([5.0, 3.0, 1.0]), "b": np.array([4.0, 2.0, 0.5])}
    keeps, tw = estimate_truncation(s, 3, 1e-16)
    # values >= s_[3] = 2.0 are kept (the reference keeps the Mmax+1-th value as well)
    assert keeps == {"a": 2, "b": 2}
    assert np.isclose(tw, (1.0 + 0.25) / (25 + 9 + 1 + 16 + 4 + 0.25))
    keeps, _ = estimate_truncation(s, 10, 0.3)
    assert keeps == {"a": 2, "b": 2}  # relative cutoff 0.3 * 5 = 1.5


def test_right_normalize():
    lay, data = chain(L=6, total=(3, 3), m=5, seed=2)
    phys2 = fuse_index(TWOU1_SITE, TWOU1_SITE)
    off = twosite_offsets(TWOU1_SITE)
    # the represented state is unchanged by canonisation: compare <psi|psi> of the first two sites
    right_normalize(lay, data)
    for k in range(1, 6):
        for lb in range(len(lay[k].left)):
            blk = rp_block(lay[k], data[k], lb)
            assert np.allclose(blk @ blk.T, np.eye(blk.shape[0]), atol=1e-12)
    assert np.isclose(np.linalg.norm(data[0]), 1.0)
