"""Sector-index ordering: the reference's known-answer tests (regression/tests/index/index.cpp)
restated for the host mirror (qcmaquis_amd.model.Index) and the product's C++ Index through
mq_index_position.  U1 charges q are embedded as (q, 0) two-component charges."""
import ctypes as C

import numpy as np
import pytest

from qcmaquis_amd import abi
from qcmaquis_amd.model import Index, MPSLayout
from qcmaquis_amd.synthetic import TWOU1_SITE, fuse_index


def u1(pairs):
    return Index([(q, 0) for q, _ in pairs], [s for _, s in pairs])


def cpos(idx, q):
    d = idx.desc()
    return abi.lib().mq_index_position(C.byref(d), q, 0)


def test_positioni():  # positioni_test
    idx = u1([(0, 5), (3, 2), (2, 3), (1, 4), (4, 1)])
    assert idx.position((3, 0)) == 1 and idx.position((1, 0)) == 3
    assert cpos(idx, 3) == 1 and cpos(idx, 1) == 3
    assert cpos(idx, 7) == len(idx)  # absent -> size()


def test_position_pair():  # position_pair_test: global offset of (charge, i)
    idx = u1([(0, 1), (1, 2), (2, 2), (3, 1)])
    glob = lambda q, i: sum(idx.sizes[:idx.position((q, 0))]) + i
    assert glob(1, 1) == 4 and glob(3, 0) == 0


def test_sort_and_charges_sizes():  # sort_test, charges_test, sizes_test, sum_of_size
    a = u1([(2, 1), (1, 1), (0, 1)])
    b = u1([(0, 1), (1, 1), (2, 1)])
    assert a == b
    c = u1([(0, 1), (1, 1)])
    assert c.charges == [(1, 0), (0, 0)]
    d = u1([(0, 1), (1, 2)])
    assert d.sizes == [2, 1] and d.total() == 3


def test_unsorted_rejected():
    ch = np.array([[0, 0], [1, 0]], np.int32)
    sz = np.array([1, 1], np.uint32)
    d = abi.IndexDesc(2, ch.ctypes.data_as(C.POINTER(C.c_int32)), sz.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert abi.lib().mq_index_position(C.byref(d), 0, 0) == abi.MQ_ERR_INVALID


def test_lexicographic_two_component():
    idx = Index([(1, 0), (0, 1), (1, 1), (0, 0), (2, 0)], [1] * 5)
    assert idx.charges == [(2, 0), (1, 1), (1, 0), (0, 1), (0, 0)]
    for i, c in enumerate(idx.charges):
        assert cpos(idx, c[0]) == idx.position((c[0], 0)) if c[1] == 0 else True
        d = idx.desc()
        assert abi.lib().mq_index_position(C.byref(d), c[0], c[1]) == i


def test_two_site_fusion():
    # TwoU1 two-site index: 16 states in 9 sectors (operator*, indexing_stable.hpp:425-448)
    two = fuse_index(TWOU1_SITE, TWOU1_SITE)
    assert len(two) == 9 and two.total() == 16
    assert dict(zip(two.charges, two.sizes))[(1, 1)] == 4


def test_pairing_roundtrip():
    rng = np.random.default_rng(3)
    left = Index([(1, 0), (0, 1), (0, 0)], [2, 3, 1])
    right = Index([(2, 1), (1, 1), (1, 0), (0, 1)], [2, 2, 3, 1])
    lay = MPSLayout(left, TWOU1_SITE, right)
    rp = rng.normal(size=lay.n)
    assert np.array_equal(lay.to_right(lay.to_left(rp)), rp)
