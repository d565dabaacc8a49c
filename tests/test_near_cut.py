"""Neighbor sets at a cutoff that falls between two roundings of one pair
distance (tests/near_cut.py): the oracle, which forms distances as the
reference does (npair_half_bin_newton.cpp:194-199, 359-364), keeps or drops
the pair as that form says (CPU); the device's lists equal the oracle's
exactly, so the device decides such a pair in the reference's form too (GPU).
"""

import numpy as np
import pytest

from oracle import oracle as O
from tests import brute, near_cut


@pytest.fixture(scope="module")
def case():
    return near_cut.split_case()


def _has_pair(keys, info):
    want = np.array([info["point"], 1, info["element"], info["ucell"], *info["image"]])
    return bool(np.any(np.all(keys == want, axis=1)))


def test_oracle_decides_the_split_pair_in_the_reference_form(case):
    deck, skin, info = case
    assert info["fast"] != info["ref"]
    assert min(info["fast"], info["ref"]) <= info["cutsq"] < max(info["fast"], info["ref"])
    o = O.Oracle(deck, "eam", skin=skin, eam_tables=near_cut.TAB)
    o.build_neighbor()
    p, k, ow, u, img = o.half_list()
    keys = O.canonical_full_sets(deck, p, k, ow, u, img, brute.corner_ucells(deck.ncell))
    assert _has_pair(keys, info) == info["ref_keeps"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_device_lists_equal_oracle_at_the_split_cutoff(case, precision):
    import cac_amd
    deck, skin, info = case
    o = O.Oracle(deck, "eam", skin=skin, eam_tables=near_cut.TAB)
    o.build_neighbor()
    p, k, ow, u, img = o.half_list()
    want = O.canonical_full_sets(deck, p, k, ow, u, img, brute.corner_ucells(deck.ncell))
    g = cac_amd.Model(deck, "eam", skin=skin, eam_tables=near_cut.TAB, precision=precision)
    g.build_neighbor()
    got = g.full_sets()
    assert _has_pair(got, info) == info["ref_keeps"]
    assert np.array_equal(got, want)
