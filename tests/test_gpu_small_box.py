"""GPU parity on periodic boxes shorter than 2.5 (rc + skin) along some axis,
where the neighbor cell grid holds two to four cells and a site is met as
more than one image: the reference's ghost shell then holds several images of
an atom or element (comm_brick.cpp:163-515, maxneed > 1).  Sorted per-point
sets equal the oracle's, FP64 forces, energy and virial within 1e-12, 20 NVE
steps within 1e-10 (rebuilds included), mixed forces within 1e-5 of FP64."""

import numpy as np
import pytest

from cac_amd import decks
from tests.test_gpu_parity import check_forces, gpu_model, oracle_sets, rel, setup_pair

pytestmark = pytest.mark.gpu

DECKS = {
    # 32 atoms, L = 3.17 sigma against rc + skin = 2.8 sigma: two cells per axis,
    # about 84 neighbors per atom, most of them images of each other
    "lj_two_cells": (lambda: decks.lj_block(n=2, jitter=0.05, T=0.2, seed=41), "lj", 0.3, 0.005),
    # three cells along x (L = 10.8 A), nine along y and z
    "cac_three_cells_x": (lambda: decks.cac_eam(box_cells=(1, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0,
                                                seed=42), "eam", 2.0, 0.001),
    # four cells along x and y (L = 14.5 A), elements of 4^3 sites
    "cac_four_cells_xy": (lambda: decks.cac_eam(box_cells=(1, 1, 3), core_cells=2, ncell=4, jitter=0.03, T=300.0,
                                                seed=43), "eam", 2.0, 0.001),
}


@pytest.mark.parametrize("name", sorted(DECKS))
def test_short_box_sets_forces_trajectory(name):
    import cac_amd
    make, pair, skin, dt = DECKS[name]
    deck = make()
    cut = (decks.cu_eam_tables().cut if pair == "eam" else 2.5) + skin
    L = deck.boxhi - deck.boxlo
    assert np.any(L < 2.5 * cut) and np.all(L >= cut)
    o, g = setup_pair(deck, pair, skin, delay=0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)
    th_o = o.run(20, dt, thermo_every=10)
    th_g = g.run(20, dt, thermo_every=10)
    assert np.allclose(th_g[:, 1:4], th_o[:, 1:4], rtol=1e-10, atol=1e-10)
    assert g.neigh_stats()[0] == o.neigh_stats()[0]
    if deck.natoms:
        assert rel(g.atoms()[0], o.atoms()[0]) < 1e-12
    if deck.nelements:
        assert rel(g.elements()["nodex"], o.elements()["nodex"]) < 1e-12
    g64 = gpu_model(deck, pair, skin)
    g64.setup()
    gm = gpu_model(deck, pair, skin, precision=cac_amd.MIXED)
    gm.setup()

    def forces(m):
        parts = [m.atoms()[2]] if deck.natoms else []
        if deck.nelements:
            parts.append(m.elements()["gaussf"].reshape(-1, 3))
        return np.concatenate(parts)

    assert rel(forces(gm), forces(g64)) < 1e-5

