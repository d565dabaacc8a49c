"""GPU parity of the observables around the force pass, through the C ABI,
against the oracle on the same inputs.

* Per-point tallies of Pair::compute (eatom, enode, vatom, vnode; pair.h:13-17,
  pair.cpp:441-476, pair_eam.cpp:863-918, 979-984) for ENERGY_ATOM,
  VIRIAL_ATOM and VIRIAL_CENTROID: FP64 within 1e-12 of the oracle (max |a-b|
  over max |b|); mixed precision within 1e-5 of FP64 for energies and 5e-5
  for virials.
* Device temperature and pressure (cac_hip_thermo and the thermo rows of the
  resident loop; compute_temp.cpp:73-157, compute_pressure.cpp:140-246)
  against the oracle's within 1e-12 (setup) and 1e-10 (after 20 steps).
* Elements of 11^3 sites (the large system's element type): sets exact,
  FP64 forces 1e-12, a 20-step trajectory 1e-10, mixed forces 1e-5.
"""

import numpy as np
import pytest

from cac_amd import decks
from tests.test_gpu_parity import cac_lj_deck, check_forces, gpu_model, oracle_model, oracle_sets, rel, setup_pair

pytestmark = pytest.mark.gpu

TAB = decks.cu_eam_tables()

PERATOM_DECKS = {
    "cac_eam_n3": (lambda: decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.04, T=300.0, seed=21),
                   "eam", 1.0),
    "cac_eam_n3_no_atoms": (lambda: decks.cac_eam(box_cells=(3, 3, 3), core_cells=0, ncell=3, jitter=0.04, T=300.0,
                                                  seed=22), "eam", 1.0),
    "cac_eam_n9": (lambda: decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=9, jitter=0.03, T=300.0, seed=23),
                   "eam", 2.0),
    "lj_atoms": (lambda: decks.lj_block(n=6, jitter=0.05, T=0.1, seed=24), "lj", 0.3),
    "cac_lj": (lambda: cac_lj_deck(seed=25), "lj", 0.3),
}


def _oracle_peratom(o):
    ea, en = o.peratom_energy()
    va, vn = o.peratom_virial()
    return dict(eatom=ea, enode=en, vatom=va, vnode=vn)


def _peratom_rel(a, b, key):
    if b[key].size == 0:
        return 0.0
    return rel(a[key], b[key])


@pytest.mark.parametrize("name", sorted(PERATOM_DECKS))
def test_peratom_tallies_match_oracle(name):
    import cac_amd
    make, pair, skin = PERATOM_DECKS[name]
    deck = make()
    o = oracle_model(deck, pair, skin)
    o.setup(eflag=3, vflag=5)
    g = gpu_model(deck, pair, skin)
    g.setup(eflag=3, vflag=5)
    check_forces(o, g, deck)
    po, pg = _oracle_peratom(o), g.peratom()
    for key in ("eatom", "enode", "vatom", "vnode"):
        assert _peratom_rel(pg, po, key) < 1e-12, (key, _peratom_rel(pg, po, key))
    # the per-point tallies weighted by nodal_weight sum to the global ones
    nw = float(np.prod(deck.ncell)) / 8.0 if deck.nelements else 0.0
    e, vir = g.energy()
    assert pg["eatom"].sum() + nw * pg["enode"].sum() == pytest.approx(e, rel=1e-11)
    assert np.allclose(pg["vatom"].sum(0) + nw * pg["vnode"].sum((0, 1)), vir, rtol=1e-10,
                       atol=1e-10 * np.abs(vir).max())
    # mixed precision, with the same lists: per-point energies within 1e-5 of
    # FP64; per-point virials within 5e-5, since a virial near equilibrium is
    # a small sum of large pair terms of both signs, each carrying the FP32
    # pair term's rounding (1e-7 relative)
    gm = gpu_model(deck, pair, skin, precision=cac_amd.MIXED)
    gm.setup(eflag=3, vflag=5)
    pm = gm.peratom()
    for key, bar in (("eatom", 1e-5), ("enode", 1e-5), ("vatom", 5e-5), ("vnode", 5e-5)):
        assert _peratom_rel(pm, pg, key) < bar, (key, _peratom_rel(pm, pg, key))


def test_peratom_flags():
    """VIRIAL_CENTROID gives the per-atom two-body virial (pair.cpp:280-281);
    ENERGY_ATOM alone leaves eng_vdwl as it was (ev_setup zeroes and tallies
    it only under ENERGY_GLOBAL); asking for tallies the last compute did not
    form is CAC_HIP_ESTATE; flags outside force.h's bits are CAC_HIP_EARG."""
    import cac_amd
    make, pair, skin = PERATOM_DECKS["cac_eam_n3"]
    deck = make()
    g = gpu_model(deck, pair, skin)
    g.setup(eflag=1, vflag=1)
    e1, v1 = g.energy()
    g.compute(eflag=2, vflag=4)
    e2, v2 = g.energy()
    assert e2 == e1 and np.array_equal(v2, v1)
    va = g.peratom(energy=False)["vatom"]
    g.compute(eflag=0, vflag=8)
    vc = g.peratom(energy=False)["vatom"]
    assert np.array_equal(va, vc)
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.peratom(virial=False)  # the last compute asked for no energies
    assert ei.value.code == 7
    g.compute(eflag=0, vflag=0)
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.peratom()
    assert ei.value.code == 7
    for ef, vf in ((4, 0), (0, 16), (-1, 0)):
        with pytest.raises(cac_amd.CacHipError) as ei:
            g.compute(eflag=ef, vflag=vf)
        assert ei.value.code == 1


# ------------------------------------------------------------------ temperature and pressure

THERMO_DECKS = {
    "cac_eam_n3": (lambda: decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=26),
                   "eam", 1.0, 0.001),
    "cac_eam_n9": (lambda: decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=9, jitter=0.03, T=300.0, seed=27),
                   "eam", 2.0, 0.001),
    "lj_atoms": (lambda: decks.lj_block(n=6, jitter=0.02, T=0.5, seed=28), "lj", 0.3, 0.005),
}


@pytest.mark.parametrize("name", sorted(THERMO_DECKS))
def test_device_temperature_and_pressure(name):
    """cac_hip_thermo's PE, KE, temperature (consistent-mass KE of the
    elements, dof = 3 (atoms + virtual atoms)) and pressure (dof k T plus the
    pair virial, over 3V) equal the oracle's compute_temp / compute_pressure
    after setup; then every thermo column of 20 resident steps (pressure
    included) equals the oracle's run."""
    make, pair, skin, dt = THERMO_DECKS[name]
    deck = make()
    o, g = setup_pair(deck, pair, skin)
    to, tg = o.thermo(), g.thermo()
    for key in ("pe", "ke", "temp", "press"):
        assert tg[key] == pytest.approx(to[key], rel=1e-12, abs=1e-12), (key, tg[key], to[key])
    assert to["temp"] > 0 and abs(to["press"]) > 0
    th_o = o.run(20, dt, thermo_every=10)
    th_g = g.run(20, dt, thermo_every=10)
    assert th_g.shape == th_o.shape
    assert np.array_equal(th_g[:, 0], th_o[:, 0])
    for col in (1, 2, 3, 4):
        assert np.allclose(th_g[:, col], th_o[:, col], rtol=1e-10, atol=1e-10 * np.abs(th_o[:, col]).max()), col
    to, tg = o.thermo(), g.thermo()
    assert tg["temp"] == pytest.approx(to["temp"], rel=1e-10)
    assert tg["press"] == pytest.approx(to["press"], rel=1e-10, abs=1e-10 * abs(to["press"]))


# ------------------------------------------------------------------ 11^3 elements

def test_ncell11_sets_forces_trajectory():
    """The large system's element type (Hex of 11^3 sites, 27 sub-elements in
    the reference's pruning) at a box the oracle finishes: sorted per-point
    sets equal, FP64 forces and energy within 1e-12, 20 NVE steps within
    1e-10, mixed forces within 1e-5 of FP64."""
    import cac_amd
    deck = decks.cac_eam(box_cells=(2, 2, 2), core_cells=2, ncell=11, jitter=0.03, T=300.0, seed=29)
    assert deck.nelements > 0 and deck.natoms > 0
    o, g = setup_pair(deck, "eam", 2.0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)
    ar_o, nr_o = o.rho()
    ar_g, nr_g = g.rho()
    assert rel(ar_g, ar_o) < 1e-13 and rel(nr_g, nr_o) < 1e-13
    th_o = o.run(20, 0.001, thermo_every=10)
    th_g = g.run(20, 0.001, thermo_every=10)
    assert np.allclose(th_g[:, 1:5], th_o[:, 1:5], rtol=1e-10, atol=1e-9)
    assert rel(g.elements()["nodex"], o.elements()["nodex"]) < 1e-12
    gm = gpu_model(deck, "eam", 2.0, precision=cac_amd.MIXED)
    gm.setup()
    g64 = gpu_model(deck, "eam", 2.0)
    g64.setup()
    f64 = np.concatenate([g64.atoms()[2], g64.elements()["gaussf"].reshape(-1, 3)])
    f32 = np.concatenate([gm.atoms()[2], gm.elements()["gaussf"].reshape(-1, 3)])
    assert rel(f32, f64) < 1e-5
