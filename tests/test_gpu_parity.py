"""GPU parity: the HIP path through the C ABI against the oracle on the same inputs.

Bars: neighbor lists equal as sorted per-point sets (exact); FP64 forces,
energies, virials and densities within 1e-12 relative (max |a - b| over the
array divided by max |b|); mixed precision forces within 1e-5 of FP64.
"""

import dataclasses

import numpy as np
import pytest

from cac_amd import decks
from tests import brute

pytestmark = pytest.mark.gpu

TAB = decks.cu_eam_tables()
G2U9 = None


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def oracle_model(deck, pair, skin, **kw):
    from oracle import oracle as O
    o = O.Oracle(deck, pair, skin=skin, eam_tables=TAB if pair == "eam" else None, **kw)
    return o


def gpu_model(deck, pair, skin, **kw):
    import cac_amd
    return cac_amd.Model(deck, pair, skin=skin, eam_tables=TAB if pair == "eam" else None, **kw)


def oracle_sets(o, deck):
    from oracle import oracle as O
    p, k, ow, u, img = o.half_list()
    return O.canonical_full_sets(deck, p, k, ow, u, img, brute.corner_ucells(deck.ncell))


def check_forces(o, g, deck, tol=1e-12):
    _, _, fo = o.atoms()
    _, _, fg = g.atoms()
    if deck.natoms:
        assert rel(fg, fo) < tol, rel(fg, fo)
    if deck.nelements:
        eo, eg = o.elements(), g.elements()
        assert rel(eg["gaussf"], eo["gaussf"]) < tol, rel(eg["gaussf"], eo["gaussf"])
        assert rel(eg["nodef"], eo["nodef"]) < tol, rel(eg["nodef"], eo["nodef"])
    e_o, v_o = o.energy()
    e_g, v_g = g.energy()
    assert abs(e_g - e_o) <= tol * abs(e_o), (e_g, e_o)
    assert rel(v_g, v_o) < tol, (v_g, v_o)


def setup_pair(deck, pair, skin, **kw):
    o = oracle_model(deck, pair, skin, **kw)
    o.setup()
    g = gpu_model(deck, pair, skin, **kw)
    g.setup()
    return o, g


# ------------------------------------------------------------------ single force evaluations

def test_lj_atoms_sets_and_forces():
    deck = decks.lj_block(n=6, jitter=0.05, T=0.1, seed=3)
    o, g = setup_pair(deck, "lj", 0.3, delay=0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)


def test_eam_atoms_sets_forces_density():
    a = 3.615
    rng = np.random.default_rng(5)
    x = np.mod(decks.fcc_sites(a, (6, 6, 6)) + rng.normal(0, 0.05, (864, 3)), 6 * a)
    deck = decks.Deck(units="metal", boxlo=np.zeros(3), boxhi=np.full(3, 6 * a), masses=np.array([63.546]), x=x,
                      v=np.zeros_like(x), atype=np.ones(864, dtype=np.int32))
    o, g = setup_pair(deck, "eam", 2.0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)
    assert rel(g.rho()[0], o.rho()[0]) < 1e-13


# core 0: elements only; ncell 2: elements whose sites are all nodes (the atomistic limit)
@pytest.mark.parametrize("nc,box,core,skin", [(3, (3, 3, 3), 8, 1.0), (5, (2, 2, 2), 4, 2.0), (9, (2, 2, 2), 4, 2.0),
                                              (3, (3, 3, 3), 0, 1.0), (2, (4, 4, 4), 8, 1.0), (2, (4, 4, 4), 0, 1.0)])
def test_cac_eam_sets_forces(nc, box, core, skin):
    deck = decks.cac_eam(box_cells=box, core_cells=core, ncell=nc, jitter=0.03, T=300.0, seed=nc + 1)
    o, g = setup_pair(deck, "eam", skin)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)
    ar_o, nr_o = o.rho()
    ar_g, nr_g = g.rho()
    assert (ar_o.size == 0 or rel(ar_g, ar_o) < 1e-13) and rel(nr_g, nr_o) < 1e-13


def test_dislocation_deck_sets_forces_trajectory():
    """The large-system geometry (cylindrical core around an edge dislocation
    dipole, Hex elements) at a size the oracle finishes: exact sets, FP64
    forces within 1e-12, 20 steps within 1e-10, mixed forces within 1e-5."""
    import cac_amd
    deck = decks.cac_dislocation(box_cells=(4, 4, 4), core_cells=40, ncell=3, separation=4)
    o, g = setup_pair(deck, "eam", 2.0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)
    th_o = o.run(20, 0.001, thermo_every=10)
    th_g = g.run(20, 0.001, thermo_every=10)
    assert np.allclose(th_g[:, 1:4], th_o[:, 1:4], rtol=1e-10, atol=1e-9)
    gm = gpu_model(deck, "eam", 2.0, precision=cac_amd.MIXED)
    gm.setup()
    g64 = gpu_model(deck, "eam", 2.0)
    g64.setup()
    f64 = np.concatenate([g64.atoms()[2], g64.elements()["gaussf"].reshape(-1, 3)])
    f32 = np.concatenate([gm.atoms()[2], gm.elements()["gaussf"].reshape(-1, 3)])
    assert rel(f32, f64) < 1e-5


def test_dislocation_full_size_mixed_vs_fp64():
    """The full large system (2,000,493 atoms + 40,097 elements of 11^3 sites,
    2,321,269 points): mixed forces within 1e-5 of FP64, list statistics, and
    50 mixed steps that run (lists grow as needed) and stay finite."""
    import cac_amd
    deck = decks.cac_dislocation()
    res = {}
    for prec in (cac_amd.FP64, cac_amd.MIXED):
        g = gpu_model(deck, "eam", 2.0, precision=prec)
        g.setup()
        res[prec] = np.concatenate([g.atoms()[2], g.elements()["gaussf"].reshape(-1, 3)])
        info = g.neigh_info()
        assert info["rows"] == deck.atom_equivalents
        assert 100 < info["entries"] / info["rows"] < 200, info
        if prec == cac_amd.MIXED:
            th = g.run(50, 0.001, thermo_every=25)
            assert np.all(np.isfinite(th))
            x, _, f = g.atoms()
            assert np.all(np.isfinite(x)) and np.all(np.isfinite(f))
        g.close()
    assert rel(res[cac_amd.MIXED], res[cac_amd.FP64]) < 1e-5


def test_cac_lj_sets_forces():
    deck = cac_lj_deck()
    o, g = setup_pair(deck, "lj", 0.3, delay=0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)


@pytest.mark.parametrize("style", [1, 2])
def test_nodal_force_styles(style):
    deck = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=9)
    o, g = setup_pair(deck, "eam", 1.0, nodal_force_style=style)
    check_forces(o, g, deck)


# ------------------------------------------------------------------ full-size workloads

def test_cac_eam_full_workload():
    """The 32,076-atom + 1,748-element EAM workload: sets and forces."""
    deck = decks.cac_eam()
    o, g = setup_pair(deck, "eam", 2.0)
    assert np.array_equal(g.full_sets(), oracle_sets(o, deck))
    check_forces(o, g, deck)


def test_cac_eam_x8_workload_forces():
    """The benchmark's 255,879-atom + 13,985-element workload (367,759 points,
    49.3 M full-list entries) at full size: FP64 forces on atoms, Gauss points
    and nodes, densities, energy and virial within 1e-12 of the oracle's, the
    full-list size equal to that of the oracle's half list expanded (a pair
    whose site is an atom or a Gauss point counts twice), and mixed-precision
    forces within 1e-5 of FP64.  The sets themselves: the next test."""
    import cac_amd
    deck = decks.cac_eam(box_cells=(14, 16, 16), core_cells=351, ncell=9, jitter=0.02, T=300.0, seed=2024)
    o, g = setup_pair(deck, "eam", 2.0)
    check_forces(o, g, deck)
    ar_o, nr_o = o.rho()
    ar_g, nr_g = g.rho()
    assert rel(ar_g, ar_o) < 1e-12 and rel(nr_g, nr_o) < 1e-12
    p, k, _, u, _ = o.half_list()
    mirrored = (k == 0) | ((k == 1) & np.isin(u, brute.corner_ucells(deck.ncell)))
    info = g.neigh_info()
    assert info["rows"] == deck.atom_equivalents
    assert info["entries"] == len(p) + int(mirrored.sum()), (info, len(p), int(mirrored.sum()))
    f64 = np.concatenate([g.atoms()[2], g.elements()["gaussf"].reshape(-1, 3)])
    g.close()
    o.close()
    gm = gpu_model(deck, "eam", 2.0, precision=cac_amd.MIXED)
    gm.setup()
    f32 = np.concatenate([gm.atoms()[2], gm.elements()["gaussf"].reshape(-1, 3)])
    assert rel(f32, f64) < 1e-5, rel(f32, f64)
    gm.close()


def test_cac_eam_x8_workload_sets():
    """The benchmark workload's sorted per-point sets at full size (49.3 M
    full-list entries, 367,759 points) equal the oracle's half list expanded,
    compared as sorted packed int64 keys (one per row)."""
    from oracle import oracle as O
    deck = decks.cac_eam(box_cells=(14, 16, 16), core_cells=351, ncell=9, jitter=0.02, T=300.0, seed=2024)
    o = oracle_model(deck, "eam", 2.0)
    o.build_neighbor()
    p, k, ow, u, img = o.half_list()
    o.close()
    ref = O.canonical_full_keys(deck, p, k, ow, u, img, brute.corner_ucells(deck.ncell))
    del p, k, ow, u, img
    g = gpu_model(deck, "eam", 2.0)
    g.build_neighbor()
    got = np.sort(O.pack_set_rows(deck, g.full_sets(sort=False)))
    g.close()
    assert got.size == ref.size and got.size > 4.9e7
    assert np.array_equal(got, ref)


def test_lj_full_workload_momentum():
    """32,000-atom LJ: forces equal the oracle's and sum to zero."""
    deck = decks.lj_block()
    o, g = setup_pair(deck, "lj", 0.3, delay=0)
    check_forces(o, g, deck)
    _, _, fg = g.atoms()
    assert np.abs(fg.sum(axis=0)).max() < 1e-9


# ------------------------------------------------------------------ dynamics

def test_cac_trajectory_matches_oracle():
    deck = decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=9, jitter=0.02, T=300.0, seed=4)
    o, g = setup_pair(deck, "eam", 2.0, delay=0)
    tho = o.run(20, 0.001, thermo_every=10)
    thg = g.run(20, 0.001, thermo_every=10)
    assert np.allclose(thg[:, 1:4], tho[:, 1:4], rtol=1e-10, atol=1e-9)
    xo, vo, _ = o.atoms()
    xg, vg, _ = g.atoms()
    assert np.abs(xg - xo).max() < 1e-10
    eo, eg = o.elements(), g.elements()
    assert np.abs(eg["nodex"] - eo["nodex"]).max() < 1e-10


def test_lj_rebuild_schedule_matches_oracle():
    deck = decks.lj_block(n=6, jitter=0.02, T=0.5, seed=8)
    o, g = setup_pair(deck, "lj", 0.3, delay=0)
    o.run(60, 0.005)
    g.run(60, 0.005)
    assert g.neigh_stats() == o.neigh_stats()
    xo, _, _ = o.atoms()
    xg, _, _ = g.atoms()
    assert np.abs(xg - xo).max() < 1e-9


def test_run_is_bitwise_deterministic():
    deck = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=2)
    out = []
    for _ in range(2):
        g = gpu_model(deck, "eam", 2.0)
        g.setup()
        g.run(15, 0.001)
        out.append((g.atoms()[0], g.elements()["nodex"], g.energy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


# ------------------------------------------------------------------ mixed precision

def cac_lj_deck(core=20, seed=1):
    rng = np.random.default_rng(seed)
    L, x, cells, nodes = decks.cac_geometry(1.55, 3, (4, 4, 4), core)
    nodes = nodes + rng.normal(0, 0.005, nodes.shape)
    x = np.mod(x + rng.normal(0, 0.005, x.shape), L)
    E = nodes.shape[0]
    return decks.Deck(units="lj", boxlo=np.zeros(3), boxhi=L, masses=np.array([1.0]), x=x, v=np.zeros_like(x),
                      atype=np.ones(len(x), dtype=np.int32), ncell=(3, 3, 3), nodex=nodes,
                      nodev=np.zeros_like(nodes), etype=np.ones(E, dtype=np.int32), ctype=np.ones(E, dtype=np.int32))


MIXED_DECKS = {
    "cac_eam_n9": (lambda: decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=9, jitter=0.03, T=300.0, seed=6),
                   "eam", 2.0),
    "cac_eam_elements_only": (lambda: decks.cac_eam(box_cells=(3, 3, 3), core_cells=0, ncell=3, jitter=0.03,
                                                    T=300.0, seed=8), "eam", 1.0),
    "cac_eam_n2": (lambda: decks.cac_eam(box_cells=(4, 4, 4), core_cells=8, ncell=2, jitter=0.03, T=300.0, seed=9),
                   "eam", 1.0),
    "cac_lj": (lambda: cac_lj_deck(), "lj", 0.3),
}


@pytest.mark.parametrize("name", sorted(MIXED_DECKS))
def test_mixed_precision_force_error(name):
    """Mixed-precision forces within 1e-5 (max |a - b| / max |b|) and energy
    within 1e-6 of the FP64 path on the same deck, with the same outer lists."""
    import cac_amd
    build, pair, skin = MIXED_DECKS[name]
    deck = build()
    g64 = gpu_model(deck, pair, skin)
    g64.setup()
    g32 = gpu_model(deck, pair, skin, precision=cac_amd.MIXED)
    g32.setup()
    assert np.array_equal(g32.full_sets(), g64.full_sets())
    f64 = np.concatenate([g64.atoms()[2], g64.elements()["gaussf"].reshape(-1, 3)])
    f32 = np.concatenate([g32.atoms()[2], g32.elements()["gaussf"].reshape(-1, 3)])
    assert rel(f32, f64) < 1e-5, rel(f32, f64)
    e64, _ = g64.energy()
    e32, _ = g32.energy()
    assert abs(e32 - e64) < 1e-6 * abs(e64), (e32, e64)


def test_mixed_nve_drift_within_twice_fp64():
    """NVE on a 4,000-atom EAM block at 300 K, 2,000 steps of 1 fs: the mixed
    path's total-energy excursion and drift stay within twice FP64's."""
    import cac_amd
    deck = decks.eam_block(n=10, seed=31)
    res = {}
    for prec in (cac_amd.FP64, cac_amd.MIXED):
        g = gpu_model(deck, "eam", 2.0, precision=prec)
        g.setup()
        th = g.run(2000, 0.001, thermo_every=50)
        g.close()
        t, e = th[:, 0] * 0.001, th[:, 3]
        res[prec] = (float(np.max(np.abs(e - e[0]))), abs(float(np.polyfit(t, e, 1)[0])))
    (x64, d64), (x32, d32) = res[cac_amd.FP64], res[cac_amd.MIXED]
    floor = 1e-9 * deck.natoms
    assert x32 <= 2 * x64 + floor, res
    assert d32 <= 2 * d64 + floor, res


def test_nve_drift_over_1e4_steps_matches_oracle():
    """NVE on a 500-atom EAM block at 300 K, 10,000 steps of 1 fs, from the same
    state in the oracle, the FP64 path and the mixed path.  The trajectories
    part after a few hundred steps (the summation orders differ, and MD
    amplifies differences), so what must match is energy conservation: E(0)
    equal, and each run's largest excursion |E(t) - E(0)| and drift (slope of a
    linear fit of E(t), per atom) within twice the oracle's (the mixed run's
    also within twice FP64's), with a floor well below a faulty integrator's."""
    import cac_amd
    deck = decks.eam_block(n=5, seed=31)
    o = oracle_model(deck, "eam", 2.0)
    o.setup()
    runs = {"oracle": o.run(10000, 0.001, thermo_every=100)}
    o.close()
    for name, prec in (("fp64", cac_amd.FP64), ("mixed", cac_amd.MIXED)):
        g = gpu_model(deck, "eam", 2.0, precision=prec)
        g.setup()
        runs[name] = g.run(10000, 0.001, thermo_every=100)
        g.close()
    stats = {}
    for name, th in runs.items():
        t, e = th[:, 0] * 0.001, th[:, 3]
        assert len(e) == 101 and np.all(np.isfinite(e)), name
        stats[name] = (float(np.max(np.abs(e - e[0])) / abs(e[0])), abs(float(np.polyfit(t, e, 1)[0])) / deck.natoms)
    e0 = runs["oracle"][0, 3]
    assert abs(runs["fp64"][0, 3] - e0) <= 1e-12 * abs(e0)
    assert abs(runs["mixed"][0, 3] - e0) <= 1e-6 * abs(e0)
    xo, do = stats["oracle"]
    x_floor, d_floor = 1e-7, 2e-7  # relative excursion; eV per atom per ps
    for name in ("fp64", "mixed"):
        x, d = stats[name]
        assert x <= 2 * xo + x_floor and d <= 2 * do + d_floor, stats
    assert stats["mixed"][0] <= 2 * stats["fp64"][0] + x_floor, stats
    assert stats["mixed"][1] <= 2 * stats["fp64"][1] + d_floor, stats


# ------------------------------------------------------------------ host handover, closed forms, launch shapes

def test_host_handover_matches_oracle():
    """A caller that integrates on the host: positions in, decide, compute, forces out."""
    deck = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=12)
    g = gpu_model(deck, "eam", 2.0)
    g.setup()
    rng = np.random.default_rng(4)
    moved = dataclasses.replace(deck, x=deck.x + rng.normal(0, 0.03, deck.x.shape),
                                nodex=deck.nodex + rng.normal(0, 0.03, deck.nodex.shape))
    g.sync_positions(moved.x, moved.nodex)
    if g.decide():
        g.build_neighbor()
    g.compute(1, 1)
    f, nodef, gaussf = g.forces()
    o = oracle_model(moved, "eam", 2.0)
    o.setup()
    _, _, fo = o.atoms()
    eo = o.elements()
    assert rel(f, fo) < 1e-12
    assert rel(gaussf, eo["gaussf"]) < 1e-12
    assert rel(nodef, eo["nodef"]) < 1e-12
    assert abs(g.energy()[0] - o.energy()[0]) < 1e-12 * abs(o.energy()[0])


@pytest.mark.parametrize("delay,expect", [(0, 6), (10, 5)])
def test_rebuild_schedule_ballistic(delay, expect):
    """One free atom at 0.07 A per step, skin 1.0 A: builds every 8 steps with
    delay 0 (6 in 50), every 10 with delay 10, each of them dangerous (5)."""
    g0 = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3) * 10.0 + 1.0
    v = np.zeros_like(g0)
    v[5] = (70.0, 0.0, 0.0)
    deck = decks.Deck(units="metal", boxlo=np.zeros(3), boxhi=np.full(3, 40.0), masses=np.array([63.546]), x=g0,
                      v=v, atype=np.ones(len(g0), dtype=np.int32))
    g = gpu_model(deck, "eam", 1.0, delay=delay)
    g.setup()
    g.run(50, 0.001, thermo=False)
    assert g.neigh_stats() == (expect, expect if delay else 0)
    x, _, _ = g.atoms()
    assert np.allclose(x[5], np.mod(g0[5] + v[5] * 0.05, 40.0), atol=1e-12)


def test_periodic_wrap_moves_elements_rigidly():
    deck = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.0, T=0.0, jitter_atoms=False)
    V = np.array([900.0, -400.0, 250.0])
    deck.v = np.tile(V, (deck.natoms, 1))
    deck.nodev = np.tile(V, (deck.nelements, 8, 1))
    g = gpu_model(deck, "eam", 2.0, delay=0)
    g.setup()
    g.run(60, 0.001, thermo=False)
    L = deck.boxhi - deck.boxlo
    x, _, f = g.atoms()
    el = g.elements()
    assert np.abs(f).max() < 1e-9 and np.abs(el["gaussf"]).max() < 1e-9
    assert np.allclose(x, deck.boxlo + np.mod(deck.x + V * 0.06 - deck.boxlo, L), atol=1e-9)
    nb = deck.nodex + V * 0.06
    k = np.floor((nb.mean(axis=1) - deck.boxlo) / L)
    assert np.allclose(el["nodex"], nb - k[:, None, :] * L, atol=1e-9)


def test_launch_shape_does_not_change_results(monkeypatch):
    """Workgroup size and entries in flight change the schedule only: the FP64
    and mixed forces are bitwise the same for every pair-kernel shape."""
    import cac_amd
    deck = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=13)
    out = {}
    for prec in (cac_amd.FP64, cac_amd.MIXED):
        for shape in range(4):
            monkeypatch.setenv("CAC_HIP_PAIR_SHAPE", str(shape))
            g = gpu_model(deck, "eam", 2.0, precision=prec)
            g.setup()
            g.run(5, 0.001, thermo=False)
            out[(prec, shape)] = np.concatenate([g.atoms()[2].ravel(), g.elements()["gaussf"].ravel()])
        for shape in range(1, 4):
            assert np.array_equal(out[(prec, 0)], out[(prec, shape)]), (prec, shape)
