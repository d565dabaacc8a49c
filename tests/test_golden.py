"""The frozen oracle fixtures of tests/golden/ (made by
tests/golden/make_oracle_fixtures.py): each deck is rebuilt from its fixture
alone.  On the CPU the oracle must reproduce its stored outputs; on the GPU the
HIP path must match them to the parity bars (neighbor sets exact, FP64 forces,
energy and virial within 1e-12 relative, the 10-step state within 1e-10,
mixed-precision forces within 1e-5)."""

import hashlib
import os

import numpy as np
import pytest

from cac_amd import decks

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("lj_atoms_500", "cac_eam_ncell3", "cac_eam_ncell9")


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def load(name):
    d = np.load(os.path.join(HERE, name + ".npz"))  # no pickles: allow_pickle stays False
    deck = decks.Deck(units="metal" if int(d["units"]) else "lj", boxlo=d["boxlo"], boxhi=d["boxhi"],
                      masses=d["masses"], x=d["x"], v=d["v"], atype=d["atype"], ncell=tuple(int(c) for c in d["ncell"]),
                      nodex=d["nodex"], nodev=d["nodev"], etype=d["etype"], ctype=d["ctype"], name=name)
    pair = "eam" if int(d["pair"]) else "lj"
    tables = None
    if pair == "eam":
        tables = decks.EamTables(nrho=int(d["nrho"]), drho=float(d["drho"]), nr=int(d["nr"]), dr=float(d["dr"]),
                                 cut=float(d["cut"]), frho=d["frho"], rhor=d["rhor"], z2r=d["z2r"],
                                 mass=float(d["mass"]), lattice_a=float(d["lattice_a"]))
    run = dict(pair=pair, skin=float(d["skin"]), delay=int(d["delay"]), dt=float(d["dt"]), nsteps=int(d["nsteps"]))
    out = {k[4:]: d[k] for k in d.files if k.startswith("out_")}
    return deck, tables, run, out


def check_sets(sets, deck, out):
    rows = np.ascontiguousarray(sets, dtype=np.int64)
    assert np.array_equal(np.bincount(rows[:, 0], minlength=deck.atom_equivalents), out["sets_count"])
    assert hashlib.sha256(rows.tobytes()).hexdigest() == str(out["sets_sha256"])
    if "sets" in out:
        assert np.array_equal(rows, out["sets"].astype(np.int64))


def check_setup(f, el, e, vir, out, tol):
    if f.size:
        assert rel(f, out["f0"]) < tol, rel(f, out["f0"])
    if el["gaussf"].size:
        assert rel(el["gaussf"], out["gaussf0"]) < tol
        assert rel(el["nodef"], out["nodef0"]) < tol
    assert abs(e - float(out["e0"])) <= tol * abs(float(out["e0"]))
    assert rel(vir, out["vir0"]) < tol


def test_fixtures_are_small_and_hold_no_pickles():
    total = 0
    for name in CASES:
        path = os.path.join(HERE, name + ".npz")
        total += os.path.getsize(path)
        d = np.load(path)
        assert all(d[k].dtype != object for k in d.files)
    assert total < 2 * 1024 * 1024


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    from oracle import oracle as O
    deck, tables, run, out = load(name)
    o = O.Oracle(deck, run["pair"], skin=run["skin"], delay=run["delay"], eam_tables=tables)
    o.setup()
    p, k, ow, u, img = o.half_list()
    nc = deck.ncell
    g2u = np.array([i * nc[1] * nc[2] + j * nc[2] + q for i in (0, nc[0] - 1) for j in (0, nc[1] - 1)
                    for q in (0, nc[2] - 1)])
    check_sets(O.canonical_full_sets(deck, p, k, ow, u, img, g2u), deck, out)
    _, _, f = o.atoms()
    e, vir = o.energy()
    check_setup(f, o.elements(), e, vir, out, 1e-13)
    if run["pair"] == "eam":
        ar, nr = o.rho()
        assert rel(ar, out["arho0"]) < 1e-13 and (nr.size == 0 or rel(nr, out["nrho0"]) < 1e-13)
    th = o.run(run["nsteps"], run["dt"])
    x, v, _ = o.atoms()
    assert rel(x, out["x_end"]) < 1e-13 and rel(v, out["v_end"]) < 1e-12
    if deck.nelements:
        el = o.elements()
        assert rel(el["nodex"], out["nodex_end"]) < 1e-13 and rel(el["nodev"], out["nodev_end"]) < 1e-12
    assert rel(th[-1], out["thermo_end"]) < 1e-12
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_path_matches_fixture(name):
    import cac_amd
    deck, tables, run, out = load(name)
    g = cac_amd.Model(deck, run["pair"], skin=run["skin"], delay=run["delay"], eam_tables=tables)
    g.setup()
    check_sets(g.full_sets(), deck, out)
    _, _, f = g.atoms()
    e, vir = g.energy()
    check_setup(f, g.elements(), e, vir, out, 1e-12)
    if run["pair"] == "eam":
        ar, nr = g.rho()
        assert rel(ar, out["arho0"]) < 1e-12 and (nr.size == 0 or rel(nr, out["nrho0"]) < 1e-12)
    th = g.run(run["nsteps"], run["dt"])
    x = g.atoms()[0]
    assert rel(x, out["x_end"]) < 1e-10
    if deck.nelements:
        assert rel(g.elements()["nodex"], out["nodex_end"]) < 1e-10
    assert abs(th[-1][1] - out["thermo_end"][1]) <= 1e-10 * abs(out["thermo_end"][1])  # potential energy
    g.close()
    # mixed precision: forces at set-up within 1e-5 of the frozen FP64 ones
    m = cac_amd.Model(deck, run["pair"], skin=run["skin"], delay=run["delay"], eam_tables=tables,
                      precision=cac_amd.MIXED)
    m.setup()
    fm = m.atoms()[2]
    if fm.size:
        assert rel(fm, out["f0"]) < 1e-5, rel(fm, out["f0"])
    if deck.nelements:
        assert rel(m.elements()["gaussf"], out["gaussf0"]) < 1e-5
    m.close()
