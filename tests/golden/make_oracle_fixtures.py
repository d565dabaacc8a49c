#!/usr/bin/env python3
"""Frozen outputs of the oracle on three small decks, together with the decks.

The reference holds no golden vectors of its own (no tests, decks, potential
files or logs).  These fixtures are therefore not reference outputs: they are
the oracle's results (oracle/, the single-rank C restatement of the
reference's algorithm, pinned by tests/test_oracle_pins.py), frozen so that a
later change to the oracle or to the HIP path that moves them shows up against
fixed numbers.  tests/test_golden.py rebuilds each deck from its fixture alone
(inputs are stored, not regenerated), runs the oracle on the CPU and the HIP
path on the GPU, and compares both with the stored outputs.

Each .npz holds, for one deck: the box, masses, atoms, elements and
velocities; the pair style, skin, neighbor delay and time step; for EAM the
setfl arrays; and the oracle's outputs: canonical full neighbor sets (rows of
point, kind, owner, unit cell, image), forces on atoms, Gauss points and nodes,
energy and virial at set-up, densities (EAM), and positions, velocities and
the thermo row after NSTEPS velocity-Verlet steps.

Regenerate only on purpose, from the repository root:

    python tests/golden/make_oracle_fixtures.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NSTEPS = 10

CASES = {
    # 500-atom LJ FCC block (the smallest whose box exceeds 2.5 (rc + skin)),
    # the reference's LJ settings (rc 2.5, skin 0.3)
    "lj_atoms_500": dict(deck=lambda d: d.lj_block(n=5, jitter=0.05, T=0.1, seed=3), pair="lj", skin=0.3,
                         delay=0, dt=0.005),
    # CAC EAM with ncell 3 elements around an atomistic core
    "cac_eam_ncell3": dict(deck=lambda d: d.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03,
                                                    T=300.0, seed=4), pair="eam", skin=1.0, delay=10, dt=0.001),
    # CAC EAM with the workload's 9^3 elements
    "cac_eam_ncell9": dict(deck=lambda d: d.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=9, jitter=0.03,
                                                    T=300.0, seed=10), pair="eam", skin=2.0, delay=10, dt=0.001),
}


MAX_SETS = 50_000  # larger neighbor sets are stored as per-point counts and a digest


def sets_digest(sets, npoints):
    """Per-point counts and the SHA-256 of the canonical (sorted, int64) set rows."""
    import hashlib
    rows = np.ascontiguousarray(sets, dtype=np.int64)
    return {"sets_count": np.bincount(rows[:, 0], minlength=npoints).astype(np.int32),
            "sets_sha256": np.array(hashlib.sha256(rows.tobytes()).hexdigest())}


def oracle_outputs(deck, pair, skin, delay, dt, tables):
    """The oracle's set-up outputs and its state after NSTEPS steps."""
    from oracle import oracle as O
    g2u = np.array([i * deck.ncell[1] * deck.ncell[2] + j * deck.ncell[2] + k
                    for i in (0, deck.ncell[0] - 1) for j in (0, deck.ncell[1] - 1) for k in (0, deck.ncell[2] - 1)])
    o = O.Oracle(deck, pair, skin=skin, delay=delay, eam_tables=tables if pair == "eam" else None)
    o.setup()
    out = {}
    p, k, ow, u, img = o.half_list()
    sets = O.canonical_full_sets(deck, p, k, ow, u, img, g2u).astype(np.int64)
    out.update(sets_digest(sets, deck.atom_equivalents))
    if len(sets) <= MAX_SETS:
        out["sets"] = sets.astype(np.int32)
    _, _, f = o.atoms()
    out["f0"] = f
    el = o.elements()
    out["gaussf0"], out["nodef0"] = el["gaussf"], el["nodef"]
    e, vir = o.energy()
    out["e0"], out["vir0"] = np.array(e), vir
    if pair == "eam":
        out["arho0"], out["nrho0"] = o.rho()
    th = o.run(NSTEPS, dt)
    x, v, _ = o.atoms()
    el = o.elements()
    out["x_end"], out["v_end"] = x, v
    out["nodex_end"], out["nodev_end"] = el["nodex"], el["nodev"]
    out["thermo_end"] = th[-1]
    o.close()
    return out


def main():
    from cac_amd import decks
    tables = decks.cu_eam_tables()
    for name, c in CASES.items():
        deck = c["deck"](decks)
        arrays = dict(
            units=np.array(1 if deck.units == "metal" else 0), boxlo=deck.boxlo, boxhi=deck.boxhi,
            masses=deck.masses, x=deck.x, v=deck.v, atype=deck.atype.astype(np.int32),
            ncell=np.array(deck.ncell, dtype=np.int32), nodex=deck.nodex, nodev=deck.nodev,
            etype=deck.etype.astype(np.int32), ctype=deck.ctype.astype(np.int32),
            pair=np.array(1 if c["pair"] == "eam" else 0), skin=np.array(c["skin"]), delay=np.array(c["delay"]),
            dt=np.array(c["dt"]), nsteps=np.array(NSTEPS))
        if c["pair"] == "eam":
            arrays.update(nrho=np.array(tables.nrho), drho=np.array(tables.drho), nr=np.array(tables.nr),
                          dr=np.array(tables.dr), cut=np.array(tables.cut), frho=tables.frho, rhor=tables.rhor,
                          z2r=tables.z2r, mass=np.array(tables.mass), lattice_a=np.array(tables.lattice_a))
        out = oracle_outputs(deck, c["pair"], c["skin"], c["delay"], c["dt"], tables)
        arrays.update({"out_" + k: v for k, v in out.items()})
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {deck.natoms} atoms, {deck.nelements} elements, {int(out['sets_count'].sum())} neighbor pairs, "
              f"{os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
