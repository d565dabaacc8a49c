"""The C++ host layer (cac_amd/host): data-file reader on the CPU; the
Pair / Neighbor / FixNVE / Verlet mirrors driving the device through the C ABI
(GPU tests), checked against the oracle."""

import os
import re
import subprocess

import numpy as np
import pytest

from cac_amd import decks
from tests import brute

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cac_amd", "cac_run")
TAB = decks.cu_eam_tables()


def run(*args, check=True):
    r = subprocess.run([RUNNER, *map(str, args)], capture_output=True, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError(f"cac_run failed ({r.returncode}): {r.stderr}")
    return r


def parse_thermo(out):
    lines = out.splitlines()
    k = lines.index("Step PotEng KinEng TotEng Press")
    rows = []
    for ln in lines[k + 1:]:
        w = ln.split()
        if len(w) != 5 or not w[0].isdigit():
            break
        rows.append([float(x) for x in w])
    builds = int([ln for ln in lines if ln.startswith("Neighbor list builds")][0].split("=")[1])
    return np.array(rows), builds


def test_host_layer_built():
    assert os.path.exists(RUNNER)
    assert os.path.exists(os.path.join(ROOT, "cac_amd", "libcachost.so"))


def test_read_data_round_trip(tmp_path):
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=1)
    p = tmp_path / "deck.data"
    decks.write_data(str(p), d)
    out = run(p, "--parse-only").stdout.split("\n")
    assert out[0].split() == ["atoms", str(d.natoms), "elements", str(d.nelements), "types", "1", "ncell", "3", "3", "3"]
    box = [float(x) for x in out[1].split()[1:]]
    assert np.allclose(box, [d.boxlo[0], d.boxhi[0], d.boxlo[1], d.boxhi[1], d.boxlo[2], d.boxhi[2]], rtol=0, atol=0)
    sums = [float(x) for x in out[2].split()[1:]]
    assert np.allclose(sums, [d.x.sum(), d.v.sum(), d.nodex.sum(), d.nodev.sum()], rtol=1e-12)


def test_read_data_rejects_other_gauss_layouts(tmp_path):
    d = decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=5, jitter=0.0, T=0.0)
    p = tmp_path / "deck.data"
    decks.write_data(str(p), d)
    txt = p.read_text().replace("1 Hex 1 5 5 5 2 2 2", "1 Hex 1 5 5 5 3 3 3")
    p.write_text(txt)
    r = run(p, "--parse-only", check=False)
    assert r.returncode == 1 and "2 2 2 Gauss cells" in r.stderr


WRITE_DATA_DECKS = {
    "cac": lambda: decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=1),
    "atoms": lambda: decks.lj_block(n=4, jitter=0.05, T=0.5, seed=2),
    "elements": lambda: decks.cac_eam(box_cells=(2, 2, 2), core_cells=0, ncell=5, jitter=0.02, T=100.0, seed=3),
}


@pytest.mark.parametrize("name", sorted(WRITE_DATA_DECKS))
def test_write_data_round_trip(tmp_path, name):
    """write_data of the host model (write_data.cpp:188-600): the header and
    section lines as the reference prints them, every coordinate and velocity
    at %-1.16e, so read_data reads each double back exactly; writing the
    re-read model again gives the same bytes."""
    d = WRITE_DATA_DECKS[name]()
    a, b, c = tmp_path / "a.data", tmp_path / "b.data", tmp_path / "c.data"
    decks.write_data(str(a), d)
    run(a, "--parse-only", "--write-data", b)
    e = decks.read_data(str(b))
    for key in ("x", "v", "nodex", "nodev", "boxlo", "boxhi", "masses", "atype", "etype", "ctype"):
        assert np.array_equal(np.asarray(getattr(e, key)), np.asarray(getattr(d, key))), key
    lines = b.read_text().splitlines()
    assert lines[0] == "CAC data file via write_data, timestep = 0" and lines[1] == ""
    assert ("%d atoms" % d.natoms in lines) == (d.natoms > 0)
    assert ("%d elements" % d.nelements in lines) == (d.nelements > 0)
    assert ("1 element types" in lines) == (d.nelements > 0)
    assert lines[lines.index("Masses") + 2] == "1 %g" % d.masses[0]
    num = r"-?\d\.\d{16}e[+-]\d{2}"
    assert re.fullmatch(rf"{num} {num} xlo xhi", [ln for ln in lines if ln.endswith("xlo xhi")][0])
    if d.nelements:
        nc = d.ncell
        assert lines[lines.index("Element Types") + 2] == f"1 Hex 1 {nc[0]} {nc[1]} {nc[2]} 2 2 2 1 1 1"
        assert re.fullmatch(rf"1 1 1 1 {num} {num} {num}", lines[lines.index("Nodes") + 2])
        assert re.fullmatch(rf"1 1 1 {num} {num} {num}", lines[lines.index("Node Velocities") + 2])
        # Elements: the centre of each element's nodes
        cen = np.array([[float(v) for v in ln.split()[2:]] for ln in
                        lines[lines.index("Elements") + 2:lines.index("Elements") + 2 + d.nelements]])
        assert np.allclose(cen, d.nodex.mean(axis=1), rtol=0, atol=1e-12)
    if d.natoms:
        assert re.fullmatch(rf"1 1 {num} {num} {num}", lines[lines.index("Atoms") + 2])
        assert re.fullmatch(rf"1 {num} {num} {num}", lines[lines.index("Velocities") + 2])
    run(b, "--parse-only", "--write-data", c)
    assert c.read_bytes() == b.read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_write_data_restart_continues_bitwise(tmp_path, mixed):
    """A checkpoint from device state: 40 steps, write_data (Domain::pbc on the
    device, then the state written), and a second run of 40 steps in the same
    process; then read_data of the checkpoint in a new process and 40 steps.
    Both end states, written by write_data, agree byte for byte (positions and
    velocities of atoms and nodes), in FP64 and in mixed precision.  The
    checkpoint holds every point inside the box and differs from the start."""
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=600.0, seed=31)
    data, setfl = tmp_path / "deck.data", tmp_path / "cu.setfl"
    ck, fa, fb = tmp_path / "ck.data", tmp_path / "a.data", tmp_path / "b.data"
    decks.write_data(str(data), d)
    decks.write_setfl(str(setfl), TAB)
    prec = ["--mixed"] if mixed else []
    run(data, "--eam", setfl, "--skin", 1.0, "--steps", 40, *prec, "--write-data", ck, "--then", 40, fa)
    run(ck, "--eam", setfl, "--skin", 1.0, "--steps", 40, *prec, "--write-data", fb)
    la, lb = fa.read_text().splitlines(), fb.read_text().splitlines()
    assert la[0].endswith("timestep = 80") and lb[0].endswith("timestep = 40")
    assert la[1:] == lb[1:]
    c = decks.read_data(str(ck))
    L = c.boxhi - c.boxlo
    assert np.all((c.x >= c.boxlo) & (c.x < c.boxhi))
    cen = c.nodex.mean(axis=1)
    assert np.all((cen >= c.boxlo - 1e-9 * L) & (cen < c.boxhi + 1e-9 * L))
    assert not np.array_equal(c.x, d.x) and not np.array_equal(c.nodev, d.nodev)


@pytest.mark.gpu
@pytest.mark.parametrize("loop", ["host", "device"])
def test_host_layer_run_matches_oracle(tmp_path, loop):
    from oracle import oracle as O
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=21)
    data, setfl = tmp_path / "deck.data", tmp_path / "cu.setfl"
    decks.write_data(str(data), d)
    decks.write_setfl(str(setfl), TAB)
    rows, builds = parse_thermo(run(data, "--eam", setfl, "--skin", 1.0, "--delay", 0, "--steps", 30, "--thermo",
                                    10, "--loop", loop).stdout)
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=decks.read_setfl(str(setfl)), delay=0)
    o.setup()
    th = o.run(30, 0.001, thermo_every=10)
    assert rows.shape == th.shape
    assert np.array_equal(rows[:, 0], th[:, 0])
    assert np.allclose(rows[:, 1:4], th[:, 1:4], rtol=1e-10, atol=1e-9)
    assert builds == o.neigh_stats()[0]


# ------------------------------------------------------------------ output taps

def read_dump(path):
    """One snapshot of a LAMMPS-style text dump: (timestep, box[3,2], columns, rows[n, ncol])."""
    lines = open(path).read().splitlines()
    assert lines[0] == "ITEM: TIMESTEP" and lines[2] == "ITEM: NUMBER OF ATOMS"
    n = int(lines[3])
    assert lines[4].startswith("ITEM: BOX BOUNDS")
    box = np.array([[float(v) for v in lines[5 + d].split()] for d in range(3)])
    assert lines[8].startswith("ITEM: ATOMS ")
    cols = lines[8].split()[2:]
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[9:9 + n]]).reshape(n, len(cols))
    return int(lines[1]), box, cols, rows


DUMP_COLS = "id type x y z vx vy vz fx eid igauss inode iucell"


def dump_from_data(tmp_path, deck, fmt, cols=DUMP_COLS, extra="precision high average yes"):
    data = tmp_path / "deck.data"
    decks.write_data(str(data), deck)
    out = tmp_path / "snap.*.atom"
    run(data, "--parse-only", "--dump", f"1 all custom 1 {out} format {fmt} {extra} {cols}")
    return read_dump(str(tmp_path / "snap.0.atom"))


def test_dump_custom_node_and_gauss_formats(tmp_path):
    """dump custom, formats lammps/node and lammps/gauss/point: atoms then each
    element's 8 points, ids after the largest atom id, positions and velocities
    exact at 17 digits (precision high)."""
    d = decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=3, jitter=0.03, T=300.0, seed=3)
    A, E = d.natoms, d.nelements
    corners = brute.corner_ucells(d.ncell)
    for fmt in ("lammps/node", "lammps/gauss/point"):
        step, box, cols, r = dump_from_data(tmp_path, d, fmt)
        assert step == 0 and cols == DUMP_COLS.split()
        assert np.array_equal(box, np.stack([d.boxlo, d.boxhi], axis=1))
        assert r.shape[0] == A + 8 * E
        assert np.array_equal(r[:A, 0], np.arange(1, A + 1))
        assert np.array_equal(r[:A, 2:5], d.x) and np.array_equal(r[:A, 5:8], d.v)
        assert np.all(r[:A, 9:] == -1)
        el = r[A:].reshape(E, 8, -1)
        assert np.array_equal(el[:, :, 0], A + 1 + np.arange(8 * E).reshape(E, 8))
        assert np.array_equal(el[:, :, 9], np.repeat(np.arange(1, E + 1)[:, None], 8, axis=1))
        N = brute.shape_hex(brute.ucell_natural(d.ncell))
        node_of_gauss = np.array([int(np.argmax(N[u])) for u in corners])
        gauss_of_node = np.argsort(node_of_gauss)
        if fmt == "lammps/node":
            node, gauss = np.arange(8), gauss_of_node
        else:
            node, gauss = node_of_gauss, np.arange(8)
        assert np.array_equal(el[:, :, 2:5], d.nodex[:, node]) and np.array_equal(el[:, :, 5:8], d.nodev[:, node])
        assert np.array_equal(el[0, :, 10], gauss) and np.array_equal(el[0, :, 11], node)
        assert np.array_equal(el[0, :, 12], corners[gauss])
    assert not np.any(r[:, 8])  # no forces in a data file


def test_dump_custom_full_map_surface_center(tmp_path):
    """lammps/full/map interpolates every virtual atom; lammps/surface keeps the
    outer unit cells; lammps/center writes one line per element at its centre."""
    d = decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=5, jitter=0.03, T=300.0, seed=4)
    A, E, nu = d.natoms, d.nelements, d.nucell
    pos, kind, owner, ucell = brute.all_sites(d)
    _, _, cols, r = dump_from_data(tmp_path, d, "lammps/full/map")
    assert r.shape[0] == A + nu * E
    assert np.allclose(r[A:, 2:5], pos[A:], rtol=0, atol=1e-12)
    assert np.array_equal(r[A:, 12], ucell[A:]) and np.array_equal(r[A:, 0], A + 1 + np.arange(nu * E))
    _, _, _, rs = dump_from_data(tmp_path, d, "lammps/surface")
    assert rs.shape[0] == A + E * (nu - 3 ** 3)
    _, _, _, rc = dump_from_data(tmp_path, d, "lammps/center", cols="id type x y z eid iucell")
    assert rc.shape[0] == A + E
    assert np.allclose(rc[A:, 2:5], d.nodex.mean(axis=1), rtol=0, atol=1e-12)
    assert np.all(rc[A:, 1] == 0) and np.all(rc[A:, 6] == -1)


def test_dump_custom_argument_rules(tmp_path):
    """The reference's argument grammar: the keyword pair after wrap or
    precision is skipped; doubles print as %g unless precision high; unknown
    columns are an error."""
    d = decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=3, jitter=0.03, T=300.0, seed=5)
    _, _, cols, r = dump_from_data(tmp_path, d, "lammps/node", cols="id x", extra="wrap yes type y")
    assert cols == ["id", "x"]
    txt = (tmp_path / "snap.0.atom").read_text().splitlines()
    assert txt[9].split()[1] == f"{d.x[0, 0]:g}"
    data = tmp_path / "deck.data"
    r = run(data, "--parse-only", "--dump", "1 all custom 1 f.atom id nope", check=False)
    assert r.returncode == 1 and "unknown column keyword" in r.stderr
    r = run(data, "--parse-only", "--dump", "1 all custom 1 f.atom id ix", check=False)
    assert r.returncode == 1 and "image" in r.stderr


@pytest.mark.gpu
def test_dump_forces_match_oracle(tmp_path):
    """dump custom fx fy fz at step 0 (atoms, and nodal forces in lammps/node)
    against the oracle's forces."""
    from oracle import oracle as O
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=22)
    data, setfl = tmp_path / "deck.data", tmp_path / "cu.setfl"
    decks.write_data(str(data), d)
    decks.write_setfl(str(setfl), TAB)
    run(data, "--eam", setfl, "--skin", 1.0, "--steps", 0, "--dump",
        f"1 all custom 5 {tmp_path}/f.*.atom format lammps/node precision high average yes id x y z fx fy fz")
    _, _, cols, r = read_dump(str(tmp_path / "f.0.atom"))
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=decks.read_setfl(str(setfl)))
    o.setup()
    _, _, f = o.atoms()
    nodef = o.elements()["nodef"].reshape(-1, 3)
    want = np.concatenate([f, nodef])
    assert np.max(np.abs(r[:, 4:7] - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.gpu
def test_neighbor_debug_rows_match_oracle(tmp_path):
    """write_dump_neighbor_debug: list lengths of every row, and one atom row and
    one Gauss row entry by entry, against the oracle's lists as full sets."""
    from oracle import oracle as O
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=23)
    data, setfl = tmp_path / "deck.data", tmp_path / "cu.setfl"
    decks.write_data(str(data), d)
    decks.write_setfl(str(setfl), TAB)
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=decks.read_setfl(str(setfl)))
    o.setup()
    p, k, ow, u, img = o.half_list()
    sets = O.canonical_full_sets(d, p, k, ow, u, img, brute.corner_ucells(d.ncell))
    A, E = d.natoms, d.nelements
    counts = np.bincount(sets[:, 0], minlength=A + 8 * E)
    run(data, "--eam", setfl, "--skin", 1.0, "--steps", 0, "--neighbor-debug", f"{tmp_path}/nl")
    _, _, cols, r = read_dump(str(tmp_path / "nl_proc_0.dat"))
    assert cols == "id type x y z tag numneigh igcell ibasis".split()
    assert np.array_equal(r[:, 6], counts)
    assert np.array_equal(r[:A, 5], np.arange(1, A + 1)) and np.all(r[:A, 7] == -1)
    assert np.array_equal(r[A:, 7], np.tile(np.arange(8), E))
    L = d.boxhi - d.boxlo
    x_o = o.atoms()[0]
    N = brute.shape_hex(brute.ucell_natural(d.ncell))
    vpos = np.einsum("uk,ekd->eud", N, o.elements()["nodex"])
    for args, point in ((f"{tmp_path}/one atom 7", 6), (f"{tmp_path}/two gauss 5 0 3", A + 8 * 4 + 3)):
        run(data, "--eam", setfl, "--skin", 1.0, "--steps", 0, "--neighbor-debug", args)
        name = args.split()[0] + "_proc_0"
        _, _, cols, r = read_dump(name)
        assert cols == "id type x y z tag iucell ibasis".split()
        mine = r[1:]
        ref = sets[sets[:, 0] == point]
        assert mine.shape[0] == ref.shape[0]
        want_type = np.where(ref[:, 1] == 0, 2, 3)
        want_tag = ref[:, 2] + 1
        want_uc = np.where(ref[:, 1] == 0, -1, ref[:, 3])
        want_x = np.where(ref[:, 1:2] == 0, x_o[np.minimum(ref[:, 2], A - 1)],
                          vpos[np.minimum(ref[:, 2], E - 1), ref[:, 3]]) + ref[:, 4:7] * L
        left = list(range(ref.shape[0]))
        for row in mine:  # each written neighbor matches one oracle entry: kind, tag, unit cell, position
            hit = [q for q in left if (want_type[q], want_tag[q], want_uc[q]) == tuple(row[[1, 5, 6]])
                   and np.max(np.abs(want_x[q] - row[2:5])) < 1e-3]
            assert hit, row
            left.remove(hit[0])
