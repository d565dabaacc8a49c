"""The C++ host layer (cac_amd/host): data-file reader on the CPU; the
Pair / Neighbor / FixNVE / Verlet mirrors driving the device through the C ABI
(GPU tests), checked against the oracle."""

import os
import subprocess

import numpy as np
import pytest

from cac_amd import decks

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
