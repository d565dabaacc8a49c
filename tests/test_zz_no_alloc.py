"""cac_hip_thermo_noalloc, the entry point that makes no device allocation,
against cac_hip_thermo on the same state: PE equal, KE, temperature and
pressure within 1e-12 relative (the kinetic-energy sum runs in another fixed
tree).  Two calls in a row on one GPU, the second on a smaller model, so the
second reads only partials it wrote itself."""

import numpy as np
import pytest

from cac_amd import decks

pytestmark = pytest.mark.gpu

TAB = decks.cu_eam_tables()


def _check(g):
    a, b = g.thermo(), g.thermo(noalloc=True)
    assert b["pe"] == a["pe"]
    for key in ("ke", "temp", "press"):
        assert b[key] == pytest.approx(a[key], rel=1e-12), (key, a[key], b[key])
    assert b["temp"] > 0


def test_thermo_noalloc_matches_thermo():
    import cac_amd
    # 32,000 atoms: 8 partial sums; then a CAC model of 259 atoms + elements: 1
    big = cac_amd.Model(decks.lj_block(n=20, jitter=0.01, T=0.1, seed=61), "lj", skin=0.3)
    big.setup()
    _check(big)
    small = cac_amd.Model(decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=62),
                          "eam", skin=1.0, eam_tables=TAB)
    small.setup()
    _check(small)
    # the state moves on; the small model's call left its partials in the array
    big.run(10, 0.005, thermo_every=0, thermo=False)
    _check(big)
    small.run(10, 0.001, thermo_every=0, thermo=False)
    _check(small)
