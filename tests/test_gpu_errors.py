"""Error behaviour of the C ABI on the GPU: each failure the reference reports
through error->all comes back as its CAC_HIP_E* code with the reference's
message, and the context stays usable where the reference would go on (a
growing list is rebuilt; an overflow past neigh_modify one is an error)."""

import ctypes as C

import numpy as np
import pytest

from cac_amd import decks

pytestmark = pytest.mark.gpu

TAB = decks.cu_eam_tables()


def small_cac(seed=3):
    return decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=seed)


def test_list_overflow_past_one_is_an_error():
    import cac_amd
    g = cac_amd.Model(small_cac(), "eam", skin=2.0, eam_tables=TAB, one=60)
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.setup()
    assert ei.value.code == 2 and "neigh_modify one" in str(ei.value)
    g.close()


def test_non_numeric_position_is_an_error():
    import cac_amd
    d = small_cac()
    d.x[5, 1] = np.nan
    g = cac_amd.Model(d, "eam", skin=2.0, eam_tables=TAB)
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.setup()
    assert ei.value.code == 3 and "Non-numeric" in str(ei.value)
    g.close()


def test_element_deformed_twice_its_size_is_an_error():
    """Element::update_element_bound_box's check (element.cpp:1452-1461): an
    element stretched to three times its initial extent stops the next build."""
    import cac_amd
    d = small_cac()
    g = cac_amd.Model(d, "eam", skin=2.0, eam_tables=TAB)
    g.setup()
    nodex = g.elements()["nodex"].copy()
    c = nodex[0].mean(axis=0)
    nodex[0] = c + 3.0 * (nodex[0] - c)
    g.sync_positions(g.atoms()[0], nodex)
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.build_neighbor()
    assert ei.value.code == 4 and "deformed" in str(ei.value)
    g.close()


def test_box_shorter_than_the_cutoff_is_an_error():
    """A box shorter than rc + skin would need images beyond the nearest."""
    import cac_amd
    d = decks.lj_block(n=1, jitter=0.0, T=0.0)  # 1.59 sigma box, rc + skin = 2.8
    g = cac_amd.Model(d, "lj", skin=0.3)
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.setup()
    assert ei.value.code == 1 and "box shorter than rc + skin" in str(ei.value)
    g.close()


def test_calls_out_of_order_and_bad_arguments():
    import cac_amd
    L = cac_amd.lib()
    h = C.c_void_p()
    assert L.cac_hip_create(C.byref(h), 0, 1) == 0
    try:
        lo, hi = np.zeros(3), np.array([10.0, -1.0, 10.0])
        assert L.cac_hip_set_box(h, lo, hi) == 1  # hi must exceed lo
        assert L.cac_hip_setup(h, 1, 1) == 7  # no box yet
        assert L.cac_hip_set_box(h, lo, np.full(3, 20.0)) == 0
        assert L.cac_hip_setup(h, 1, 1) == 7  # no atoms or elements
        assert b"no atoms or elements" in L.cac_hip_last_error(h)
        assert L.cac_hip_set_masses(h, 0, np.ones(1)) == 1
        assert L.cac_hip_set_precision(h, 5) == 1
        th = np.zeros(5)
        assert L.cac_hip_run(h, 1, 0.001, 0, th.ctypes.data_as(C.c_void_p)) == 7  # not set up
    finally:
        L.cac_hip_destroy(h)


def test_pbc_wraps_and_leaves_the_lists_stale():
    """cac_hip_pbc (Domain::pbc as write_data runs it): an atom and an element
    pushed past a periodic face come back inside the box, shifted by a box
    length; a run or a neighbor decision before the next setup is
    CAC_HIP_ESTATE; setup rebuilds and the model runs on."""
    import cac_amd
    d = small_cac()
    L = d.boxhi - d.boxlo
    a = int(np.argmax(d.x[:, 0]))
    d.x[a, 0] = d.boxhi[0] + 0.25  # past +x
    e = 0
    d.nodex[e, :, 2] -= d.nodex[e, :, 2].mean() - (d.boxlo[2] - 0.5)  # centre below -z
    g = cac_amd.Model(d, "eam", skin=2.0, eam_tables=TAB)
    g.pbc()
    x = g.atoms()[0]
    nx = g.elements()["nodex"]
    assert x[a, 0] == (d.boxhi[0] + 0.25) - L[0]
    # rigid shift by the wrapped centre minus the node mean: L up to rounding
    assert np.allclose(nx[e, :, 2], d.nodex[e, :, 2] + L[2], rtol=0, atol=1e-12)
    assert np.array_equal(nx[e, :, :2], d.nodex[e, :, :2])
    assert np.all((x >= d.boxlo) & (x < d.boxhi))
    cen = nx.mean(axis=1)
    assert np.all((cen >= d.boxlo - 1e-12) & (cen < d.boxhi + 1e-12))
    g.close()
    g = cac_amd.Model(small_cac(), "eam", skin=2.0, eam_tables=TAB)
    g.setup()
    g.pbc()
    with pytest.raises(cac_amd.CacHipError) as ei:
        g.run(1, 0.001)
    assert ei.value.code == 7
    g.setup()
    th = g.run(5, 0.001, thermo_every=5)
    assert np.all(np.isfinite(th))
    g.close()


def test_capacity_grows_without_changing_lists():
    """The first build sizes rows from the mean site density.  An LJ slab that
    fills a quarter of its box has rows about four times that estimate: the
    build overflows, grows the capacity and builds again, and the lists equal
    the oracle's."""
    import cac_amd
    from oracle import oracle as O
    d = decks.lj_block(n=6, jitter=0.02, T=0.1, seed=4)
    d.boxhi = d.boxhi.copy()
    d.boxhi[0] = d.boxlo[0] + 4.0 * (d.boxhi[0] - d.boxlo[0])
    cutneigh = 2.5 + 0.3
    vol = float(np.prod(d.boxhi - d.boxlo))
    est = d.natoms / vol * 4.0 / 3.0 * np.pi * cutneigh ** 3
    first_cap = (int(1.5 * est) + 32 + 15) // 16 * 16
    g = cac_amd.Model(d, "lj", skin=0.3, delay=0)
    g.setup()
    o = O.Oracle(d, "lj", skin=0.3, delay=0)
    o.setup()
    p, k, ow, u, img = o.half_list()
    ref = O.canonical_full_sets(d, p, k, ow, u, img, np.zeros(8, dtype=np.int64))
    rows = np.bincount(ref[:, 0], minlength=d.natoms)
    assert rows.max() > first_cap  # the first attempt overflowed
    assert np.array_equal(g.full_sets(), ref)
    assert g.neigh_info()["capacity"] >= rows.max()
    g.close()


def test_section_timing_modes():
    """Timing mode 2 times only the pair force pass, one event pair per launch
    read back after the run; mode 1 times every section; other modes are an
    argument error."""
    import cac_amd
    g = cac_amd.Model(small_cac(), "eam", skin=2.0, eam_tables=TAB, precision=cac_amd.MIXED)
    g.setup()
    g.enable_timing(2)
    g.run(12, 0.001, thermo=False)
    us, n = g.kernel_time("eam_force")
    assert n == 12 and us > 0.0
    assert g.kernel_time("eam_density") == (0.0, 0)
    g.enable_timing(1)
    g.run(3, 0.001, thermo=False)
    for name in ("eam_density", "eam_embed", "eam_force"):
        us, n = g.kernel_time(name)
        assert n == 3 and us > 0.0, name
    us, n = g.kernel_time("site_refresh")
    assert n >= 3 and us > 0.0
    with pytest.raises(cac_amd.CacHipError):
        g.enable_timing(3)
    g.enable_timing(0)
    g.close()
