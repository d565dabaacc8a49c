"""Pins for the CPU restatement (the oracle).

The reference ships no tests, decks or golden outputs and cannot be built here
(it needs MPI and its whole 184-file program), so every oracle function is
pinned against mathematics that a plausible mistake would break: closed forms,
brute force over explicit periodic images, finite differences, invariants and
the ncell = 2 limit in which a CAC element is exactly eight atoms.
"""

import numpy as np
import pytest

from cac_amd import decks
from oracle import oracle as O
from tests import brute

O.build()
TAB = decks.cu_eam_tables()


# ---------------------------------------------------------------- splines (pair_eam.cpp:586-614)

def test_spline_reproduces_cubic():
    """The 5-point derivative is exact for cubics and the Hermite segment then
    reproduces the cubic; a wrong coefficient or sign breaks this."""
    n, dx = 60, 0.05
    xs = np.arange(n) * dx  # knot m (1-based) at (m-1)*dx

    def f(x):
        return 0.3 - 1.1 * x + 0.7 * x ** 2 - 0.9 * x ** 3

    def fp(x):
        return -1.1 + 1.4 * x - 2.7 * x ** 2

    spl = O.interpolate(np.concatenate([[0.0], f(xs)]), dx)
    for m in range(3, n - 3):
        for p in (0.0, 0.25, 0.5, 0.9):
            x = (m - 1 + p) * dx
            val, der = brute.spline_eval(spl, m, p)
            assert abs(val - f(x)) < 1e-12
            assert abs(der - fp(x)) < 1e-10
    # end knots use the 2-point derivative, and the last knot has no cubic term
    assert spl[n, 3] == 0.0 and spl[n, 4] == 0.0
    assert spl[1, 5] == pytest.approx(f(xs[1]) - f(xs[0]), abs=1e-15)


def test_spline_quartic_derivative_exact():
    n, dx = 40, 0.1
    xs = np.arange(n) * dx
    spl = O.interpolate(np.concatenate([[0.0], xs ** 4]), dx)
    for m in range(3, n - 1):
        assert spl[m, 5] == pytest.approx(4 * xs[m - 1] ** 3 * dx, rel=1e-12, abs=1e-13)


# ---------------------------------------------------------------- element tables

@pytest.mark.parametrize("nc", [2, 3, 9, 11])
def test_element_tables_closed_forms(nc):
    t = O.etype_tables((nc, nc, nc))
    nu = nc ** 3
    assert t["nucell"] == nu and t["ngcell"] == 8
    corners = brute.corner_ucells((nc, nc, nc))
    assert list(t["g2u"]) == list(corners)
    # nodes follow from N = 1: Hex natural signs vs the i-major corner order
    assert list(t["g2n"]) == [0, 4, 3, 7, 1, 5, 2, 6]
    assert all(t["g2n"][t["n2g"][k]] == k for k in range(8))
    assert t["weight"].sum() == pytest.approx(nu)
    assert np.allclose(t["weight"], nu / 8.0)
    assert t["nodal_weight"] == nu / 8.0
    shape = t["shape"]
    assert np.allclose(shape.sum(axis=1), 1.0, atol=1e-15)
    assert np.allclose(shape, brute.shape_hex(brute.ucell_natural((nc, nc, nc))), atol=1e-15)
    for g, u in enumerate(t["g2u"]):
        e = np.zeros(8)
        e[t["g2n"][g]] = 1.0
        assert np.array_equal(shape[u], e)
    u2g = t["u2g"]
    assert (u2g >= 0).sum() == 8
    # CONSISTENT weighted shape: N * w / nucell = delta / 8 at the nodes
    assert np.allclose(t["wsa"], np.eye(8)[t["g2n"]] / 8.0)
    # sub-elements partition the unit cells exactly once
    allu = np.concatenate([t["us2u"][s, :t["nucell_sub"][s]] for s in range(t["nsub"])])
    assert np.array_equal(np.sort(allu), np.arange(nu))
    es = int(round(np.sqrt(np.cbrt(nu))))
    assert t["nsub"] == es ** 3


def test_mass_matrices():
    m, mi = O.mass_matrices()
    assert np.allclose(m @ mi, np.eye(8), atol=1e-13)
    assert m.sum() == pytest.approx(1.0)
    assert set(np.unique(mi)) == {64.0, -32.0, 16.0, -8.0}


# ---------------------------------------------------------------- LJ vs brute force

def _lj_deck(n=5, jitter=0.05, seed=1):
    return decks.lj_block(n=n, jitter=jitter, T=0.1, seed=seed)


def _lj_energy(x, L, cut=2.5):
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.einsum("ijk,ijk->ij", d, d)
    np.fill_diagonal(r2, np.inf)
    m = r2 < cut * cut
    r6 = np.where(m, 1.0 / np.where(m, r2, 1.0) ** 3, 0.0)
    return 0.5 * np.sum(4.0 * (r6 * r6 - r6))


def test_lj_energy_forces_brute_force():
    d = _lj_deck()
    o = O.Oracle(d, "lj", skin=0.3, delay=0)
    o.setup()
    x, _, f = o.atoms()
    L = d.boxhi - d.boxlo
    e, vir = o.energy()
    assert e == pytest.approx(_lj_energy(x, L), rel=1e-12)
    assert np.abs(f.sum(axis=0)).max() < 1e-10
    h = 1e-6
    for i in (0, 77, 311):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = -(_lj_energy(xp, L) - _lj_energy(xm, L)) / (2 * h)
            assert f[i, k] == pytest.approx(fd, rel=1e-6, abs=1e-6)
    # pair virial sum_{i<j} r_ij (x) f_ij equals sum_i x_i . f_i for a periodic
    # system only up to image terms; check trace against the direct pair sum
    dd = x[:, None, :] - x[None, :, :]
    dd -= L * np.round(dd / L)
    r2 = np.einsum("ijk,ijk->ij", dd, dd)
    np.fill_diagonal(r2, np.inf)
    msk = r2 < 6.25
    r2i = np.where(msk, 1.0 / np.where(msk, r2, 1.0), 0.0)
    r6 = r2i ** 3
    fpair = r6 * (48 * r6 - 24) * r2i
    vxx = 0.5 * np.sum(dd[..., 0] ** 2 * fpair)
    assert vir[0] == pytest.approx(vxx, rel=1e-11)


def test_nve_time_reversible():
    d = _lj_deck(jitter=0.02)
    o = O.Oracle(d, "lj", skin=0.3, delay=0)
    o.setup()
    x0, v0, _ = o.atoms()
    o.run(50, 0.005)
    x1, v1, _ = o.atoms()
    d2 = decks.Deck(units="lj", boxlo=d.boxlo, boxhi=d.boxhi, masses=d.masses, x=x1, v=-v1, atype=d.atype)
    o2 = O.Oracle(d2, "lj", skin=0.3, delay=0)
    o2.setup()
    o2.run(50, 0.005)
    x2, v2, _ = o2.atoms()
    L = d.boxhi - d.boxlo
    dx = x2 - x0
    dx -= L * np.round(dx / L)
    assert np.abs(dx).max() < 1e-9
    assert np.abs(v2 + v0).max() < 1e-9


def test_nve_energy_error_scales_dt2():
    """Velocity Verlet's energy error falls as dt^2 (smooth tapered EAM; the
    unshifted LJ cutoff would add dt-independent jumps)."""
    d = _eam_atoms_deck(n=6, jitter=0.05)
    rng = np.random.default_rng(3)
    d.v = rng.normal(0.0, 2.0, d.x.shape)
    errs = []
    for dt in (0.002, 0.001):
        o = O.Oracle(d, "eam", skin=2.0, eam_tables=TAB)
        o.setup()
        th = o.run(int(round(0.2 / dt)), dt, thermo_every=int(round(0.01 / dt)))
        errs.append(np.abs(th[:, 3] - th[0, 3]).max())
    assert 3.0 < errs[0] / errs[1] < 5.5


# ---------------------------------------------------------------- EAM, pure atoms

def _eam_atoms_deck(n=6, jitter=0.05, seed=7):
    a = 3.615
    rng = np.random.default_rng(seed)
    x = decks.fcc_sites(a, (n, n, n)) + rng.normal(0, jitter, (4 * n ** 3, 3))
    L = n * a
    x = np.mod(x, L)
    return decks.Deck(units="metal", boxlo=np.zeros(3), boxhi=np.full(3, L), masses=np.array([63.546]), x=x,
                      v=np.zeros_like(x), atype=np.ones(len(x), dtype=np.int32))


def test_eam_atoms_brute_force_and_fd():
    d = _eam_atoms_deck()
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    x, _, f = o.atoms()
    L = d.boxhi - d.boxlo
    e, _ = o.energy()
    assert e == pytest.approx(brute.eam_atoms_energy(x, L, TAB), rel=1e-12)
    # the splines track the analytic functions they tabulate
    assert e == pytest.approx(brute.eam_atoms_energy(x, L, TAB, analytic=True), rel=1e-6)
    assert np.abs(f.sum(axis=0)).max() < 1e-9
    h = 1e-5
    for i in (3, 500):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = -(brute.eam_atoms_energy(xp, L, TAB) - brute.eam_atoms_energy(xm, L, TAB)) / (2 * h)
            assert f[i, k] == pytest.approx(fd, rel=1e-5, abs=1e-5)


# ---------------------------------------------------------------- CAC closed forms

def _perfect_energy_per_site(a=3.615, strain=(0.0, 0.0, 0.0)):
    x = decks.fcc_sites(a, (6, 6, 6)) * (1.0 + np.asarray(strain))
    L = 6 * a * (1.0 + np.asarray(strain))
    return brute.eam_atoms_energy(x, L, TAB) / x.shape[0]


@pytest.mark.parametrize("nc,box,core", [(2, (4, 4, 4), 20), (3, (3, 3, 3), 8), (5, (2, 2, 2), 4), (9, (2, 2, 2), 4)])
def test_cac_perfect_lattice_energy_and_zero_force(nc, box, core):
    """A perfect lattice represented by atoms and elements has the energy of
    every represented site at the cohesive energy (nodal weights, energy scales
    and list completeness) and no force anywhere."""
    d = decks.cac_eam(box_cells=box, core_cells=core, ncell=nc, jitter=0.0, T=0.0, jitter_atoms=False)
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    e, _ = o.energy()
    assert e / d.represented_sites == pytest.approx(_perfect_energy_per_site(), rel=1e-12)
    _, _, f = o.atoms()
    el = o.elements()
    assert np.abs(f).max() < 1e-11
    assert np.abs(el["gaussf"]).max() < 1e-11
    assert np.abs(el["nodef"]).max() < 1e-10


def test_cac_affine_strain_zero_force():
    strain = np.array([0.01, -0.005, 0.0])
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.0, T=0.0, jitter_atoms=False)
    F = 1.0 + strain
    d.x = d.x * F
    d.nodex = d.nodex * F
    d.boxhi = d.boxhi * F
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    e, _ = o.energy()
    assert e / d.represented_sites == pytest.approx(_perfect_energy_per_site(strain=strain), rel=1e-12)
    assert np.abs(o.elements()["gaussf"]).max() < 1e-11


def test_cac_ncell2_is_atomistic():
    """With 2x2x2 sites per element every site is a node and a Gauss point, so
    the CAC Gauss forces, densities and energy equal the all-atom ones."""
    d = decks.cac_eam(box_cells=(4, 4, 4), core_cells=20, ncell=2, jitter=0.04, T=0.0, seed=11)
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    e_cac, vir_cac = o.energy()
    _, _, fa = o.atoms()
    el = o.elements()
    ar, nr = o.rho()
    allx = np.concatenate([d.x, d.nodex.reshape(-1, 3)], axis=0)
    da = decks.Deck(units="metal", boxlo=d.boxlo, boxhi=d.boxhi, masses=d.masses, x=np.mod(allx, d.boxhi),
                    v=np.zeros_like(allx), atype=np.ones(len(allx), dtype=np.int32))
    oa = O.Oracle(da, "eam", skin=1.0, eam_tables=TAB)
    oa.setup()
    e_at, vir_at = oa.energy()
    _, _, f_at = oa.atoms()
    rho_at, _ = oa.rho()
    A = d.natoms
    assert e_cac == pytest.approx(e_at, rel=1e-12)
    assert np.allclose(vir_cac, vir_at, rtol=1e-11, atol=1e-9)
    assert np.allclose(fa, f_at[:A], atol=1e-10)
    # Gauss cell g sits at node g2n[g]; the all-atom index of node k of element e is A + 8e + k
    g2n = np.array([0, 4, 3, 7, 1, 5, 2, 6])
    fnode = f_at[A:].reshape(-1, 8, 3)
    assert np.allclose(el["gaussf"], fnode[:, g2n, :], atol=1e-10)
    assert np.allclose(ar, rho_at[:A], rtol=1e-13)
    assert np.allclose(nr, rho_at[A:].reshape(-1, 8), rtol=1e-13)
    # consistent nodal assembly: 8 M nodef = gaussf at the nodes (element.cpp:2178-2211)
    m, _ = O.mass_matrices()
    gf_node = np.zeros_like(el["gaussf"])
    gf_node[:, g2n, :] = el["gaussf"]
    assert np.allclose(8.0 * np.einsum("ab,ebd->ead", m, el["nodef"]), gf_node, atol=1e-9)


# ---------------------------------------------------------------- neighbor sets

# core 0: elements only, no atoms
@pytest.mark.parametrize("nc,box,core,skin", [(3, (3, 3, 3), 8, 1.0), (5, (2, 2, 2), 4, 2.0), (3, (3, 3, 3), 0, 1.0)])
def test_neighbor_sets_match_brute_force(nc, box, core, skin):
    """Half/bin/newton lists over atoms, Gauss points and virtual atoms, ghosts
    and sub-element pruning expand to exactly the brute-force set of
    (point, site, image) pairs within rc + skin."""
    d = decks.cac_eam(box_cells=box, core_cells=core, ncell=nc, jitter=0.05, T=0.0, seed=nc)
    o = O.Oracle(d, "eam", skin=skin, eam_tables=TAB)
    o.build_neighbor()
    p, k, ow, u, img = o.half_list()
    g2u = brute.corner_ucells(d.ncell)
    got = O.canonical_full_sets(d, p, k, ow, u, img, g2u)
    want = brute.full_sets(d, TAB.cut + skin)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    # every unordered point-point pair is stored exactly once
    pts = k == 0
    assert len(np.unique(np.stack([p[pts], ow[pts], img[pts, 0], img[pts, 1], img[pts, 2]], 1), axis=0)) == pts.sum()


def test_cac_force_is_deterministic():
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=300.0, seed=2)
    res = []
    for _ in range(2):
        o = O.Oracle(d, "eam", skin=2.0, eam_tables=TAB)
        o.setup()
        o.run(20, 0.001)
        res.append(o.elements()["nodex"])
    assert np.array_equal(res[0], res[1])


# ---------------------------------------------------------------- composite CAC force

@pytest.mark.parametrize("nc,box,core", [(3, (3, 3, 3), 8), (5, (2, 2, 2), 4), (3, (3, 3, 3), 0)])
def test_cac_eam_matches_all_pairs_definition(nc, box, core):
    """Forces, densities and energy of the half/bin/newton machinery (asymmetric
    pair storage, reactions only on atoms and Gauss points, per-pair F' of
    interpolated atoms from interpolated nodal densities, nodal energy weights)
    equal an all-pairs evaluation from the definitions."""
    d = decks.cac_eam(box_cells=box, core_cells=core, ncell=nc, jitter=0.06, T=0.0, seed=17 + nc)
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    fa, fg, ra, rn, e = brute.cac_eam(d, TAB)
    _, _, f = o.atoms()
    el = o.elements()
    ar, nr = o.rho()
    eo, _ = o.energy()
    scale = np.abs(fg).max()
    if d.natoms:
        assert np.abs(f - fa).max() < 1e-10 * scale
    assert np.abs(el["gaussf"] - fg).max() < 1e-10 * scale
    assert np.allclose(ar, ra, rtol=1e-12) and np.allclose(nr, rn, rtol=1e-12)
    assert eo == pytest.approx(e, rel=1e-12)


# ---------------------------------------------------------------- thermo closed forms

def test_thermo_uniform_translation():
    """A uniform velocity V on every atom and node: the consistent mass matrix
    sums to one per element, so KE = 1/2 m V^2 (atoms + nucell * elements) and
    T = m V^2 / (3 k_B) whatever the geometry (compute_temp.cpp:96-150)."""
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.03, T=0.0, seed=5)
    V = np.array([0.7, -1.1, 0.4])  # A/ps
    d.v = np.tile(V, (d.natoms, 1))
    d.nodev = np.tile(V, (d.nelements, 8, 1))
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    th = o.thermo()
    mvv2e, boltz = 1.0364269e-4, 8.617343e-5
    m = 63.546
    n = d.natoms + 27 * d.nelements
    ke = 0.5 * m * float(V @ V) * n * mvv2e
    assert th["ke"] == pytest.approx(ke, rel=1e-12)
    assert th["temp"] == pytest.approx(m * float(V @ V) * mvv2e / (3.0 * boltz), rel=1e-12)


def test_pressure_of_a_dilute_gas():
    """Atoms farther apart than the cutoff: no virial, so P = N k_B T / V in
    pressure units (compute_pressure.cpp:140-180)."""
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3) * 8.0 + 1.0
    v = rng.normal(0, 3.0, g.shape)
    d = decks.Deck(units="metal", boxlo=np.zeros(3), boxhi=np.full(3, 32.0), masses=np.array([63.546]), x=g, v=v,
                   atype=np.ones(len(g), dtype=np.int32))
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB)
    o.setup()
    th = o.thermo()
    boltz, nktv2p = 8.617343e-5, 1.6021765e6
    assert th["pe"] == 0.0
    assert th["press"] == pytest.approx(len(g) * boltz * th["temp"] / 32.0 ** 3 * nktv2p, rel=1e-12)


# ---------------------------------------------------------------- rebuild trigger and periodic wrap

def _dilute_gas(u, n=4, spacing=10.0):
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3) * spacing + 1.0
    v = np.zeros_like(g)
    v[5] = u
    L = n * spacing
    return decks.Deck(units="metal", boxlo=np.zeros(3), boxhi=np.full(3, L), masses=np.array([63.546]), x=g, v=v,
                      atype=np.ones(len(g), dtype=np.int32))


@pytest.mark.parametrize("delay,expect", [(0, 6), (10, 5)])
def test_rebuild_schedule_ballistic(delay, expect):
    """One free atom moving 0.07 A per step, skin 1.0 A: the trigger (moved more
    than skin/2 since the last build, neighbor.cpp:1617-1683) first holds 8
    steps after a build, so with delay 0 a build lands every 8 steps (6 in 50
    steps); with delay 10 the first check after a build is 10 steps on and
    already trips, so every build is a dangerous one (5 in 50)."""
    dt = 0.001
    d = _dilute_gas(np.array([70.0, 0.0, 0.0]))
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB, delay=delay)
    o.setup()
    o.run(50, dt)
    ncalls, ndanger = o.neigh_stats()
    assert ncalls == expect
    assert ndanger == (expect if delay else 0)
    x, _, _ = o.atoms()
    x0 = d.x[5] + d.v[5] * 50 * dt
    assert np.allclose(x[5], np.mod(x0, d.boxhi), atol=1e-12)


def test_periodic_wrap_moves_elements_rigidly():
    """A perfect CAC lattice translating at a uniform velocity feels no force;
    after crossing the box every atom sits at its wrapped ballistic position and
    every element at its ballistic position shifted by whole box lengths chosen
    by its centre (domain.cpp:671-737)."""
    d = decks.cac_eam(box_cells=(3, 3, 3), core_cells=8, ncell=3, jitter=0.0, T=0.0, jitter_atoms=False)
    V = np.array([900.0, -400.0, 250.0])  # A/ps: the box is crossed within the run
    d.v = np.tile(V, (d.natoms, 1))
    d.nodev = np.tile(V, (d.nelements, 8, 1))
    dt, n = 0.001, 60
    o = O.Oracle(d, "eam", skin=2.0, eam_tables=TAB, delay=0)
    o.setup()
    o.run(n, dt)
    L = d.boxhi - d.boxlo
    x, _, f = o.atoms()
    el = o.elements()
    assert np.abs(f).max() < 1e-9 and np.abs(el["gaussf"]).max() < 1e-9
    xa = d.x + V * n * dt
    assert np.allclose(x, d.boxlo + np.mod(xa - d.boxlo, L), atol=1e-9)
    nb = d.nodex + V * n * dt
    c = nb.mean(axis=1)
    k = np.floor((c - d.boxlo) / L)
    assert np.allclose(el["nodex"], nb - k[:, None, :] * L, atol=1e-9)


# ---------------------------------------------------------------- setfl reader (pair_eam_alloy.cpp:29-315)

# A two-element setfl written by hand: every table has distinct values and
# tokens are split across lines irregularly, as grab (pair_eam.cpp:292-304)
# allows.  Blocks: Cu F(rho), Cu rho(r), Ni F(rho), Ni rho(r), then z2r for
# (Cu,Cu), (Ni,Cu), (Ni,Ni) (lower triangle, row order).
SETFL_NRHO, SETFL_DRHO, SETFL_NR, SETFL_DR, SETFL_CUT = 6, 0.5, 7, 0.25, 1.5
SETFL_TABLES = {
    "Cu": ([0.0, -1.0, -1.5, -1.75, -1.9, -2.0], [3.0, 2.5, 2.0, 1.4, 0.9, 0.4, 0.0]),
    "Ni": ([0.0, -1.2, -1.7, -2.1, -2.3, -2.4], [4.0, 3.1, 2.2, 1.5, 0.8, 0.3, 0.1]),
}
SETFL_Z2R = {("Cu", "Cu"): [9.0, 5.0, 2.0, 0.5, -0.5, -0.2, 0.0],
             ("Ni", "Cu"): [8.0, 4.0, 1.5, 0.25, -0.6, -0.3, 0.0],
             ("Ni", "Ni"): [7.0, 3.5, 1.0, 0.0, -0.7, -0.4, -0.1]}


def _write_setfl2(path, names_line="2 Cu Ni", truncate=0):
    def lines_of(vals, per):
        return [" ".join(repr(v) for v in vals[i:i + per]) for i in range(0, len(vals), per)]

    out = ["comment 1", "comment 2", "comment 3", names_line,
           f"{SETFL_NRHO} {SETFL_DRHO} {SETFL_NR} {SETFL_DR} {SETFL_CUT}"]
    for el, per in (("Cu", 4), ("Ni", 5)):
        out.append(f"{29 if el == 'Cu' else 28} {63.546 if el == 'Cu' else 58.69} 3.615 fcc")
        frho, rhor = SETFL_TABLES[el]
        out += lines_of(frho, per) + lines_of(rhor, per + 1)
    for key in (("Cu", "Cu"), ("Ni", "Cu"), ("Ni", "Ni")):
        out += lines_of(SETFL_Z2R[key], 3)
    text = "\n".join(out) + "\n"
    if truncate:
        text = text[:-truncate]
    path.write_text(text)


@pytest.mark.parametrize("element", ["Cu", "Ni"])
def test_setfl_reader_picks_the_mapped_element(tmp_path, element):
    """The element an atom type maps to supplies F(rho), rho(r) and the z2r
    block (i, i) of the lower triangle; reading any other block (the other
    element, or the mixed (Ni, Cu) pair) gives different splines."""
    path = tmp_path / "two.setfl"
    _write_setfl2(path)
    fs, rs, zs = O.setfl_splines(str(path), element)

    def spl(vals, delta):
        return O.interpolate(np.concatenate([[0.0], vals]), delta)

    frho, rhor = SETFL_TABLES[element]
    assert np.array_equal(fs, spl(frho, SETFL_DRHO))
    assert np.array_equal(rs, spl(rhor, SETFL_DR))
    assert np.array_equal(zs, spl(SETFL_Z2R[(element, element)], SETFL_DR))
    other = "Ni" if element == "Cu" else "Cu"
    assert not np.array_equal(fs, spl(SETFL_TABLES[other][0], SETFL_DRHO))
    assert not np.array_equal(zs, spl(SETFL_Z2R[("Ni", "Cu")], SETFL_DR))
    # the 7th coefficient of each knot is the tabulated value itself (1-based knots)
    assert np.array_equal(zs[1:, 6], SETFL_Z2R[(element, element)])


def test_setfl_reader_errors(tmp_path):
    """The reference's errors: an unknown element, a names line whose word
    count is not nelements + 1, a short file."""
    path = tmp_path / "two.setfl"
    _write_setfl2(path)
    with pytest.raises(O.OracleError, match="No matching element"):
        O.setfl_splines(str(path), "Fe")
    _write_setfl2(path, names_line="2 Cu")
    with pytest.raises(O.OracleError, match="Incorrect element names"):
        O.setfl_splines(str(path), "Cu")
    _write_setfl2(path, truncate=8)
    with pytest.raises(O.OracleError, match="truncated"):
        O.setfl_splines(str(path), "Ni")


def test_setfl_round_trip_of_the_benchmark_potential(tmp_path):
    """The one-element file decks.write_setfl writes for the benchmark
    potential reads back into the same splines as the arrays themselves."""
    path = tmp_path / "cu.setfl"
    decks.write_setfl(str(path), TAB)
    fs, rs, zs = O.setfl_splines(str(path))
    o = O.Oracle(decks.lj_block(n=2), "eam", skin=1.0, eam_tables=TAB)
    fs2, rs2, zs2 = o.eam_splines()
    assert np.array_equal(fs, fs2[: fs.shape[0]]) and np.array_equal(rs, rs2) and np.array_equal(zs, zs2)


# ---------------------------------------------------------------- nodal force styles (element.cpp:2178-2246)

def _nodef_style(style, nc=5):
    d = decks.cac_eam(box_cells=(2, 2, 2), core_cells=4, ncell=nc, jitter=0.05, T=0.0, seed=31)
    o = O.Oracle(d, "eam", skin=1.0, eam_tables=TAB, nodal_force_style=style)
    o.setup()
    return o.elements()


def _node_gets_its_gauss_force(el, n2g):
    """nodef[k] == gaussf[n2g[k]] exactly, for every element and node."""
    return np.array_equal(el["nodef"], el["gaussf"][:, n2g, :])


@pytest.mark.parametrize("style", [1, 2])
def test_lumped_and_nodeonly_closed_form(style):
    """2x2x2 Gauss cells at the corners with the UNIFORM weight scale: the
    LUMPED weights N(g) w_g / nodal_weight are delta(g, n2g[k]) exactly
    (element_vec_cac.cpp:3255-3260), so LUMPED, like NODEONLY
    (element.cpp:2229-2240), gives each node its own Gauss cell's force."""
    t = O.etype_tables((5, 5, 5))
    el = _nodef_style(style)
    assert np.abs(el["gaussf"]).max() > 1e-3
    assert _node_gets_its_gauss_force(el, t["n2g"])


def test_lumped_check_catches_consistent_weights():
    """The closed-form check above rejects the plausible mistakes: the
    CONSISTENT weights N w / nucell in the LUMPED branch (nodef = gaussf / 8),
    and the CONSISTENT style itself (Minv applied)."""
    t = O.etype_tables((5, 5, 5))
    el = _nodef_style(1)
    planted = dict(el)
    planted["nodef"] = np.einsum("gk,egd->ekd", t["wsa"], el["gaussf"])  # sum_g wsa[g][k] gaussf[g]
    assert np.allclose(planted["nodef"], el["gaussf"][:, t["n2g"], :] / 8.0, rtol=1e-14, atol=0)
    assert not _node_gets_its_gauss_force(planted, t["n2g"])
    assert not _node_gets_its_gauss_force(_nodef_style(0), t["n2g"])


# ---------------------------------------------------------------- per-atom tallies (pair.cpp:266-476)

def _lj_peratom(x, L, cut=2.5):
    """Per-atom LJ energy 1/2 sum phi and virial 1/2 sum d d fpair over all
    minimum images within the cutoff (eps = sigma = 1)."""
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.einsum("ijk,ijk->ij", d, d)
    np.fill_diagonal(r2, np.inf)
    m = r2 < cut * cut
    r2i = np.where(m, 1.0 / np.where(m, r2, 1.0), 0.0)
    r6 = r2i ** 3
    e = 0.5 * np.sum(4.0 * (r6 * r6 - r6), axis=1)
    fpair = r6 * (48.0 * r6 - 24.0) * r2i
    comps = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    v = np.stack([0.5 * np.sum(d[..., a] * d[..., b] * fpair, axis=1) for a, b in comps], 1)
    return e, v


@pytest.mark.parametrize("vflag", [4, 8])
def test_lj_peratom_tallies_match_all_pairs(vflag):
    """ENERGY_ATOM and VIRIAL_ATOM (or VIRIAL_CENTROID, which falls back to the
    two-body per-atom virial for pair styles with CENTROID_SAME) give each atom
    half of each of its pair energies and pair virials (pair.cpp:280-281, 441-476)."""
    d = _lj_deck()
    o = O.Oracle(d, "lj", skin=0.3, delay=0)
    o.setup(eflag=3, vflag=1 | vflag)
    x, _, _ = o.atoms()
    e_b, v_b = _lj_peratom(x, d.boxhi - d.boxlo)
    ea, _ = o.peratom_energy()
    va, _ = o.peratom_virial()
    assert np.allclose(ea, e_b, rtol=1e-12, atol=1e-13)
    assert np.abs(va - v_b).max() < 1e-11 * np.abs(v_b).max()
    e, vir = o.energy()
    assert ea.sum() == pytest.approx(e, rel=1e-12)
    assert np.allclose(va.sum(0), vir, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("nc,box,core", [(3, (3, 3, 3), 8), (3, (3, 3, 3), 0), (5, (2, 2, 2), 4)])
def test_cac_eam_peratom_tallies_match_definition(nc, box, core):
    """Per-point energies and virials of the half list with reactions equal the
    definitions: an atom or node gets F(rho) plus half of each of its pair
    energies, and half of each pair virial (interpolated non-Gauss sites get
    nothing, pair_eam.cpp:863-918, 971-984); weighting nodes by nodal_weight
    the sums are the global tallies."""
    dk = decks.cac_eam(box_cells=box, core_cells=core, ncell=nc, jitter=0.06, T=0.0, seed=41 + nc)
    o = O.Oracle(dk, "eam", skin=1.0, eam_tables=TAB)
    o.setup(eflag=3, vflag=5)
    *_, e_b, pa = brute.cac_eam(dk, TAB, peratom=True)
    ea, en = o.peratom_energy()
    va, vn = o.peratom_virial()
    escale = max(np.abs(pa["enode"]).max(), np.abs(pa["eatom"]).max() if dk.natoms else 0.0)
    vscale = max(np.abs(pa["vnode"]).max(), np.abs(pa["vatom"]).max() if dk.natoms else 0.0)
    if dk.natoms:
        assert np.abs(ea - pa["eatom"]).max() < 1e-12 * escale
        assert np.abs(va - pa["vatom"]).max() < 1e-10 * vscale
    assert np.abs(en - pa["enode"]).max() < 1e-12 * escale
    assert np.abs(vn - pa["vnode"]).max() < 1e-10 * vscale
    nw = nc ** 3 / 8.0
    e, vir = o.energy()
    assert ea.sum() + nw * en.sum() == pytest.approx(e, rel=1e-12)
    assert np.allclose(va.sum(0) + nw * vn.sum((0, 1)), vir, rtol=1e-10, atol=1e-9)
    # a planted mistake: tallying the nodal-weighted share per node instead of
    # the half misses the definition
    assert np.abs(nw * en - pa["enode"]).max() > 1e-3 * escale
