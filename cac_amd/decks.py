"""Synthetic CAC model decks: geometry, velocities, potential tables, data files.

Every generator is seeded and deterministic.  Positions follow the reference's
conventions:

* atoms are ``x[n][3]`` in box coordinates (``atom.h:36-39``);
* an element is a Hex with one basis atom per unit cell; its 8 nodes are listed
  in the Hex natural-coordinate order (-,-,-), (+,-,-), (+,+,-), (-,+,-),
  (-,-,+), (+,-,+), (+,+,+), (-,+,+) (``element_vec_cac.cpp:1756-1764``);
* element unit cells run i-major over the natural axes, so a site's ucell index
  is ``i*ny*nz + j*nz + k`` (``element_vec_cac.cpp:1881-1968``).

The CAC geometry puts an FCC atomistic core inside a shell of rhombohedral
elements whose edges lie along the primitive FCC vectors b1 = a/2(0,1,1),
b2 = a/2(1,0,1), b3 = a/2(1,1,0).  Element unit cells tile the lattice with a
stride of ``ncell`` sites, so no lattice site is shared
(``create_elements.cpp:603-608``).  A periodic box of ``ncell*a*(nx,ny,nz)`` is
commensurate with that element superlattice and holds 4 nx ny nz superlattice
cells; the cells nearest the box centre are atomistic and every other cell is
an element, so atoms and elements tile the crystal with no void.  (The
benchmark EAM has a positive cohesive energy at a = 3.615 A, so a free surface
or void inside the model evaporates; a void-free tiling is required.)

Units: LJ decks are in LJ units; EAM decks are in metal units (Angstrom, eV,
ps, amu) with the reference's constants (``update.cpp:155-173``).
"""

from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

BOLTZ_METAL = 8.617343e-5
MVV2E_METAL = 1.0364269e-4

HEX_SIGNS = np.array(
    [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
    dtype=np.int64,
)
FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])


@dataclasses.dataclass
class Deck:
    """One CAC model: box, atoms, elements (Hex, apc = 1) and velocities."""

    units: str
    boxlo: np.ndarray
    boxhi: np.ndarray
    masses: np.ndarray  # per atom type, type 1 first
    x: np.ndarray  # [A,3]
    v: np.ndarray  # [A,3]
    atype: np.ndarray  # [A] 1-based
    ncell: tuple = (9, 9, 9)  # element type 1
    nodex: np.ndarray = None  # [E,8,3]
    nodev: np.ndarray = None  # [E,8,3]
    etype: np.ndarray = None  # [E] 1-based
    ctype: np.ndarray = None  # [E] 1-based
    name: str = ""

    def __post_init__(self):
        if self.nodex is None:
            self.nodex = np.zeros((0, 8, 3))
            self.nodev = np.zeros((0, 8, 3))
            self.etype = np.zeros(0, dtype=np.int32)
            self.ctype = np.zeros(0, dtype=np.int32)

    @property
    def natoms(self) -> int:
        return int(self.x.shape[0])

    @property
    def nelements(self) -> int:
        return int(self.nodex.shape[0])

    @property
    def nucell(self) -> int:
        return int(self.ncell[0] * self.ncell[1] * self.ncell[2])

    @property
    def atom_equivalents(self) -> int:
        """Real atoms plus Gauss points (8 per element with 2x2x2 Gauss cells)."""
        return self.natoms + 8 * self.nelements

    @property
    def represented_sites(self) -> int:
        return self.natoms + self.nucell * self.nelements


# ----------------------------------------------------------------------------
# potentials


def lj_params():
    """C0: epsilon = sigma = 1, rc = 2.5 sigma."""
    return dict(eps=1.0, sigma=1.0, cut=2.5)


@dataclasses.dataclass
class EamTables:
    nrho: int
    drho: float
    nr: int
    dr: float
    cut: float
    frho: np.ndarray  # [nrho], F at rho = m*drho
    rhor: np.ndarray  # [nr], f at r = m*dr
    z2r: np.ndarray  # [nr], r*phi at r = m*dr
    mass: float
    lattice_a: float
    name: str = "Cu"


def _taper(r, r_in=4.45, r_out=4.95):
    """Cosine taper 0.5*(1 + cos(pi (r - r_in)/(r_out - r_in))) on [r_in, r_out]."""
    r = np.asarray(r, dtype=np.float64)
    t = np.ones_like(r)
    mid = (r > r_in) & (r < r_out)
    t[mid] = 0.5 * (1.0 + np.cos(np.pi * (r[mid] - r_in) / (r_out - r_in)))
    t[r >= r_out] = 0.0
    return t


def cu_phi(r):
    """Pair term phi(r) = 0.59 exp(-8 (r/2.556 - 1)) eV, tapered."""
    return 0.59 * np.exp(-8.0 * (np.asarray(r) / 2.556 - 1.0)) * _taper(r)


def cu_f(r):
    """Density function f(r) = exp(-5.85 (r/2.556 - 1)), tapered."""
    return np.exp(-5.85 * (np.asarray(r) / 2.556 - 1.0)) * _taper(r)


def cu_F(rho):
    """Embedding F(rho) = -2 sqrt(rho/12) eV."""
    return -2.0 * np.sqrt(np.maximum(np.asarray(rho, dtype=np.float64), 0.0) / 12.0)


def cu_eam_tables(nr: int = 2000, nrho: int = 2000, cut: float = 4.95, rhomax: float = 40.0) -> EamTables:
    """Analytic Cu-like EAM tabulated as a one-element setfl.

    Knot m (0-based) sits at r = m*dr and rho = m*drho, the setfl convention the
    reference's spline indexing assumes (``pair_eam.cpp:799-805``).  The taper
    multiplies both phi and f.
    """
    dr = cut / (nr - 1)
    drho = rhomax / (nrho - 1)
    r = np.arange(nr, dtype=np.float64) * dr
    rho = np.arange(nrho, dtype=np.float64) * drho
    return EamTables(
        nrho=nrho, drho=drho, nr=nr, dr=dr, cut=cut,
        frho=cu_F(rho), rhor=cu_f(r), z2r=r * cu_phi(r), mass=63.546, lattice_a=3.615,
    )


def write_setfl(path: str, t: EamTables) -> None:
    """Write a one-element setfl file (format read by pair_eam_alloy.cpp:100-196)."""
    with open(path, "w") as fh:
        fh.write("Cu-like analytic EAM (synthetic benchmark potential)\n")
        fh.write("phi = 0.59 exp(-8(r/2.556-1)); f = exp(-5.85(r/2.556-1)); F = -2 sqrt(rho/12)\n")
        fh.write("cosine taper 4.45-4.95 A on phi and f\n")
        fh.write(f"1 {t.name}\n")
        fh.write(f"{t.nrho} {t.drho:.17g} {t.nr} {t.dr:.17g} {t.cut:.17g}\n")
        fh.write(f"29 {t.mass:.17g} {t.lattice_a:.17g} fcc\n")
        for arr in (t.frho, t.rhor, t.z2r):
            for k in range(0, arr.size, 5):
                fh.write(" ".join(f"{v:.17e}" for v in arr[k:k + 5]) + "\n")


def read_setfl(path: str) -> EamTables:
    """Read the first element of a setfl file (pair_eam_alloy.cpp:100-196)."""
    with open(path) as fh:
        lines = fh.readlines()
    words = lines[3].split()
    nelem = int(words[0])
    nrho, drho, nr, dr, cut = lines[4].split()[:5]
    nrho, nr = int(nrho), int(nr)
    drho, dr, cut = float(drho), float(dr), float(cut)
    head = lines[5].split()
    mass = float(head[1])
    if nelem != 1:
        raise ValueError("only one-element setfl files are read")
    vals = np.array(" ".join(lines[6:]).split(), dtype=np.float64)
    frho = vals[:nrho]
    rhor = vals[nrho:nrho + nr]
    z2r = vals[nrho + nr:nrho + 2 * nr]
    return EamTables(nrho=nrho, drho=drho, nr=nr, dr=dr, cut=cut, frho=frho, rhor=rhor, z2r=z2r,
                     mass=mass, lattice_a=float(head[2]), name=words[1])


# ----------------------------------------------------------------------------
# geometry


def fcc_sites(a: float, n: tuple, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Cubic FCC sites of an n[0] x n[1] x n[2] block of unit cells, cell-major order."""
    ix, iy, iz = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    cells = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float64)
    pos = (cells[:, None, :] + FCC_BASIS[None, :, :]).reshape(-1, 3) * a
    return pos + np.asarray(origin, dtype=np.float64)


def primitive_vectors(a: float) -> np.ndarray:
    """Rows b1, b2, b3 of the primitive FCC lattice in the cubic frame."""
    return 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def element_nodes(a: float, ncell: int, cells: np.ndarray) -> np.ndarray:
    """Node positions [E,8,3] of superlattice cells (I,J,K): corner site
    (ncell*I, ncell*J, ncell*K) in primitive coordinates plus (ncell-1) steps
    along each primitive vector for the + corners."""
    b = primitive_vectors(a)
    corner = (ncell * cells.astype(np.float64)) @ b  # [E,3]
    steps = (HEX_SIGNS > 0).astype(np.float64) * (ncell - 1)  # [8,3] in primitive units
    return corner[:, None, :] + (steps @ b)[None, :, :]


def element_sites(a: float, ncell: int, cells: np.ndarray) -> np.ndarray:
    """All lattice sites [E, ncell^3, 3] of superlattice cells, ucell order i-major."""
    b = primitive_vectors(a)
    corner = (ncell * cells.astype(np.float64)) @ b
    g = np.arange(ncell)
    pi, pj, pk = np.meshgrid(g, g, g, indexing="ij")
    loc = np.stack([pi.ravel(), pj.ravel(), pk.ravel()], axis=1).astype(np.float64) @ b
    return corner[:, None, :] + loc[None, :, :]


def superlattice_cells(a: float, ncell: int, box_cells) -> np.ndarray:
    """One representative (I,J,K) per periodic class of element cells in the box
    [0, ncell*a*n)^3: the cell whose centre lies inside the box.  A box of
    ncell*a*(nx, ny, nz) is commensurate with the element superlattice
    (a x = b2 + b3 - b1 and cyclic) and holds 4 nx ny nz cells."""
    b = primitive_vectors(a)
    nb = np.asarray(box_cells, dtype=np.int64)
    L = ncell * a * nb.astype(np.float64)
    span = int(3 * nb.max() + 3)
    rng = np.arange(-span, span + 1)
    I, J, K = np.meshgrid(rng, rng, rng, indexing="ij")
    cells = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    centre = (ncell * cells + 0.5 * (ncell - 1)).astype(np.float64) @ b
    tol = 1e-9 * L
    keep = np.all((centre >= -tol) & (centre < L - tol), axis=1)
    cells = cells[keep]
    assert cells.shape[0] == 4 * int(np.prod(nb)), (cells.shape, nb)
    return cells


def cac_geometry(a: float, ncell: int, box_cells, core_cells: int, core_axis=None):
    """Void-free CAC geometry: the periodic box ncell*a*(nx, ny, nz) is tiled by
    the element superlattice; the ``core_cells`` cells nearest the box centre
    (nearest the box's central line along ``core_axis`` when one is given: a
    cylindrical core) are atomistic (their ncell^3 lattice sites become atoms)
    and every other cell is an element.  Returns (box lengths, atom positions,
    element cells, nodes)."""
    b = primitive_vectors(a)
    L = ncell * a * np.asarray(box_cells, dtype=np.float64)
    cells = superlattice_cells(a, ncell, box_cells)
    centre = (ncell * cells + 0.5 * (ncell - 1)).astype(np.float64) @ b
    off = centre - 0.5 * L
    if core_axis is not None:
        off[:, core_axis] = 0.0
    d = np.linalg.norm(off, axis=1)
    idx = np.arange(cells.shape[0])
    order = idx[np.lexsort((idx, np.round(d, 9)))]
    core = np.sort(order[:core_cells])
    rest = np.sort(order[core_cells:])
    atoms = element_sites(a, ncell, cells[core]).reshape(-1, 3) if core_cells else np.zeros((0, 3))
    atoms = atoms - L * np.floor(atoms / L)
    nodes = element_nodes(a, ncell, cells[rest])
    return L, atoms, cells[rest], nodes


# ----------------------------------------------------------------------------
# velocities


def _mass_matrix_hex() -> np.ndarray:
    """Consistent Hex mass matrix normalised to unit volume (element.cpp:1999-2177)."""
    m1 = np.array([[1.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 1.0 / 3.0]])
    sgn = (HEX_SIGNS > 0).astype(int)
    M = np.ones((8, 8))
    for d in range(3):
        M *= m1[sgn[:, d][:, None], sgn[:, d][None, :]]
    return M


def cac_temperature(v, mass_atom, nodev, mass_node, nucell, boltz, mvv2e) -> float:
    """ComputeTemp::compute_scalar with the consistent mass matrix (compute_temp.cpp:73-157)."""
    t = float(np.sum(mass_atom * np.sum(v * v, axis=1)))
    if nodev.shape[0]:
        M = _mass_matrix_hex()
        vv = np.einsum("ekd,eld->ekl", nodev, nodev)
        t += float(np.sum(np.einsum("kl,ekl->e", M, vv) * mass_node * nucell))
    dof = 3.0 * (v.shape[0] + nucell * nodev.shape[0])
    return t * mvv2e / (dof * boltz)


def maxwell(rng, n_atoms, n_elem, mass, T, boltz, mvv2e, nucell):
    """Gaussian velocities with zero net momentum, rescaled so the reference's
    temperature (consistent mass matrix, dof = 3*(atoms + virtual atoms)) is T."""
    sig = math.sqrt(boltz * T / (mass * mvv2e))
    v = rng.normal(0.0, sig, size=(n_atoms, 3))
    nv = rng.normal(0.0, sig, size=(n_elem, 8, 3))
    # zero total momentum: atoms carry m, an element carries m*nucell at its mean node velocity
    p = mass * v.sum(axis=0) + mass * nucell * nv.mean(axis=1).sum(axis=0)
    mtot = mass * (n_atoms + nucell * n_elem)
    dv = p / mtot
    v -= dv
    nv -= dv
    if T > 0:
        t_now = cac_temperature(v, mass, nv, mass, nucell, boltz, mvv2e)
        s = math.sqrt(T / t_now)
        v *= s
        nv *= s
    else:
        v[:] = 0.0
        nv[:] = 0.0
    return v, nv


# ----------------------------------------------------------------------------
# the benchmark decks


def lj_block(n: int = 20, a: float = 1.5874, jitter: float = 0.01, T: float = 0.1, seed: int = 12345) -> Deck:
    """Pure-atomistic LJ: periodic FCC n^3 cells, Gaussian jitter, Maxwell velocities
    at T* with zero net momentum.  n = 20 is the 32,000-atom workload."""
    rng = np.random.default_rng(seed)
    x = fcc_sites(a, (n, n, n))
    x = x + rng.normal(0.0, jitter, size=x.shape)
    L = n * a
    x = x - L * np.floor(x / L)
    v, _ = maxwell(rng, x.shape[0], 0, 1.0, T, 1.0, 1.0, 1)
    return Deck(units="lj", boxlo=np.zeros(3), boxhi=np.full(3, L), masses=np.array([1.0]), x=x, v=v,
                atype=np.ones(x.shape[0], dtype=np.int32), name=f"lj_fcc_{n}")


def eam_block(n: int = 20, a: float = 3.615, jitter: float = 0.02, T: float = 300.0, seed: int = 2024,
              mass: float = 63.546) -> Deck:
    """Pure-atomistic EAM Cu: periodic FCC n^3 cells (n = 20: 32,000 atoms, the
    atomistic core of the 1x CAC workload as a periodic block), Gaussian jitter,
    Maxwell velocities at T with zero net momentum."""
    rng = np.random.default_rng(seed)
    x = fcc_sites(a, (n, n, n))
    x = x + rng.normal(0.0, jitter, size=x.shape)
    L = n * a
    x = x - L * np.floor(x / L)
    v, _ = maxwell(rng, x.shape[0], 0, mass, T, BOLTZ_METAL, MVV2E_METAL, 1)
    return Deck(units="metal", boxlo=np.zeros(3), boxhi=np.full(3, L), masses=np.array([mass]), x=x, v=v,
                atype=np.ones(x.shape[0], dtype=np.int32), name=f"eam_fcc_{n}")


def cac_eam(box_cells=(7, 8, 8), core_cells: int = 44, ncell: int = 9, jitter: float = 0.02, T: float = 300.0,
            seed: int = 2024, a: float = 3.615, mass: float = 63.546, jitter_atoms: bool = True) -> Deck:
    """EAM CAC deck on the void-free geometry of :func:`cac_geometry`.  Defaults
    are the benchmark workload: 44 core cells = 32,076 atoms inside 1,748 Hex
    elements of 9^3 sites, box 63a x 72a x 72a; box (14,16,16) with 351 core
    cells is its 8x scale (255,879 atoms, 13,985 elements).  Nodes (and atoms)
    get Gaussian jitter; velocities are Maxwell at T with zero net momentum."""
    rng = np.random.default_rng(seed)
    L, x, cells, nodes = cac_geometry(a, ncell, box_cells, core_cells)
    nodes = nodes + rng.normal(0.0, jitter, size=nodes.shape)
    if jitter_atoms and x.shape[0]:
        x = x + rng.normal(0.0, jitter, size=x.shape)
        x = x - L * np.floor(x / L)
    v, nv = maxwell(rng, x.shape[0], nodes.shape[0], mass, T, BOLTZ_METAL, MVV2E_METAL, ncell ** 3)
    E = nodes.shape[0]
    return Deck(units="metal", boxlo=np.zeros(3), boxhi=L.copy(), masses=np.array([mass]), x=x, v=v,
                atype=np.ones(x.shape[0], dtype=np.int32), ncell=(ncell, ncell, ncell), nodex=nodes,
                nodev=nv, etype=np.ones(E, dtype=np.int32), ctype=np.ones(E, dtype=np.int32),
                name=f"cac_eam_box{'x'.join(map(str, box_cells))}_core{core_cells}_n{ncell}")


def edge_dislocation_field(x1, x2, b: float, nu: float, theta: float):
    """Isotropic Volterra displacement (u1, u2) of a straight edge dislocation
    at the origin of the (x1, x2) plane with Burgers vector b along the
    direction at angle theta from x1 (displace.cpp:956-974): in the frame
    rotated by theta, u1 = b/2pi [atan2(r2, r1) + pi + r1 r2 / (2 (1 - nu) r^2)],
    u2 = -b / (8 pi (1 - nu)) [(1 - 2 nu) ln r^2 + (r1^2 - r2^2) / r^2], then
    rotated back.  The cut (the jump of b in u1) lies on the half-plane r2 = 0,
    r1 < 0."""
    c, s = np.cos(theta), np.sin(theta)
    r1 = x1 * c + x2 * s
    r2 = -x1 * s + x2 * c
    rsq = r1 * r1 + r2 * r2
    if np.any(rsq < 1.0e-4):
        raise ValueError("a site is too close to the dislocation line")
    u1 = b / (2.0 * np.pi) * ((np.arctan2(r2, r1) + np.pi) + r1 * r2 / (2.0 * (1.0 - nu) * rsq))
    u2 = -b / (8.0 * np.pi * (1.0 - nu)) * ((1.0 - 2.0 * nu) * np.log(rsq) + (r1 * r1 - r2 * r2) / rsq)
    return u1 * c - u2 * s, u1 * s + u2 * c


def cac_dislocation(box_cells=(20, 20, 26), core_cells: int = 1503, ncell: int = 11, T: float = 300.0,
                    seed: int = 2025, a: float = 3.615, mass: float = 63.546, nu: float = 0.34,
                    separation: int = 24) -> Deck:
    """Large single system: FCC Cu holding edge dislocations with b = a/2 [110]
    and lines along z, inserted by the Volterra field of
    :func:`edge_dislocation_field` (nu = 0.34) on atoms and element nodes.

    A single Volterra dislocation in a periodic box leaves its cut open where
    it meets the box boundary (nodes of neighbouring elements 0.7 A apart on
    this box), so the periodic share holds it as a dipole: the dislocation and
    one of opposite sign ``separation`` Burgers vectors away along b.  Their
    cuts cancel outside the strip between the lines, which lies in the
    atomistic core, and the far field decays.  The lines sit at (a/8, 3a/8)
    from a lattice row in (x, y), so that no row lies on a line or on the cut
    plane (a row on the cut would slip by b onto its neighbour).

    The atomistic core is the cylinder of ``core_cells`` superlattice cells
    nearest the box's central z line; the rest are Hex elements of ncell^3
    sites.  Defaults: 1,503 cells of 11^3 sites = 2,000,493 atoms in 40,097
    elements, 55.4 M represented sites, box 795 x 795 x 1034 A.  Velocities
    Maxwell at T with zero net momentum."""
    rng = np.random.default_rng(seed)
    L, x, cells, nodes = cac_geometry(a, ncell, box_cells, core_cells, core_axis=2)
    bvec, theta = a / np.sqrt(2.0), 0.25 * np.pi
    step = 0.5 * a * np.array([1.0, 1.0])  # one Burgers vector along [110]
    c1 = 0.5 * L[:2] + a * np.array([0.125, 0.375]) - 0.5 * separation * step
    c2 = c1 + separation * step

    def displaced(p):
        u1, u2 = edge_dislocation_field(p[..., 0] - c1[0], p[..., 1] - c1[1], bvec, nu, theta)
        w1, w2 = edge_dislocation_field(p[..., 0] - c2[0], p[..., 1] - c2[1], -bvec, nu, theta)
        q = p.copy()
        q[..., 0] += u1 + w1
        q[..., 1] += u2 + w2
        return q

    x = displaced(x)
    x = x - L * np.floor(x / L)
    nodes = displaced(nodes)
    v, nv = maxwell(rng, x.shape[0], nodes.shape[0], mass, T, BOLTZ_METAL, MVV2E_METAL, ncell ** 3)
    E = nodes.shape[0]
    return Deck(units="metal", boxlo=np.zeros(3), boxhi=L.copy(), masses=np.array([mass]), x=x, v=v,
                atype=np.ones(x.shape[0], dtype=np.int32), ncell=(ncell, ncell, ncell), nodex=nodes,
                nodev=nv, etype=np.ones(E, dtype=np.int32), ctype=np.ones(E, dtype=np.int32),
                name=f"cac_dislocation_box{'x'.join(map(str, box_cells))}_core{core_cells}_n{ncell}")


def strain_uniaxial(deck: Deck, axis: int, eps: float) -> Deck:
    """The deck stretched by (1 + eps) along one box axis, about boxlo: box,
    atom positions and nodes scale together (a uniaxial pre-strain)."""
    s = np.ones(3)
    s[axis] += eps
    lo = deck.boxlo
    out = dataclasses.replace(deck, boxhi=lo + (deck.boxhi - lo) * s, x=lo + (deck.x - lo) * s)
    if deck.nelements:
        out.nodex = lo + (deck.nodex - lo) * s
    return out


def serving_box(size: float, ratio: float):
    """Superlattice box (nx, ny, nz) and core cells of a CAC replica of about
    ``size`` atom-equivalents whose atomistic core holds ``ratio`` times as many
    cells as its coarse region (9^3-site cells: 729 atoms per core cell, 8 Gauss
    points per element); boxes up to 3:1 are taken, a size error weighing
    more than the aspect ratio."""
    per_cell = (729.0 * ratio + 8.0) / (1.0 + ratio)
    m = max(size / per_cell / 4.0, 1.0)
    best = None
    top = int(np.ceil(m ** (1.0 / 3.0))) + 1
    for nx in range(1, top + 1):
        for ny in range(nx, int(np.sqrt(m / nx)) + 2):
            nz = max(ny, int(round(m / (nx * ny))))
            cells = 4 * nx * ny * nz
            core = int(round(ratio / (1.0 + ratio) * cells))
            ae = 729 * core + 8 * (cells - core)
            if nz > 3 * nx:
                continue
            key = abs(np.log(ae / size)) + 0.05 * np.log(nz / nx)
            if best is None or key < best[0]:
                best = (key, (nx, ny, nz), core)
    return best[1], best[2]


def serving_stream(n: int = 512, seed: int = 77, rate: float = 25.0, size_range=(2000.0, 200000.0),
                   ratio_range=(0.05, 0.3), t_range=(100.0, 600.0), strain: float = 0.02):
    """Replica requests of a serving run: Poisson arrivals at ``rate`` per
    second; size log-uniform over ``size_range`` atom-equivalents; core/coarse
    cell ratio, temperature and uniaxial pre-strain (random axis) uniform.
    Returns a list of dicts (arrival s, size, ratio, T, axis, strain, seed);
    :func:`serving_deck` builds the deck of one."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for k in range(n):
        out.append(dict(arrival=float(t[k]), size=float(np.exp(rng.uniform(*np.log(size_range)))),
                        ratio=float(rng.uniform(*ratio_range)), T=float(rng.uniform(*t_range)),
                        axis=int(rng.integers(0, 3)), strain=float(rng.uniform(-strain, strain)),
                        seed=int(rng.integers(0, 2 ** 31))))
    return out


def serving_deck(spec: dict) -> Deck:
    """The EAM CAC deck of one serving request (9^3-site Hex elements)."""
    box, core = serving_box(spec["size"], spec["ratio"])
    d = cac_eam(box_cells=box, core_cells=core, ncell=9, jitter=0.02, T=spec["T"], seed=spec["seed"])
    return strain_uniaxial(d, spec["axis"], spec["strain"])


def wrap_atoms(deck: Deck) -> Deck:
    L = deck.boxhi - deck.boxlo
    deck.x = deck.boxlo + np.mod(deck.x - deck.boxlo, L)
    return deck


# ----------------------------------------------------------------------------
# CAC data file (read_data sections, read_data.cpp:338-375)


def write_data(path: str, deck: Deck) -> None:
    """Write a CAC data file: Masses, Element Types, Atoms, Velocities, Elements,
    Nodes, Node Velocities (formats of atom_vec_atomic.cpp and element_vec_cac.cpp)."""
    with open(path, "w") as fh:
        fh.write(f"CAC data file {deck.name}\n\n")
        fh.write(f"{deck.natoms} atoms\n{deck.nelements} elements\n{len(deck.masses)} atom types\n")
        fh.write("1 element types\n\n")
        for d, nm in enumerate("xyz"):
            fh.write(f"{deck.boxlo[d]:.17g} {deck.boxhi[d]:.17g} {nm}lo {nm}hi\n")
        fh.write("\nMasses\n\n")
        for i, mm in enumerate(deck.masses):
            fh.write(f"{i + 1} {mm:.17g}\n")
        fh.write("\nElement Types\n\n")
        nc = deck.ncell
        fh.write(f"1 Hex 1 {nc[0]} {nc[1]} {nc[2]} 2 2 2\n")
        if deck.natoms:
            fh.write("\nAtoms\n\n")
            for i in range(deck.natoms):
                p = deck.x[i]
                fh.write(f"{i + 1} {deck.atype[i]} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
            fh.write("\nVelocities\n\n")
            for i in range(deck.natoms):
                p = deck.v[i]
                fh.write(f"{i + 1} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
        if deck.nelements:
            fh.write("\nElements\n\n")
            c = deck.nodex.mean(axis=1)
            for e in range(deck.nelements):
                fh.write(f"{e + 1} {deck.etype[e]} {c[e, 0]:.17g} {c[e, 1]:.17g} {c[e, 2]:.17g}\n")
            fh.write("\nNodes\n\n")
            for e in range(deck.nelements):
                for k in range(8):
                    p = deck.nodex[e, k]
                    fh.write(f"{e + 1} {k + 1} 1 {deck.ctype[e]} {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")
            fh.write("\nNode Velocities\n\n")
            for e in range(deck.nelements):
                for k in range(8):
                    p = deck.nodev[e, k]
                    fh.write(f"{e + 1} {k + 1} 1 {p[0]:.17g} {p[1]:.17g} {p[2]:.17g}\n")


def read_data(path: str, units: str = "metal") -> Deck:
    """Read a CAC data file written by :func:`write_data` (same section formats)."""
    with open(path) as fh:
        lines = [ln.split("#")[0].rstrip() for ln in fh]
    lo, hi = np.zeros(3), np.zeros(3)
    na = ne = 0
    i = 1
    while i < len(lines) and lines[i].strip() not in ("Masses", "Element Types", "Atoms", "Elements"):
        w = lines[i].split()
        if len(w) >= 2 and w[1] == "atoms":
            na = int(w[0])
        elif len(w) >= 2 and w[1] == "elements":
            ne = int(w[0])
        elif len(w) == 4 and w[2].endswith("lo"):
            d = "xyz".index(w[2][0])
            lo[d], hi[d] = float(w[0]), float(w[1])
        i += 1
    sections = {}
    cur = None
    for ln in lines[i:]:
        s = ln.strip()
        if not s:
            continue
        if s in ("Masses", "Element Types", "Atoms", "Velocities", "Elements", "Nodes", "Node Velocities"):
            cur = s
            sections[cur] = []
            continue
        sections[cur].append(s.split())
    masses = np.array([float(w[1]) for w in sorted(sections["Masses"], key=lambda w: int(w[0]))])
    et = sections.get("Element Types", [["1", "Hex", "1", "9", "9", "9", "2", "2", "2"]])[0]
    ncell = (int(et[3]), int(et[4]), int(et[5]))
    x = np.zeros((na, 3)); v = np.zeros((na, 3)); at = np.ones(na, dtype=np.int32)
    for w in sections.get("Atoms", []):
        k = int(w[0]) - 1
        at[k] = int(w[1]); x[k] = [float(w[2]), float(w[3]), float(w[4])]
    for w in sections.get("Velocities", []):
        v[int(w[0]) - 1] = [float(w[1]), float(w[2]), float(w[3])]
    nodex = np.zeros((ne, 8, 3)); nodev = np.zeros((ne, 8, 3))
    ety = np.ones(ne, dtype=np.int32); cty = np.ones(ne, dtype=np.int32)
    for w in sections.get("Elements", []):
        ety[int(w[0]) - 1] = int(w[1])
    for w in sections.get("Nodes", []):
        e, k = int(w[0]) - 1, int(w[1]) - 1
        cty[e] = int(w[3]); nodex[e, k] = [float(w[4]), float(w[5]), float(w[6])]
    for w in sections.get("Node Velocities", []):
        e, k = int(w[0]) - 1, int(w[1]) - 1
        nodev[e, k] = [float(w[3]), float(w[4]), float(w[5])]
    return Deck(units=units, boxlo=lo, boxhi=hi, masses=masses, x=x, v=v, atype=at, ncell=ncell,
                nodex=nodex, nodev=nodev, etype=ety, ctype=cty, name=path)
