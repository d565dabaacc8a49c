"""Brute-force references for the tests: all-pairs enumeration of points and
sites with explicit periodic images, no bins, no ghosts, no half lists."""

import numpy as np

from cac_amd import decks


def shape_hex(p):
    """Trilinear Hex shape functions at natural coordinates p[...,3] (Hex node order)."""
    s = decks.HEX_SIGNS.astype(np.float64)
    return 0.125 * np.prod(1.0 + p[..., None, :] * s, axis=-1)


def ucell_natural(ncell):
    g = [np.linspace(-1.0, 1.0, n) for n in ncell]
    pi, pj, pk = np.meshgrid(g[0], g[1], g[2], indexing="ij")
    return np.stack([pi.ravel(), pj.ravel(), pk.ravel()], axis=1)


def all_sites(deck):
    """Sites: atoms (kind 0) then every virtual atom of every element (kind 1).
    Returns pos[S,3], kind[S], owner[S], ucell[S]."""
    N = shape_hex(ucell_natural(deck.ncell))  # [nu,8]
    vpos = np.einsum("uk,ekd->eud", N, deck.nodex).reshape(-1, 3)
    nu = N.shape[0]
    E = deck.nelements
    pos = np.concatenate([deck.x, vpos], axis=0)
    kind = np.concatenate([np.zeros(deck.natoms, np.int64), np.ones(E * nu, np.int64)])
    owner = np.concatenate([np.arange(deck.natoms), np.repeat(np.arange(E), nu)])
    ucell = np.concatenate([np.zeros(deck.natoms, np.int64), np.tile(np.arange(nu), E)])
    return pos, kind, owner, ucell


def corner_ucells(ncell):
    nx, ny, nz = ncell
    out = []
    for i in (0, nx - 1):
        for j in (0, ny - 1):
            for k in (0, nz - 1):
                out.append(i * ny * nz + j * nz + k)
    return np.array(out)


def full_sets(deck, cut, images=1):
    """Canonical full neighbor set: rows (point, kind, owner, ucell, ix, iy, iz)
    for every point (atoms, then Gauss cells 8e+g) and every site image within
    cut (r <= cut), self excluded, sorted lexicographically."""
    pos, kind, owner, ucell = all_sites(deck)
    L = deck.boxhi - deck.boxlo
    g2u = corner_ucells(deck.ncell)
    nu = int(np.prod(deck.ncell))
    A = deck.natoms
    ppos = np.concatenate([deck.x, pos[A:].reshape(-1, nu, 3)[:, g2u, :].reshape(-1, 3)], axis=0) if deck.nelements else deck.x
    psite = np.concatenate([np.arange(A), A + (np.arange(deck.nelements)[:, None] * nu + g2u[None, :]).ravel()])
    rng = np.arange(-images, images + 1)
    shifts = np.array([(a, b, c) for a in rng for b in rng for c in rng])
    rows = []
    cut2 = cut * cut
    for p in range(ppos.shape[0]):
        for s in shifts:
            d = ppos[p] - (pos + s * L)
            r2 = np.einsum("ij,ij->i", d, d)
            hit = np.nonzero(r2 <= cut2)[0]
            if not np.any(s):
                hit = hit[hit != psite[p]]
            for h in hit:
                rows.append((p, kind[h], owner[h], ucell[h], s[0], s[1], s[2]))
    arr = np.array(rows, dtype=np.int64).reshape(-1, 7)
    order = np.lexsort(tuple(arr[:, c] for c in range(6, -1, -1)))
    return arr[order]


def spline_eval(spl, m, p):
    c = spl[m]
    val = ((c[..., 3] * p + c[..., 4]) * p + c[..., 5]) * p + c[..., 6]
    der = (c[..., 0] * p + c[..., 1]) * p + c[..., 2]
    return val, der


def eam_atoms_energy(x, L, tab, analytic=False):
    """EAM energy of a periodic all-atom configuration, all pairs with minimum
    image (box >= 2 cut).  analytic=True evaluates the defining functions
    instead of the tabulated splines."""
    from oracle import oracle as O
    n = x.shape[0]
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
    np.fill_diagonal(r, np.inf)
    mask = r < tab.cut
    if analytic:
        f = np.where(mask, decks.cu_f(np.where(mask, r, 1.0)), 0.0)
        rho = f.sum(1)
        phi = np.where(mask, decks.cu_phi(np.where(mask, r, 1.0)), 0.0)
        return float(np.sum(decks.cu_F(rho)) + 0.5 * phi.sum())
    rs = O.interpolate(np.concatenate([[0.0], tab.rhor]), tab.dr)
    zs = O.interpolate(np.concatenate([[0.0], tab.z2r]), tab.dr)
    fs = O.interpolate(np.concatenate([[0.0], tab.frho]), tab.drho)
    rr = np.where(mask, r, 1.0)
    p = rr / tab.dr + 1.0
    m = np.minimum(p.astype(np.int64), tab.nr - 1)
    p = np.minimum(p - m, 1.0)
    fv, _ = spline_eval(rs, m, p)
    zv, _ = spline_eval(zs, m, p)
    rho = np.where(mask, fv, 0.0).sum(1)
    phi = np.where(mask, zv / rr, 0.0)
    pr = rho / tab.drho + 1.0
    mr = np.clip(pr.astype(np.int64), 1, tab.nrho - 1)
    pr = np.minimum(pr - mr, 1.0)
    Fv, Fp = spline_eval(fs, mr, pr)
    rhomax = (tab.nrho - 1) * tab.drho
    Fv = np.where(rho > rhomax, Fv + Fp * (rho - rhomax), Fv)
    return float(np.sum(Fv) + 0.5 * phi.sum())
