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


def _splines(tab):
    from oracle import oracle as O
    rs = O.interpolate(np.concatenate([[0.0], tab.rhor]), tab.dr)
    zs = O.interpolate(np.concatenate([[0.0], tab.z2r]), tab.dr)
    fs = O.interpolate(np.concatenate([[0.0], tab.frho]), tab.drho)
    return rs, zs, fs


def _knot_r(r, tab):
    p = r / tab.dr + 1.0
    m = np.minimum(p.astype(np.int64), tab.nr - 1)
    return m, np.minimum(p - m, 1.0)


def _embed(fs, rho, tab):
    """F(rho) with the linear extension above rhomax, and F'(rho)."""
    p = rho / tab.drho + 1.0
    m = np.clip(p.astype(np.int64), 1, tab.nrho - 1)
    p = np.minimum(p - m, 1.0)
    F, Fp = spline_eval(fs, m, p)
    rhomax = (tab.nrho - 1) * tab.drho
    return np.where(rho > rhomax, F + Fp * (rho - rhomax), F), Fp


def cac_eam(deck, tab, images=1, peratom=False):
    """Composite CAC EAM by all-pairs enumeration, from the definitions alone:
    the density of every point (atom or Gauss point) sums f(r) over every site
    image within the cutoff; an interpolated atom's density is the
    shape-function interpolation of its element's nodal densities; the force
    on a point sums -(F'_i + F'_j) rho'(r) - phi'(r), divided by r, times the
    pair vector, over every site image within the cutoff; the energy counts
    atoms once and Gauss points nodal_weight times, each with half of its
    pair energies.  No lists, bins, ghosts or half-list bookkeeping.

    Returns atom forces [A,3], Gauss forces [E,8 Gauss cells,3], atom
    densities [A], nodal densities [E,8 nodes], energy; with peratom, also
    each point's own energy F(rho) + 1/2 sum phi and virial 1/2 sum d d fpair
    (atoms [A] / [A,6], nodes [E,8] / [E,8,6] in node order)."""
    pos, kind, owner, ucell = all_sites(deck)
    L = deck.boxhi - deck.boxlo
    A, E = deck.natoms, deck.nelements
    nu = int(np.prod(deck.ncell)) if E else 1
    Nsh = shape_hex(ucell_natural(deck.ncell)) if E else np.zeros((1, 8))
    g2u = corner_ucells(deck.ncell) if E else np.zeros(8, np.int64)
    g2n = np.array([int(np.argmax(Nsh[u])) for u in g2u]) if E else np.arange(8)
    psite = np.concatenate([np.arange(A), A + (np.arange(E)[:, None] * nu + g2u[None, :]).ravel()])
    rs, zs, fs = _splines(tab)
    rng = np.arange(-images, images + 1)
    shifts = [np.array((a, b, c)) for a in rng for b in rng for c in rng]
    cut2 = tab.cut * tab.cut

    def pairs(p):
        out = []
        for s in shifts:
            d = pos[psite[p]] - (pos + s * L)
            r2 = np.einsum("ij,ij->i", d, d)
            hit = r2 < cut2
            if not np.any(s):
                hit[psite[p]] = False
            idx = np.nonzero(hit)[0]
            out.append((idx, d[idx], np.sqrt(r2[idx])))
        idx = np.concatenate([o[0] for o in out])
        return idx, np.concatenate([o[1] for o in out]), np.concatenate([o[2] for o in out])

    P = len(psite)
    pr = [pairs(p) for p in range(P)]
    rho = np.zeros(P)
    for p, (idx, d, r) in enumerate(pr):
        m, q = _knot_r(r, tab)
        rho[p] = spline_eval(rs, m, q)[0].sum()
    rho_nodes = np.zeros((E, 8))
    if E:
        rho_g = rho[A:].reshape(E, 8)
        rho_nodes[:, g2n] = rho_g
    site_rho = np.concatenate([rho[:A], np.einsum("uk,ek->eu", Nsh, rho_nodes).reshape(-1)]) if E else rho[:A]
    _, site_fp = _embed(fs, site_rho, tab)
    Fi, fp_i = _embed(fs, rho, tab)
    f = np.zeros((P, 3))
    phisum = np.zeros(P)
    vir = np.zeros((P, 6))
    for p, (idx, d, r) in enumerate(pr):
        m, q = _knot_r(r, tab)
        _, rhop = spline_eval(rs, m, q)
        z2, z2p = spline_eval(zs, m, q)
        phi = z2 / r
        phip = z2p / r - phi / r
        fpair = -((fp_i[p] + site_fp[idx]) * rhop + phip) / r
        f[p] = (d * fpair[:, None]).sum(0)
        phisum[p] = phi.sum()
        vir[p] = 0.5 * np.stack([d[:, 0] * d[:, 0], d[:, 1] * d[:, 1], d[:, 2] * d[:, 2], d[:, 0] * d[:, 1],
                                 d[:, 0] * d[:, 2], d[:, 1] * d[:, 2]], 1).T @ fpair
    nw = nu / 8.0
    energy = float(np.sum(Fi[:A] + 0.5 * phisum[:A]) + nw * np.sum(Fi[A:] + 0.5 * phisum[A:]))
    if peratom:
        ep = Fi + 0.5 * phisum
        enode, vnode = np.zeros((E, 8)), np.zeros((E, 8, 6))
        if E:
            enode[:, g2n] = ep[A:].reshape(E, 8)
            vnode[:, g2n, :] = vir[A:].reshape(E, 8, 6)
        return f[:A], f[A:].reshape(E, 8, 3), rho[:A], rho_nodes, energy, dict(
            eatom=ep[:A], enode=enode, vatom=vir[:A], vnode=vnode)
    return f[:A], f[A:].reshape(E, 8, 3), rho[:A], rho_nodes, energy
