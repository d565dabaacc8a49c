"""A CAC deck whose neighbor-list cutoff falls between two roundings of one
pair distance.

The device scans site records that hold each site wrapped into the box (an
interpolated atom as sum_k N_k x_k + W L) and forms row - (site + t L).  The
reference subtracts N_k (x_k + T L) node by node from the row point's own
position (npair_half_bin_newton.cpp:194-199, 359-364).  The two forms round
differently; for a pair whose two values straddle (rc + skin)^2 only the
reference's form gives the reference's list.  This module emulates both forms
in IEEE double (numpy elementwise operations, no fused multiply-add), finds
such a pair on a small deck, and picks the skin that puts the cutoff between
the two values.
"""

from __future__ import annotations

import numpy as np

from cac_amd import decks
from oracle import oracle as O

TAB = decks.cu_eam_tables()


def _shape_sum(shape_u, nodes):
    """sum_k N_k x_k from 0, node by node (the device's refresh arithmetic)."""
    s = np.zeros(nodes.shape[:-2] + (3,))
    for k in range(8):
        s = s + shape_u[..., k, None] * nodes[..., k, :]
    return s


def _ref_sub(row, shape_u, nodes, shift):
    """row - N_k (x_k + shift), subtracted node by node (the reference's form)."""
    d = row.copy()
    for k in range(8):
        d = d - shape_u[..., k, None] * (nodes[..., k, :] + shift)
    return d


def _rsq(d):
    return d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]


def split_case(seed=5, skin_target=1.0, box=(3, 3, 3), core=8, ncell=3):
    """Return (deck, skin, info): the skin puts (rc + skin)^2 between the
    device's fast form and the reference's form of the squared distance of one
    (point, interpolated atom, image) pair; info names the pair and both
    values."""
    deck = decks.cac_eam(box_cells=box, core_cells=core, ncell=ncell, jitter=0.05, T=0.0, seed=seed)
    cut = TAB.cut
    # every pair within a slightly larger cutoff, from the oracle's lists
    o = O.Oracle(deck, "eam", skin=skin_target + 0.05, eam_tables=TAB)
    o.build_neighbor()
    x = o.atoms()[0]
    nodes = o.elements()["nodex"]
    t = O.etype_tables(deck.ncell)
    shape, g2u, g2n, u2g = t["shape"], t["g2u"], t["g2n"], t["u2g"]
    p, k, ow, u, img = o.half_list()
    keys = O.canonical_full_sets(deck, p, k, ow, u, img, g2u)
    A = deck.natoms
    lo, L = deck.boxlo, deck.boxhi - deck.boxlo
    virt = (keys[:, 1] == 1) & (u2g[keys[:, 3]] < 0)
    keys = keys[virt]
    pt, e, uc, T = keys[:, 0], keys[:, 2], keys[:, 3], keys[:, 4:7].astype(np.float64)
    # row positions: an atom, or the node a Gauss cell sits on
    row = np.empty((len(keys), 3))
    isa = pt < A
    row[isa] = x[pt[isa]]
    ge, gg = np.divmod(pt[~isa] - A, 8)
    row[~isa] = nodes[ge, g2n[gg]]
    N = shape[uc]
    raw = _shape_sum(N, nodes[e])
    w_row = -np.floor((row - lo) / L)
    w_site = -np.floor((raw - lo) / L)
    rec_row = row + w_row * L
    rec_site = raw + w_site * L
    t_img = T - w_site + w_row
    fast = _rsq(rec_row - (rec_site + t_img * L))
    ref = _rsq(_ref_sub(row, N, nodes[e], T * L))
    want = (cut + skin_target) ** 2
    order = np.argsort(np.abs(ref - want))
    for i in order[:20000]:
        if fast[i] == ref[i]:
            continue
        a, b = sorted((float(fast[i]), float(ref[i])))
        # smallest skin whose (cut + skin)^2 reaches a, by bisection over doubles
        s_lo, s_hi = np.sqrt(a) - cut - 1e-9, np.sqrt(a) - cut + 1e-9
        for _ in range(200):
            mid = 0.5 * (s_lo + s_hi)
            if mid in (s_lo, s_hi):
                break
            cn = cut + mid
            if cn * cn >= a:
                s_hi = mid
            else:
                s_lo = mid
        s = s_hi
        cn = cut + s
        c = cn * cn
        if a <= c < b:
            info = dict(point=int(pt[i]), element=int(e[i]), ucell=int(uc[i]), image=tuple(int(v) for v in T[i]),
                        fast=float(fast[i]), ref=float(ref[i]), cutsq=c, ref_keeps=bool(ref[i] <= c))
            return deck, float(s), info
    raise RuntimeError("no pair whose two roundings straddle a reachable cutoff")
