"""ctypes binding of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker and as the timed CPU baseline.  The
product path never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_errmsg.restype = C.c_char_p
        L.orc_errmsg.argtypes = [vp]
        L.orc_set_box.argtypes = [vp, _dp, _dp]
        L.orc_set_masses.argtypes = [vp, C.c_int, _dp]
        L.orc_add_etype.argtypes = [vp] + [C.c_int] * 7
        L.orc_set_atoms.argtypes = [vp, C.c_int, _dp, _dp, _ip, C.c_void_p]
        L.orc_set_elements.argtypes = [vp, C.c_int, _ip, _ip, _dp, _dp, C.c_void_p]
        L.orc_pair_lj.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.orc_pair_eam_setfl.argtypes = [vp, C.c_char_p, C.c_void_p]
        L.orc_pair_eam_arrays.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_nodal_force_style.argtypes = [vp, C.c_int]
        L.orc_neighbor.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_setup.argtypes = [vp, C.c_int, C.c_int]
        L.orc_run.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_void_p]
        L.orc_build_neighbor.argtypes = [vp]
        L.orc_compute_forces.argtypes = [vp, C.c_int, C.c_int]
        L.orc_counts.argtypes = [vp] + [C.POINTER(C.c_int)] * 6
        L.orc_list_size.restype = C.c_long
        L.orc_list_size.argtypes = [vp]
        L.orc_list_export.restype = C.c_long
        L.orc_list_export.argtypes = [vp, _ip, _ip, _ip, _ip, _ip]
        L.orc_get_atoms.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_elements.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_get_rho.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.orc_get_energy.argtypes = [vp, C.POINTER(C.c_double), _dp]
        L.orc_get_peratom_energy.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.orc_get_peratom_virial.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.orc_thermo.argtypes = [vp] + [C.POINTER(C.c_double)] * 4
        L.orc_neigh_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_etype_tables.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                       _dp, _ip, _ip, _ip, _ip, _dp, _dp, C.POINTER(C.c_int), _ip, _ip, C.c_int, _dp]
        L.orc_mass_matrices.argtypes = [_dp, _dp]
        L.orc_eam_splines.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_interpolate.argtypes = [C.c_int, C.c_double, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleError(RuntimeError):
    pass


class Oracle:
    """One CPU model: deck + pair style + neighbor settings."""

    def __init__(self, deck, pair: str, skin: float, every: int = 1, delay: int = 10, check: int = 1,
                 oneatom: int = 2000, eam_tables=None, lj=None, nodal_force_style: int = 0):
        L = lib()
        self.L = L
        self.deck = deck
        self.h = L.orc_create(1 if deck.units == "metal" else 0)
        self._chk(L.orc_set_box(self.h, np.ascontiguousarray(deck.boxlo, dtype=np.float64),
                                np.ascontiguousarray(deck.boxhi, dtype=np.float64)))
        self._chk(L.orc_set_masses(self.h, len(deck.masses), np.ascontiguousarray(deck.masses, dtype=np.float64)))
        if deck.nelements:
            nc = deck.ncell
            self._chk(L.orc_add_etype(self.h, 1, nc[0], nc[1], nc[2], 2, 2, 2))
        self._x = np.ascontiguousarray(deck.x, dtype=np.float64)
        self._v = np.ascontiguousarray(deck.v, dtype=np.float64)
        self._t = np.ascontiguousarray(deck.atype, dtype=np.int32)
        self._chk(L.orc_set_atoms(self.h, deck.natoms, self._x, self._v, self._t, None))
        if deck.nelements:
            self._chk(L.orc_set_elements(self.h, deck.nelements, np.ascontiguousarray(deck.etype, dtype=np.int32),
                                         np.ascontiguousarray(deck.ctype, dtype=np.int32),
                                         np.ascontiguousarray(deck.nodex, dtype=np.float64),
                                         np.ascontiguousarray(deck.nodev, dtype=np.float64), None))
        if pair == "lj":
            p = lj or dict(eps=1.0, sigma=1.0, cut=2.5)
            self._chk(L.orc_pair_lj(self.h, p["eps"], p["sigma"], p["cut"]))
        elif pair == "eam":
            t = eam_tables
            self._chk(L.orc_pair_eam_arrays(self.h, t.nrho, t.drho, t.nr, t.dr, t.cut,
                                            np.ascontiguousarray(t.frho), np.ascontiguousarray(t.rhor),
                                            np.ascontiguousarray(t.z2r)))
        else:
            raise ValueError(pair)
        self._chk(L.orc_neighbor(self.h, skin, every, delay, check, oneatom))
        self._chk(L.orc_nodal_force_style(self.h, nodal_force_style))

    def _chk(self, rc):
        if rc:
            raise OracleError(f"oracle error {rc}: {self.L.orc_errmsg(self.h).decode()}")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # run control
    def setup(self, eflag=1, vflag=1):
        self._chk(self.L.orc_setup(self.h, eflag, vflag))

    def build_neighbor(self):
        self._chk(self.L.orc_build_neighbor(self.h))

    def compute_forces(self, eflag=1, vflag=1):
        self._chk(self.L.orc_compute_forces(self.h, eflag, vflag))

    def run(self, nsteps, dt, thermo_every=0):
        nrec = (nsteps // thermo_every + 2) if thermo_every else 2
        th = np.zeros((nrec + 1, 5))
        self._chk(self.L.orc_run(self.h, nsteps, dt, thermo_every, _ptr(th)))
        used = 1 + (nsteps // thermo_every if thermo_every else 0)
        if thermo_every == 0 or nsteps % thermo_every:
            used += 1
        return th[:used]

    # getters
    def counts(self):
        v = [C.c_int() for _ in range(6)]
        self.L.orc_counts(self.h, *[C.byref(x) for x in v])
        return dict(zip(["nalocal", "naghost", "nelocal", "neghost", "inum", "ainum"], [x.value for x in v]))

    def atoms(self):
        n = self.deck.natoms
        x, v, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self.L.orc_get_atoms(self.h, _ptr(x), _ptr(v), _ptr(f))
        return x, v, f

    def elements(self):
        n = self.deck.nelements
        a = [np.zeros((n, 8, 3)) for _ in range(4)]
        self.L.orc_get_elements(self.h, *[_ptr(x) for x in a])
        return dict(nodex=a[0], nodev=a[1], nodef=a[2], gaussf=a[3])

    def rho(self):
        ar, nr = np.zeros(self.deck.natoms), np.zeros((self.deck.nelements, 8))
        self.L.orc_get_rho(self.h, _ptr(ar), _ptr(nr))
        return ar, nr

    def energy(self):
        e = C.c_double()
        vir = np.zeros(6)
        self.L.orc_get_energy(self.h, C.byref(e), vir)
        return e.value, vir

    def peratom_energy(self):
        ea, en = np.zeros(self.deck.natoms), np.zeros((self.deck.nelements, 8))
        self.L.orc_get_peratom_energy(self.h, _ptr(ea), _ptr(en))
        return ea, en

    def peratom_virial(self):
        va, vn = np.zeros((self.deck.natoms, 6)), np.zeros((self.deck.nelements, 8, 6))
        self.L.orc_get_peratom_virial(self.h, _ptr(va), _ptr(vn))
        return va, vn

    def thermo(self):
        v = [C.c_double() for _ in range(4)]
        self.L.orc_thermo(self.h, *[C.byref(x) for x in v])
        return dict(zip(["pe", "ke", "temp", "press"], [x.value for x in v]))

    def neigh_stats(self):
        a, b = C.c_int(), C.c_int()
        self.L.orc_neigh_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def half_list(self):
        """Raw half list: arrays point, site kind, site owner, site ucell, image[3]."""
        n = self.L.orc_list_size(self.h)
        p, k, o, u = (np.zeros(n, dtype=np.int32) for _ in range(4))
        img = np.zeros((n, 3), dtype=np.int32)
        m = self.L.orc_list_export(self.h, p, k, o, u, img)
        assert m == n
        return p, k, o, u, img

    def eam_splines(self):
        a, b = C.c_int(), C.c_int()
        self.L.orc_eam_splines(self.h, C.byref(a), C.byref(b), None, None, None)
        fs, rs, zs = np.zeros((a.value + 1, 7)), np.zeros((b.value + 1, 7)), np.zeros((b.value + 1, 7))
        self.L.orc_eam_splines(self.h, None, None, _ptr(fs), _ptr(rs), _ptr(zs))
        return fs, rs, zs


def setfl_splines(path, element=None):
    """Read a setfl file with the oracle's reader for one atom type mapped to
    `element` (None: the file's first element) and return the spline tables it
    built: F(rho), rho(r), z2r, each [n+1][7]."""
    L = lib()
    h = L.orc_create(1)
    try:
        names = (C.c_char_p * 1)(element.encode()) if element is not None else None
        rc = L.orc_pair_eam_setfl(h, os.fsencode(path), names)
        if rc:
            raise OracleError(L.orc_errmsg(h).decode())
        a, b = C.c_int(), C.c_int()
        L.orc_eam_splines(h, C.byref(a), C.byref(b), None, None, None)
        fs, rs, zs = np.zeros((a.value + 1, 7)), np.zeros((b.value + 1, 7)), np.zeros((b.value + 1, 7))
        L.orc_eam_splines(h, None, None, _ptr(fs), _ptr(rs), _ptr(zs))
        return fs, rs, zs
    finally:
        L.orc_destroy(h)


def etype_tables(ncell=(9, 9, 9)):
    """Element-type tables of a Hex type with 2x2x2 Gauss cells."""
    L = lib()
    h = L.orc_create(1)
    try:
        rc = L.orc_add_etype(h, 1, ncell[0], ncell[1], ncell[2], 2, 2, 2)
        if rc:
            raise OracleError(L.orc_errmsg(h).decode())
        nu = ncell[0] * ncell[1] * ncell[2]
        nucell, ngcell, nsub = C.c_int(), C.c_int(), C.c_int()
        nw = C.c_double()
        shape = np.zeros((nu, 8)); g2u = np.zeros(8, dtype=np.int32); u2g = np.zeros(nu, dtype=np.int32)
        g2n = np.zeros(8, dtype=np.int32); n2g = np.zeros(8, dtype=np.int32); w = np.zeros(8)
        wsa = np.zeros((8, 8)); ncs = np.zeros(125, dtype=np.int32); us2u = np.full((125, 500), -1, dtype=np.int32)
        sacs = np.zeros((125, 8))
        L.orc_etype_tables(h, 1, C.byref(nucell), C.byref(ngcell), C.byref(nw), shape, g2u, u2g, g2n, n2g, w, wsa,
                           C.byref(nsub), ncs, us2u, 500, sacs)
        ns = nsub.value
        return dict(nucell=nucell.value, ngcell=ngcell.value, nodal_weight=nw.value, shape=shape, g2u=g2u, u2g=u2g,
                    g2n=g2n, n2g=n2g, weight=w, wsa=wsa, nsub=ns, nucell_sub=ncs[:ns], us2u=us2u[:ns], sacs=sacs[:ns])
    finally:
        L.orc_destroy(h)


def mass_matrices():
    m, mi = np.zeros((8, 8)), np.zeros((8, 8))
    lib().orc_mass_matrices(m, mi)
    return m, mi


def interpolate(f, delta):
    """PairEAM::interpolate on f[1..n] (f[0] unused); returns spline[n+1][7]."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    n = f.size - 1
    s = np.zeros((n + 1, 7))
    lib().orc_interpolate(n, delta, f, s)
    return s


def _full_rows(deck, point, skind, sowner, sucell, img, g2u):
    """Half-list rows (point, kind, owner, ucell, ix, iy, iz) plus, for every
    pair whose site is itself a point (an atom, or a virtual atom on a Gauss
    cell's unit cell), the mirrored row in that point's set with the opposite
    image; points are atoms (id a) and Gauss cells (id A + 8e + g)."""
    A = deck.natoms
    g2u = np.asarray(g2u, dtype=np.int64)
    keys = np.stack([point, skind, sowner, sucell, img[:, 0], img[:, 1], img[:, 2]], axis=1).astype(np.int64)
    k, o, u = keys[:, 1], keys[:, 2], keys[:, 3]
    u2g = np.full(int(g2u.max()) + 1 if g2u.size else 1, -1, dtype=np.int64)
    u2g[g2u] = np.arange(g2u.size)
    g = np.where((k == 1) & (u < u2g.size), u2g[np.clip(u, 0, u2g.size - 1)], -1)
    site_is_point = (k == 0) | (g >= 0)
    q = np.where(k == 0, o, A + 8 * o + g)
    p = keys[:, 0]
    e, gp = np.divmod(np.maximum(p - A, 0), 8)
    pk = (p >= A).astype(np.int64)
    po = np.where(p < A, p, e)
    pu = np.where(p < A, 0, g2u[gp] if g2u.size else 0)
    extra = np.stack([q, pk, po, pu, -keys[:, 4], -keys[:, 5], -keys[:, 6]], axis=1)[site_is_point]
    return np.concatenate([keys, extra], axis=0)


def canonical_full_sets(deck, point, skind, sowner, sucell, img, g2u):
    """Expand a half list into per-point full sets: rows (point, kind, owner,
    ucell, ix, iy, iz), sorted (see _full_rows for the mirrored pairs)."""
    allk = _full_rows(deck, point, skind, sowner, sucell, img, g2u)
    order = np.lexsort(tuple(allk[:, c] for c in range(6, -1, -1)))
    return allk[order]


def pack_set_rows(deck, rows):
    """Rows (point, kind, owner, ucell, ix, iy, iz) as one int64 each, in the
    rows' lexicographic order (images -1..1 per axis): sorting the packed keys
    sorts the rows, at a fraction of a lexsort's cost on tens of millions."""
    rows = np.asarray(rows, dtype=np.int64)
    P = deck.natoms + 8 * deck.nelements
    nucell = int(np.prod(deck.ncell)) if deck.nelements else 1
    widths = [P, 2, max(deck.natoms, deck.nelements, 1), max(nucell, 1), 3, 3, 3]
    total = 1
    for w in widths:
        total *= w
    if total >= 2 ** 63:
        raise ValueError("set rows do not fit one int64 key")
    key = np.zeros(rows.shape[0], dtype=np.int64)
    for c, w in enumerate(widths):
        v = rows[:, c] + (1 if c >= 4 else 0)
        if v.size and (v.min() < 0 or v.max() >= w):
            raise ValueError(f"set row column {c} outside 0..{w - 1}")
        key = key * w + v
    return key


def canonical_full_keys(deck, point, skind, sowner, sucell, img, g2u):
    """canonical_full_sets as sorted packed keys (pack_set_rows)."""
    return np.sort(pack_set_rows(deck, _full_rows(deck, point, skind, sowner, sucell, img, g2u)))
