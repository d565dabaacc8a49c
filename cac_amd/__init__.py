"""cac_amd -- MI355X-native CAC force + neighbor hot path.

Python binding of the C ABI in include/cac_hip.h (libcachip.so, built in-tree
by cac_amd/build.py).  :class:`Model` holds one CAC model resident in HBM and
exposes the reference's phases: neighbor build, Pair::compute, FixNVE and the
Verlet loop.  There is no CPU fallback: without the HIP library every call
raises.
"""

from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBPATH = os.path.join(HERE, "libcachip.so")
HEADER = os.path.join(ROOT, "include", "cac_hip.h")

FP64, MIXED = 0, 1
UNITS = {"lj": 0, "metal": 1}
ERRORS = {1: "EARG", 2: "EOVERFLOW", 3: "ENAN", 4: "EDEFORM", 5: "EOOM", 6: "EHIP", 7: "ESTATE", 8: "ENODEV"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Request(C.Structure):
    """cac_hip_request of include/cac_hip.h"""
    _fields_ = [("units", C.c_int), ("lo", C.c_double * 3), ("hi", C.c_double * 3), ("ntypes", C.c_int),
                ("mass", C.c_void_p), ("ncell", C.c_int * 3), ("natoms", C.c_int), ("x", C.c_void_p),
                ("v", C.c_void_p), ("type", C.c_void_p), ("nelements", C.c_int), ("ctype", C.c_void_p),
                ("nodex", C.c_void_p), ("nodev", C.c_void_p), ("skin", C.c_double), ("precision", C.c_int),
                ("nsteps", C.c_int), ("dt", C.c_double), ("keep_state", C.c_int)]


class _Result(C.Structure):
    """cac_hip_result of include/cac_hip.h"""
    _fields_ = [("status", C.c_int), ("nsteps", C.c_int), ("atom_equivalents", C.c_long),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("step_ms_mean", C.c_double), ("pe", C.c_double), ("ke", C.c_double), ("etotal", C.c_double),
                ("builds", C.c_long), ("device_step_ms_mean", C.c_double)]


class CacHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cac_hip error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def header_functions(path: str = HEADER):
    """Names of every function the C ABI header declares."""
    txt = open(path).read()
    return sorted(set(re.findall(r"\b(cac_hip_[a-z_0-9]+)\s*\(", txt)))


_lib = None


def lib():
    """Load libcachip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        L = C.CDLL(LIBPATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "cac_hip_create": [C.POINTER(vp), i, i],
            "cac_hip_destroy": [vp],
            "cac_hip_set_box": [vp, _dp, _dp],
            "cac_hip_set_masses": [vp, i, _dp],
            "cac_hip_add_etype": [vp, i, i, i, i, i, i, i],
            "cac_hip_upload_atoms": [vp, i, vp, vp, vp],
            "cac_hip_upload_elements": [vp, i, vp, vp, vp, vp],
            "cac_hip_pair_lj": [vp, d, d, d],
            "cac_hip_pair_eam": [vp, i, d, i, d, d, _dp, _dp, _dp],
            "cac_hip_pair_eam_setfl": [vp, C.c_char_p],
            "cac_hip_set_precision": [vp, i],
            "cac_hip_set_nodal_force_style": [vp, i],
            "cac_hip_neighbor_settings": [vp, d, i, i, i, i],
            "cac_hip_neigh_build": [vp],
            "cac_hip_pbc": [vp],
            "cac_hip_neigh_decide": [vp, C.POINTER(i)],
            "cac_hip_neigh_info": [vp, C.POINTER(i), C.POINTER(C.c_long), C.POINTER(i), C.POINTER(i)],
            "cac_hip_neigh_export": [vp, C.POINTER(C.c_long), vp, vp, vp, vp, vp],
            "cac_hip_compute": [vp, i, i],
            "cac_hip_get_energy": [vp, C.POINTER(d), _dp],
            "cac_hip_get_peratom": [vp, vp, vp, vp, vp],
            "cac_hip_nve_initial": [vp, d],
            "cac_hip_nve_final": [vp, d],
            "cac_hip_setup": [vp, i, i],
            "cac_hip_run": [vp, i, d, i, vp],
            "cac_hip_thermo": [vp, C.POINTER(d), C.POINTER(d), C.POINTER(d), C.POINTER(d)],
            "cac_hip_thermo_noalloc": [vp, C.POINTER(d), C.POINTER(d), C.POINTER(d), C.POINTER(d)],
            "cac_hip_download_atoms": [vp, vp, vp, vp],
            "cac_hip_download_elements": [vp, vp, vp, vp, vp],
            "cac_hip_download_rho": [vp, vp, vp],
            "cac_hip_counts": [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)],
            "cac_hip_sync": [vp],
            "cac_hip_enable_timing": [vp, i],
            "cac_hip_kernel_time": [vp, C.c_char_p, C.POINTER(d), C.POINTER(C.c_long)],
            "cac_hip_neigh_stats": [vp, C.POINTER(i), C.POINTER(i)],
            "cac_hip_inner_stats": [vp, C.POINTER(C.c_long), C.POINTER(d), C.POINTER(d), C.POINTER(i)],
            "cac_hip_set_inner_skin": [vp, d],
            "cac_hip_sync_positions": [vp, vp, vp],
            "cac_hip_download_forces": [vp, vp, vp, vp],
            "cac_hip_server_create": [C.POINTER(vp), i, i, d],
            "cac_hip_server_destroy": [vp],
            "cac_hip_server_pair_eam": [vp, i, d, i, d, d, _dp, _dp, _dp],
            "cac_hip_server_pair_lj": [vp, d, d, d],
            "cac_hip_server_submit": [vp, C.POINTER(_Request), C.POINTER(C.c_long)],
            "cac_hip_server_drain": [vp],
            "cac_hip_server_result": [vp, C.c_long, C.POINTER(_Result)],
            "cac_hip_server_state": [vp, C.c_long, vp, vp],
            "cac_hip_server_stats": [vp, C.POINTER(d), C.POINTER(d), C.POINTER(C.c_long), C.POINTER(d),
                                     C.POINTER(i), C.POINTER(d)],
            "cac_hip_server_segments": [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)],
            "cac_hip_server_device_intervals": [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                C.POINTER(C.c_long)],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.cac_hip_last_error.argtypes = [vp]
        L.cac_hip_last_error.restype = C.c_char_p
        L.cac_hip_server_last_error.argtypes = [vp]
        L.cac_hip_server_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Model:
    """One CAC model on one MI355X (a cac_hip_ctx)."""

    def __init__(self, deck, pair: str, skin: float, every: int = 1, delay: int = 10, check: int = 1,
                 one: int = 2000, eam_tables=None, lj=None, precision: int = FP64, nodal_force_style: int = 0,
                 device: int = 0):
        self.L = lib()
        self.deck = deck
        h = C.c_void_p()
        rc = self.L.cac_hip_create(C.byref(h), device, UNITS[deck.units])
        if rc:
            raise CacHipError(rc, "cac_hip_create failed (no HIP device?)")
        self.h = h
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self._chk(self.L.cac_hip_set_box(h, f64(deck.boxlo), f64(deck.boxhi)))
        self._chk(self.L.cac_hip_set_masses(h, len(deck.masses), f64(deck.masses)))
        if deck.nelements:
            nc = deck.ncell
            self._chk(self.L.cac_hip_add_etype(h, 1, nc[0], nc[1], nc[2], 2, 2, 2))
        x, v, t = f64(deck.x), f64(deck.v), i32(deck.atype)
        self._chk(self.L.cac_hip_upload_atoms(h, deck.natoms, _p(x), _p(v), _p(t)))
        if deck.nelements:
            et, ct, nx, nv = i32(deck.etype), i32(deck.ctype), f64(deck.nodex), f64(deck.nodev)
            self._chk(self.L.cac_hip_upload_elements(h, deck.nelements, _p(et), _p(ct), _p(nx), _p(nv)))
        if pair == "lj":
            p = lj or dict(eps=1.0, sigma=1.0, cut=2.5)
            self._chk(self.L.cac_hip_pair_lj(h, p["eps"], p["sigma"], p["cut"]))
        elif pair == "eam":
            tb = eam_tables
            self._chk(self.L.cac_hip_pair_eam(h, tb.nrho, tb.drho, tb.nr, tb.dr, tb.cut, f64(tb.frho), f64(tb.rhor),
                                              f64(tb.z2r)))
        else:
            raise ValueError(pair)
        self._chk(self.L.cac_hip_set_precision(h, precision))
        self._chk(self.L.cac_hip_set_nodal_force_style(h, nodal_force_style))
        self._chk(self.L.cac_hip_neighbor_settings(h, skin, every, delay, check, one))

    def _chk(self, rc):
        if rc:
            raise CacHipError(rc, self.L.cac_hip_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.cac_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # phases
    def set_precision(self, mode):
        self._chk(self.L.cac_hip_set_precision(self.h, mode))

    def build_neighbor(self):
        self._chk(self.L.cac_hip_neigh_build(self.h))

    def pbc(self):
        """Domain::pbc on the device model, as write_data runs it; the lists
        are stale until the next setup() or build_neighbor()."""
        self._chk(self.L.cac_hip_pbc(self.h))

    def compute(self, eflag=1, vflag=1):
        self._chk(self.L.cac_hip_compute(self.h, eflag, vflag))

    def setup(self, eflag=1, vflag=1):
        self._chk(self.L.cac_hip_setup(self.h, eflag, vflag))

    def run(self, nsteps, dt, thermo_every=0, thermo=True):
        nrec = nsteps // max(thermo_every, 1) + 3
        th = np.zeros((nrec, 5)) if thermo else None
        self._chk(self.L.cac_hip_run(self.h, nsteps, dt, thermo_every, _p(th)))
        if not thermo:
            return None
        used = 1 + (nsteps // thermo_every if thermo_every else 0)
        if thermo_every == 0 or nsteps % thermo_every:
            used += 1
        return th[:used]

    def sync(self):
        self._chk(self.L.cac_hip_sync(self.h))

    # data
    def atoms(self):
        n = self.deck.natoms
        x, v, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._chk(self.L.cac_hip_download_atoms(self.h, _p(x), _p(v), _p(f)))
        return x, v, f

    def elements(self):
        n = self.deck.nelements
        a = [np.zeros((n, 8, 3)) for _ in range(4)]
        self._chk(self.L.cac_hip_download_elements(self.h, *[_p(x) for x in a]))
        return dict(nodex=a[0], nodev=a[1], nodef=a[2], gaussf=a[3])

    def rho(self):
        ar, nr = np.zeros(self.deck.natoms), np.zeros((self.deck.nelements, 8))
        self._chk(self.L.cac_hip_download_rho(self.h, _p(ar), _p(nr)))
        return ar, nr

    def energy(self):
        e = C.c_double()
        vir = np.zeros(6)
        self._chk(self.L.cac_hip_get_energy(self.h, C.byref(e), vir))
        return e.value, vir

    def peratom(self, energy=True, virial=True):
        """Per-point tallies of the last compute: eatom [A], enode [E,8],
        vatom [A,6], vnode [E,8,6] (None where not asked for)."""
        A, E = self.deck.natoms, self.deck.nelements
        ea, en = (np.zeros(A), np.zeros((E, 8))) if energy else (None, None)
        va, vn = (np.zeros((A, 6)), np.zeros((E, 8, 6))) if virial else (None, None)
        self._chk(self.L.cac_hip_get_peratom(self.h, _p(ea), _p(en), _p(va), _p(vn)))
        return dict(eatom=ea, enode=en, vatom=va, vnode=vn)

    def thermo(self, noalloc=False):
        """PE, KE, temperature, pressure; noalloc: through the entry point
        that makes no device allocation."""
        v = [C.c_double() for _ in range(4)]
        fn = self.L.cac_hip_thermo_noalloc if noalloc else self.L.cac_hip_thermo
        self._chk(fn(self.h, *[C.byref(x) for x in v]))
        return dict(zip(["pe", "ke", "temp", "press"], [x.value for x in v]))

    def neigh_info(self):
        r, s, cap = C.c_int(), C.c_int(), C.c_int()
        ent = C.c_long()
        self._chk(self.L.cac_hip_neigh_info(self.h, C.byref(r), C.byref(ent), C.byref(s), C.byref(cap)))
        return dict(rows=r.value, entries=ent.value, sites=s.value, capacity=cap.value)

    def neigh_stats(self):
        a, b = C.c_int(), C.c_int()
        self._chk(self.L.cac_hip_neigh_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def inner_stats(self):
        """(prunes of the inner list, mean inner entries per point, delta, sites the inner lists reference)"""
        n, m, dl, lv = C.c_long(), C.c_double(), C.c_double(), C.c_int()
        self._chk(self.L.cac_hip_inner_stats(self.h, C.byref(n), C.byref(m), C.byref(dl), C.byref(lv)))
        return n.value, m.value, dl.value, lv.value

    def set_inner_skin(self, delta):
        self._chk(self.L.cac_hip_set_inner_skin(self.h, float(delta)))

    def sync_positions(self, x, nodex=None):
        """Copy host positions in ([A,3], [E,8,3]) without touching the lists."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        nx = np.ascontiguousarray(nodex, dtype=np.float64) if nodex is not None else None
        self._chk(self.L.cac_hip_sync_positions(self.h, x.ctypes.data if x.size else None,
                                                nx.ctypes.data if nx is not None and nx.size else None))

    def decide(self):
        """Neighbor::decide on the current positions: True when a rebuild is due."""
        f = C.c_int()
        self._chk(self.L.cac_hip_neigh_decide(self.h, C.byref(f)))
        return bool(f.value)

    def forces(self):
        """(atom forces [A,3], nodal forces [E,8,3], Gauss-cell forces [E,8,3])"""
        A, E = self.deck.natoms, self.deck.nelements
        f, nf, gf = np.zeros((A, 3)), np.zeros((E, 8, 3)), np.zeros((E, 8, 3))
        self._chk(self.L.cac_hip_download_forces(self.h, f.ctypes.data if A else None,
                                                 nf.ctypes.data if E else None, gf.ctypes.data if E else None))
        return f, nf, gf

    def full_sets(self, sort=True):
        """Full lists as rows (point, kind, owner, ucell, ix, iy, iz), sorted
        unless sort is False."""
        n = C.c_long()
        self._chk(self.L.cac_hip_neigh_export(self.h, C.byref(n), None, None, None, None, None))
        m = n.value
        row, kind, owner, ucell = (np.zeros(m, dtype=np.int32) for _ in range(4))
        img = np.zeros((m, 3), dtype=np.int32)
        self._chk(self.L.cac_hip_neigh_export(self.h, C.byref(n), _p(row), _p(kind), _p(owner), _p(ucell), _p(img)))
        arr = np.stack([row, kind, owner, ucell, img[:, 0], img[:, 1], img[:, 2]], axis=1).astype(np.int64)
        if not sort:
            return arr
        order = np.lexsort(tuple(arr[:, c] for c in range(6, -1, -1)))
        return arr[order]

    def enable_timing(self, on=True):
        """True / 1: time every kernel section; 2: only the pair force pass; False / 0: off."""
        self._chk(self.L.cac_hip_enable_timing(self.h, int(on)))

    def kernel_time(self, name):
        a, n = C.c_double(), C.c_long()
        self._chk(self.L.cac_hip_kernel_time(self.h, name.encode(), C.byref(a), C.byref(n)))
        return a.value, n.value


class Server:
    """The resident serving loop (cac_hip_server): independent replicas
    admitted by slot and HBM budget and stepped round-robin on one GPU."""

    def __init__(self, slots: int = 16, hbm_budget_bytes: float = 0.0, pair: str = "eam", eam_tables=None,
                 lj=None, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.cac_hip_server_create(C.byref(h), device, slots, hbm_budget_bytes)
        if rc:
            raise CacHipError(rc, "cac_hip_server_create failed")
        self.h = h
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        if pair == "eam":
            tb = eam_tables
            self._chk(self.L.cac_hip_server_pair_eam(h, tb.nrho, tb.drho, tb.nr, tb.dr, tb.cut, f64(tb.frho),
                                                     f64(tb.rhor), f64(tb.z2r)))
        elif pair == "lj":
            p = lj or dict(eps=1.0, sigma=1.0, cut=2.5)
            self._chk(self.L.cac_hip_server_pair_lj(h, p["eps"], p["sigma"], p["cut"]))
        else:
            raise ValueError(pair)

    def _chk(self, rc):
        if rc:
            raise CacHipError(rc, self.L.cac_hip_server_last_error(self.h).decode())

    def submit(self, deck, nsteps: int, dt: float, skin: float = 2.0, precision: int = MIXED,
               keep_state: bool = False) -> int:
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        keep = [f64(deck.masses), f64(deck.x), f64(deck.v), i32(deck.atype)]
        r = _Request()
        r.units = UNITS[deck.units]
        r.lo[:] = list(map(float, deck.boxlo))
        r.hi[:] = list(map(float, deck.boxhi))
        r.ntypes = len(deck.masses)
        r.mass = _p(keep[0])
        r.natoms = deck.natoms
        r.x, r.v, r.type = _p(keep[1]), _p(keep[2]), _p(keep[3])
        r.nelements = deck.nelements
        if deck.nelements:
            keep += [i32(deck.ctype), f64(deck.nodex), f64(deck.nodev)]
            r.ctype, r.nodex, r.nodev = _p(keep[4]), _p(keep[5]), _p(keep[6])
            r.ncell[:] = list(deck.ncell)
        r.skin, r.precision, r.nsteps, r.dt, r.keep_state = skin, precision, nsteps, dt, int(keep_state)
        rid = C.c_long()
        self._chk(self.L.cac_hip_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def drain(self):
        self._chk(self.L.cac_hip_server_drain(self.h))

    def result(self, rid: int) -> dict:
        r = _Result()
        self._chk(self.L.cac_hip_server_result(self.h, rid, C.byref(r)))
        return {k: getattr(r, k) for k, _ in _Result._fields_}

    def state(self, rid: int, deck):
        x = np.zeros((deck.natoms, 3))
        nx = np.zeros((deck.nelements, 8, 3))
        self._chk(self.L.cac_hip_server_state(self.h, rid, _p(x) if deck.natoms else None,
                                              _p(nx) if deck.nelements else None))
        return x, nx

    def stats(self) -> dict:
        p50, p99, rate, hbm = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        n, peak = C.c_long(), C.c_int()
        self._chk(self.L.cac_hip_server_stats(self.h, C.byref(p50), C.byref(p99), C.byref(n), C.byref(rate),
                                              C.byref(peak), C.byref(hbm)))
        sl, ss = C.c_long(), C.c_long()
        self._chk(self.L.cac_hip_server_segments(self.h, C.byref(sl), C.byref(ss)))
        d50, d99, dn = C.c_double(), C.c_double(), C.c_long()
        self._chk(self.L.cac_hip_server_device_intervals(self.h, C.byref(d50), C.byref(d99), C.byref(dn)))
        return dict(p50_ms=p50.value, p99_ms=p99.value, steps=n.value, atom_steps_per_s=rate.value,
                    peak_active=peak.value, peak_hbm_bytes=hbm.value, segmented_launches=sl.value,
                    segmented_steps=ss.value, device_p50_ms=d50.value, device_p99_ms=d99.value,
                    device_steps=dn.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.cac_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
