"""CPU checks of the EAM table builders in cac_amd/csrc/eam_tables.hpp, built
with g++ into a small driver (tests/eam_tables_check.cpp).

interpolate() must give the oracle's splines (PairEAM::interpolate,
pair_eam.cpp:586-614).  The mixed path's force table holds, per knot
interval of the reference's splines, cubics in the knot fraction p for
A = -rho'(r)/r, B = -phi'(r)/r and C = phi(r); evaluated in FP64 they must
reproduce the reference's per-pair FP64 arithmetic (pair_eam.cpp:933-950:
knot, clamps, z2 / r, z2' / r - phi / r) to 2e-7 of each function's range
(the FP32 rounding of the stored coefficients; the interpolation itself is
about 1e-11), and evaluated the way the FP32 force pass does within 2e-6.
Slopes set to zero would miss by about 1e-3, a dropped phi / r term by 1."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from cac_amd import decks
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TAB = decks.cu_eam_tables()


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("eamtab")
    exe = str(d / "eam_tables_check")
    subprocess.run([gxx, "-std=c++17", "-O2", "-I", os.path.join(ROOT, "cac_amd", "csrc"),
                    os.path.join(ROOT, "tests", "eam_tables_check.cpp"), "-o", exe], check=True)
    inp, out = str(d / "in.bin"), str(d / "out.bin")
    nr, dr = TAB.nr, TAB.dr
    with open(inp, "wb") as f:
        f.write(np.int32(nr).tobytes())
        f.write(np.float64(dr).tobytes())
        f.write(np.asarray(TAB.rhor, np.float64).tobytes())
        f.write(np.asarray(TAB.z2r, np.float64).tobytes())
    subprocess.run([exe, inp, out], check=True)
    raw = open(out, "rb").read()
    ns = (nr + 1) * 7 * 8
    rs = np.frombuffer(raw[:ns], np.float64).reshape(nr + 1, 7)
    zs = np.frombuffer(raw[ns:2 * ns], np.float64).reshape(nr + 1, 7)
    t = np.frombuffer(raw[2 * ns:], np.float32).reshape(3, nr + 2, 4)
    return rs, zs, t


def reference_pair(rs, zs, r):
    """The reference's per-pair FP64 arithmetic for one element
    (pair_eam.cpp:933-950): returns -rho'/r, -phi'/r, phi."""
    nr, rdr = TAB.nr, 1.0 / TAB.dr
    p = r * rdr + 1.0
    m = np.minimum(p.astype(np.int64), nr - 1)
    p = np.minimum(p - m, 1.0)
    cr, cz = rs[m], zs[m]
    rhop = (cr[:, 0] * p + cr[:, 1]) * p + cr[:, 2]
    z2p = (cz[:, 0] * p + cz[:, 1]) * p + cz[:, 2]
    z2 = ((cz[:, 3] * p + cz[:, 4]) * p + cz[:, 5]) * p + cz[:, 6]
    recip = 1.0 / r
    phi = z2 * recip
    phip = z2p * recip - phi * recip
    return -rhop * recip, -phip * recip, phi


def table_eval(t, r, dtype):
    """The mixed force pass's evaluation: x = r rdr + 1, m = int(x), p its
    fraction, Horner in p (in FP64 or FP32)."""
    rdr = dtype(1.0 / TAB.dr)
    x = (r.astype(dtype) * rdr + dtype(1.0)).astype(dtype)
    m = np.minimum(x.astype(np.int64), TAB.nr + 1)
    p = (x - m).astype(dtype)
    out = []
    for w in range(3):
        c = t[w][m].astype(dtype)
        out.append((((c[:, 0] * p + c[:, 1]) * p + c[:, 2]) * p + c[:, 3]).astype(np.float64))
    return out


def test_interpolate_equals_the_oracle(built):
    rs, zs, _ = built
    f = np.zeros(TAB.nr + 1)
    f[1:] = TAB.rhor
    assert np.array_equal(rs, O.interpolate(f, TAB.dr))
    f[1:] = TAB.z2r
    assert np.array_equal(zs, O.interpolate(f, TAB.dr))


def test_mixed_force_table_reproduces_the_reference(built):
    rs, zs, t = built
    # every knot interval from 1.5 A (closer than any Cu pair at these
    # temperatures) to the cutoff, about 100 points per interval plus the knots
    r = np.concatenate([np.linspace(1.5, TAB.cut, 200_001)[:-1], np.arange(600, TAB.nr) * TAB.dr])
    r = r[(r >= 1.5) & (r < TAB.cut)]
    ref = reference_pair(rs, zs, r)
    got64 = table_eval(t, r, np.float64)
    got32 = table_eval(t, r, np.float32)
    for w, name in enumerate(("A", "B", "C")):
        scale = np.abs(ref[w]).max()
        e64 = np.abs(got64[w] - ref[w]).max() / scale
        e32 = np.abs(got32[w] - ref[w]).max() / scale
        assert e64 < 2e-7, (name, e64)
        assert e32 < 2e-6, (name, e32)
    # the pair force of two sites with F' = -2 each, as the pass forms it
    fsum = -4.0
    fref = fsum * ref[0] + ref[1]
    f32 = fsum * got32[0] + got32[1]
    assert np.abs(f32 - fref).max() / np.abs(fref).max() < 2e-6


def test_mixed_force_table_clamp_and_zero_knot(built):
    rs, zs, t = built
    nr = TAB.nr
    end = reference_pair(rs, zs, np.array([(nr - 1) * TAB.dr * (1 - 1e-15)]))
    for w in range(3):
        # knot nr: the reference's clamp (knot nr - 1 at p = 1), constant
        assert np.all(t[w][nr][:3] == 0.0)
        assert t[w][nr][3] == pytest.approx(end[w][0], rel=1e-6, abs=1e-12)
        # knot nr + 1: all zeros, an exact zero for an entry beyond rc
        assert np.all(t[w][nr + 1] == 0.0)
        # knot 0 is never read (m = int(r rdr + 1) >= 1)
        assert np.all(t[w][0] == 0.0)
