"""The serving loop (cac_hip_server): replica requests on the CPU; on the GPU,
each replica's result is the same alone and in the mix, whatever the slot it
lands in, the replicas it shares the device with and the admission order."""

import numpy as np
import pytest

from cac_amd import decks

TAB = decks.cu_eam_tables()
NSTEP = 150
SKIN = 0.5  # small enough that thermal motion rebuilds the outer lists within NSTEP


def test_serving_stream_follows_its_distributions():
    sp = decks.serving_stream(n=512, seed=77, rate=25.0)
    t = np.array([s["arrival"] for s in sp])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 25.0) < 0.2 / 25.0  # Poisson arrivals at 25 per second
    size = np.array([s["size"] for s in sp])
    assert size.min() >= 2000 and size.max() <= 200000
    ls = np.log(size)
    assert abs(ls.mean() - 0.5 * (np.log(2000) + np.log(200000))) < 0.25  # log-uniform
    for key, lo, hi in (("ratio", 0.05, 0.3), ("T", 100, 600), ("strain", -0.02, 0.02)):
        v = np.array([s[key] for s in sp])
        assert v.min() >= lo and v.max() <= hi


def test_serving_deck_matches_its_request():
    spec = dict(arrival=0.0, size=20000.0, ratio=0.2, T=400.0, axis=1, strain=0.015, seed=5)
    d = decks.serving_deck(spec)
    assert abs(d.atom_equivalents / 20000.0 - 1) < 0.25
    box, core = decks.serving_box(20000.0, 0.2)
    L0 = 9 * 3.615 * np.array(box, dtype=float)
    assert np.allclose(d.boxhi - d.boxlo, L0 * np.array([1, 1.015, 1]))
    assert d.natoms == 729 * core


def _specs(n, seed):
    rng = np.random.default_rng(seed)
    return [dict(arrival=0.0, size=float(rng.uniform(2000, 9000)), ratio=float(rng.uniform(0.05, 0.3)),
                 T=float(rng.uniform(100, 600)), axis=int(rng.integers(0, 3)),
                 strain=float(rng.uniform(-0.02, 0.02)), seed=int(rng.integers(0, 2 ** 31))) for _ in range(n)]


@pytest.mark.gpu
def test_replicas_same_alone_and_in_the_mix(monkeypatch):
    """Six replicas of uneven size, temperature and strain, 150 mixed steps
    each: run alone through the resident loop, then together on three slots
    (so slots are reused by replicas of other sizes, and replicas overlap on
    the device), then on four slots stepped by one thread with segmented pair
    passes switched on (one launch per pass over the thread's replicas), then one at a time
    under a budget that admits a single replica: final positions bitwise
    equal, final energies equal."""
    import cac_amd
    specs = _specs(6, 11)
    dks = [decks.serving_deck(s) for s in specs]
    alone = []
    for d in dks:
        m = cac_amd.Model(d, "eam", skin=SKIN, eam_tables=TAB, precision=cac_amd.MIXED)
        m.setup()
        th = m.run(NSTEP, 0.001)
        alone.append((m.atoms()[0], m.elements()["nodex"] if d.nelements else None, th[-1]))
        m.close()
    for slots, budget, steppers in ((3, 0.0, None), (4, 0.0, "1"), (2, 1.0, None)):
        if steppers:
            monkeypatch.setenv("CAC_HIP_SERVE_STEPPERS", steppers)
            monkeypatch.setenv("CAC_HIP_SERVE_SEGMENT", "1")
        else:
            monkeypatch.delenv("CAC_HIP_SERVE_STEPPERS", raising=False)
            monkeypatch.delenv("CAC_HIP_SERVE_SEGMENT", raising=False)
        srv = cac_amd.Server(slots=slots, hbm_budget_bytes=budget, eam_tables=TAB)
        ids = [srv.submit(d, NSTEP, 0.001, skin=SKIN, keep_state=True) for d in dks]
        srv.drain()
        st = srv.stats()
        assert st["steps"] == 6 * NSTEP
        if budget == 0.0:
            assert 2 <= st["peak_active"] <= slots, st  # replicas overlapped
        else:
            assert st["peak_active"] == 1, st
        if not steppers:
            assert st["segmented_launches"] == 0, st  # off unless asked for
        if steppers == "1":
            # one thread steps up to four replicas: their pair passes were segmented
            assert st["segmented_launches"] > 0 and st["segmented_steps"] >= 2 * st["segmented_launches"], st
        # some replicas rebuilt their outer lists on the way (those steps are
        # finished by the server's rebuilder thread)
        assert max(srv.result(rid)["builds"] for rid in ids) > 1
        for rid, d, (x0, nx0, th0) in zip(ids, dks, alone):
            r = srv.result(rid)
            assert r["status"] == 0 and r["nsteps"] == NSTEP
            x, nx = srv.state(rid, d)
            assert np.array_equal(x, x0)
            if d.nelements:
                assert np.array_equal(nx, nx0)
            assert r["pe"] == th0[1] and r["ke"] == th0[2], (r, th0)
        srv.close()


@pytest.mark.gpu
def test_server_rejects_bad_requests_and_reports_failures():
    import cac_amd
    srv = cac_amd.Server(slots=2, eam_tables=TAB)
    d = decks.serving_deck(_specs(1, 3)[0])
    with pytest.raises(cac_amd.CacHipError):
        srv.submit(d, 0, 0.001)
    rid = srv.submit(d, 5, 0.001)
    srv.drain()
    assert srv.result(rid)["status"] == 0
    with pytest.raises(cac_amd.CacHipError):
        srv.state(rid, d)  # not kept
    srv.close()


@pytest.mark.gpu
def test_lj_replicas_in_lj_units_same_alone_and_in_the_mix():
    """A server with the LJ pair style takes replicas in LJ units (its slots
    take each replica's units): four FCC blocks of different size and
    temperature, 100 mixed steps each, alone and then together on two slots:
    final positions bitwise equal, final energies equal."""
    import cac_amd
    dks = [decks.lj_block(n=n, jitter=0.02, T=T, seed=s) for n, T, s in ((5, 0.1, 1), (6, 0.3, 2), (5, 0.5, 3),
                                                                           (7, 0.2, 4))]
    alone = []
    for d in dks:
        m = cac_amd.Model(d, "lj", skin=0.3, precision=cac_amd.MIXED)
        m.setup()
        th = m.run(100, 0.005)
        alone.append((m.atoms()[0], th[-1]))
        m.close()
    srv = cac_amd.Server(slots=2, pair="lj")
    ids = [srv.submit(d, 100, 0.005, skin=0.3, keep_state=True) for d in dks]
    srv.drain()
    assert srv.stats()["steps"] == 4 * 100
    for rid, d, (x0, th0) in zip(ids, dks, alone):
        r = srv.result(rid)
        assert r["status"] == 0 and r["nsteps"] == 100
        x, _ = srv.state(rid, d)
        assert np.array_equal(x, x0)
        assert r["pe"] == th0[1] and r["ke"] == th0[2], (r, th0)
    srv.close()


@pytest.mark.gpu
def test_replicas_across_the_serving_range():
    """Replicas drawn across the serving configuration's whole size range (2k
    to 200k atom-equivalents, log-spaced), core/coarse ratio, temperature and
    strain ranges, at its skin of 2.0 A, 100 mixed steps, four slots: none
    fails; every step is timed on the device; the smallest, the median and
    the largest are bitwise the same alone and in the mix."""
    import cac_amd
    rng = np.random.default_rng(99)
    sizes = np.geomspace(2000.0, 200000.0, 7)
    specs = [dict(arrival=0.0, size=float(sz), ratio=float(rng.uniform(0.05, 0.3)), T=float(rng.uniform(100, 600)),
                  axis=int(rng.integers(0, 3)), strain=float(rng.uniform(-0.02, 0.02)),
                  seed=int(rng.integers(0, 2 ** 31))) for sz in sizes]
    dks = [decks.serving_deck(s) for s in specs]
    ae = [d.atom_equivalents for d in dks]
    assert min(ae) < 3000 and max(ae) > 150000, ae
    nstep = 100
    srv = cac_amd.Server(slots=4, eam_tables=TAB)
    ids = [srv.submit(d, nstep, 0.001, skin=2.0, keep_state=True) for d in dks]
    srv.drain()
    st = srv.stats()
    assert st["steps"] == len(dks) * nstep
    assert st["device_steps"] == len(dks) * nstep  # one completion interval per step
    assert 0.0 < st["device_p50_ms"] <= st["device_p99_ms"]
    res = [srv.result(rid) for rid in ids]
    assert all(r["status"] == 0 and r["nsteps"] == nstep for r in res), res
    assert all(r["device_step_ms_mean"] > 0.0 for r in res)
    for k in (0, len(dks) // 2, len(dks) - 1):
        d = dks[k]
        m = cac_amd.Model(d, "eam", skin=2.0, eam_tables=TAB, precision=cac_amd.MIXED)
        m.setup()
        th = m.run(nstep, 0.001)
        x, nx = srv.state(ids[k], d)
        assert np.array_equal(x, m.atoms()[0])
        assert np.array_equal(nx, m.elements()["nodex"])
        assert res[k]["pe"] == th[-1][1] and res[k]["ke"] == th[-1][2]
        m.close()
    srv.close()


@pytest.mark.gpu
def test_server_closes_without_drain(monkeypatch):
    """Closing a server whose replicas are still queued or running, with more
    stepping threads than replicas (idle threads wait for work while the last
    replica finishes on another), finishes their work and returns."""
    import cac_amd
    monkeypatch.setenv("CAC_HIP_SERVE_STEPPERS", "4")
    d = decks.serving_deck(_specs(1, 5)[0])
    for nrep in (1, 3):
        srv = cac_amd.Server(slots=4, eam_tables=TAB)
        for _ in range(nrep):
            srv.submit(d, 60, 0.001, skin=SKIN)
        srv.close()
