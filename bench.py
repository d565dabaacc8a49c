#!/usr/bin/env python3
"""Benchmark of the CAC force + neighbor hot path on one MI355X.

Workload (default, cac_eam_x8): 255,879 core atoms + 13,985 Hex elements of
9^3 unit cells (2x2x2 Gauss points at the corners), 367,759 atom-equivalents
(atoms + Gauss points), 10.45 M represented sites, Cu-like analytic EAM
tabulated as setfl (Nr = Nrho = 2000, cut 4.95 A), skin 2.0 A, 300 K, dt 1 fs,
neighbor every 1 delay 10 check yes (the reference's defaults).

A step is one velocity-Verlet step of the whole model: half kick + drift,
neighbor decision (and rebuild when due), site refresh, EAM density /
embedding / force passes, nodal force assembly, half kick.  The metric is
atom-equivalent steps per second in mixed precision (FP32 pair terms, FP64
positions and sums); the FP64 path runs on the same deck for the force error
and the speedup.  Inputs are resident in HBM before the timed region.

The CAC deck (2x2x2 Gauss points on 9^3 elements) is not energy-stable
under the reference's equations (their linearization has growing
oscillations): it warms up, and its lists are rebuilt more often as it does,
so the rate of a timed window depends on its length.  `value` is the timed
--steps window, as the contract asks; `amortized` is the steady-state rate
independent of --steps (the list upkeep costs measured on the model at rest,
spread over the intervals of a fixed 220-step window).

Prints one JSON line.  `--profile-run` instead runs only the mixed model for
--steps steps and then a 1 GiB device-to-device copy, the calibration kernel
of the HBM traffic counters; it is the program the committed rocprofv3
summaries were taken on.  `--handover` instead times the per-call drop-in,
with host positions in and forces out every step (PCIe inside the step).
"""

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (deck builder, description, pair style)
    "cac_eam_x8": (lambda d: d.cac_eam(box_cells=(14, 16, 16), core_cells=351, ncell=9, jitter=0.02, T=300.0,
                                       seed=2024),
                   "255,879 atoms + 13,985 Hex(9^3) elements, EAM, mixed vs FP64", "eam"),
    "cac_eam": (lambda d: d.cac_eam(box_cells=(7, 8, 8), core_cells=44, ncell=9, jitter=0.02, T=300.0, seed=2024),
                "32,076 atoms + 1,748 Hex(9^3) elements, EAM, mixed vs FP64", "eam"),
    "cac_dislocation": (lambda d: d.cac_dislocation(),
                        "2,000,493 atoms around an edge dislocation dipole + 40,097 Hex(11^3) elements, "
                        "55.4 M represented sites, EAM, mixed vs FP64", "eam"),
    "lj_atoms": (lambda d: d.lj_block(n=20, a=1.5874, jitter=0.01, T=0.1, seed=12345),
                 "32,000-atom FCC LJ block (20^3 cells, a = 1.5874 sigma, rc = 2.5 sigma), LJ units, mixed vs FP64",
                 "lj"),
}
SKIN, DT = 2.0, 0.001  # metal units: the EAM decks
# neighbor skin and time step of each pair style's decks (update.cpp:112-184 defaults)
PAIR_UNITS = {"eam": (SKIN, DT), "lj": (0.3, 0.005)}
# the workload's pair style, skin and time step (set by main)
WL = {"pair": "eam", "skin": SKIN, "dt": DT}


def progress(msg):
    """One line on stderr per finished phase (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def model(cac_amd, deck, tab, precision):
    """A model of the workload deck in the workload's pair style."""
    return cac_amd.Model(deck, WL["pair"], skin=WL["skin"], eam_tables=tab if WL["pair"] == "eam" else None,
                         precision=precision)
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "eam_force_mixed_cac_eam_x8_pmc.md")
VALU_SUMMARY = os.path.join(ROOT, "profiles", "pair_passes_mixed_cac_eam_x8_valu.md")


def force_vector(m):
    import numpy as np
    a = m.atoms()[2]
    e = m.elements()["gaussf"].reshape(-1, 3)
    return np.concatenate([a, e])


def timed_run(m, steps, warmup, dt, force_timing=False):
    m.run(warmup, dt, thermo=False)
    m.sync()
    if force_timing:
        m.enable_timing(2)
    t0 = time.perf_counter()
    m.run(steps, dt, thermo=False)
    m.sync()
    return time.perf_counter() - t0


def cpu_baseline(deck, tab, nsteps, barrier=None):
    """The oracle (one-thread C restatement of the reference's algorithm) on
    the same deck: set-up (list build + first force), then nsteps NVE steps
    timed; with a barrier, every worker starts its timed steps together."""
    from oracle import oracle as O
    o = O.Oracle(deck, WL["pair"], skin=WL["skin"], eam_tables=tab if WL["pair"] == "eam" else None)
    t0 = time.perf_counter()
    o.setup(eflag=0, vflag=0)
    t_setup = time.perf_counter() - t0
    if barrier is not None:
        barrier.wait()
    t0 = time.perf_counter()
    o.run(nsteps, WL["dt"])
    t_run = time.perf_counter() - t0
    o.close()
    return t_setup, t_run


_POOL_DECK = None


def _pool_worker(args):
    tab, nsteps, barrier = args
    return cpu_baseline(_POOL_DECK, tab, nsteps, barrier)


def cpu_baseline_all_cores(deck, tab, nsteps, workers):
    """`workers` independent oracle replicas of the deck in as many processes
    (forked before this process touches the GPU), their timed steps started
    together: aggregate rate = workers x P x nsteps / the slowest replica's
    time."""
    import multiprocessing as mp
    global _POOL_DECK
    _POOL_DECK = deck
    ctx = mp.get_context("fork")
    barrier = ctx.Manager().Barrier(workers)
    with ctx.Pool(workers) as pool:
        res = pool.map(_pool_worker, [(tab, nsteps, barrier)] * workers)
    _POOL_DECK = None
    t_run = max(r[1] for r in res)
    return workers * deck.atom_equivalents * nsteps / t_run, t_run, max(r[0] for r in res)


def host_cores():
    """CPUs this process may run on (its affinity set), at most 16 workers:
    the share of host cores that goes with one GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return n, min(n, 16)


def nve_drift(cac_amd, deck, tab, precision, steps, every=100, skin=None):
    """Total-energy trace of an NVE run: drift = slope of a linear fit of
    E_total(t) per atom-equivalent (eV / atom / ps) and the largest excursion
    |E(t) - E(0)| / |E(0)|."""
    import numpy as np
    m = cac_amd.Model(deck, "eam", skin=SKIN if skin is None else skin, eam_tables=tab, precision=precision)
    m.setup()
    th = m.run(steps, DT, thermo_every=every)
    m.close()
    t, e = th[:, 0] * DT, th[:, 3]
    slope = float(np.polyfit(t, e, 1)[0]) / deck.atom_equivalents
    return {"drift_eV_per_atom_ps": slope, "max_rel_excursion": float(np.max(np.abs(e - e[0])) / abs(e[0])),
            "e0": float(e[0]), "e_end": float(e[-1])}


def handover_run(cac_amd, deck, tab, steps=50, warmup=5):
    """The per-call drop-in (a caller that keeps its integrator on the host):
    every step host positions go in (sync_positions), the rebuild decision is
    made (decide), forces are computed and come back out (download_forces), so
    PCIe is inside the step.  Positions stay fixed, so the host integrator's own
    cost is left out.  Returns the rate and the host time of each call; compute
    only queues the passes, whose device time shows up in the force download,
    which waits for them."""
    import numpy as np
    P = deck.atom_equivalents
    x = np.array(deck.x, dtype=np.float64)
    nx = np.array(deck.nodex, dtype=np.float64).reshape(-1, 8, 3)
    out = {}
    for name, prec in (("mixed", cac_amd.MIXED), ("fp64", cac_amd.FP64)):
        m = model(cac_amd, deck, tab, prec)
        m.setup()
        for _ in range(warmup):
            m.sync_positions(x, nx)
            m.decide()
            m.compute(0, 0)
            m.forces()
        t = [0.0] * 4
        t0 = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            m.sync_positions(x, nx)
            b = time.perf_counter()
            m.decide()
            c = time.perf_counter()
            m.compute(0, 0)
            d = time.perf_counter()
            m.forces()
            e = time.perf_counter()
            for i, dt in enumerate((b - a, c - b, d - c, e - d)):
                t[i] += dt
        wall = time.perf_counter() - t0
        m.close()
        out[name] = {"value": P * steps / wall, "ms_per_step": 1000.0 * wall / steps,
                     "ms_sync_positions": 1000.0 * t[0] / steps, "ms_decide": 1000.0 * t[1] / steps,
                     "ms_compute_queue": 1000.0 * t[2] / steps, "ms_download_forces": 1000.0 * t[3] / steps}
    return {"unit": "atom-eq steps/s", "steps": steps, "host_to_device_bytes": 24 * P,
            "device_to_host_bytes": 24 * P + 24 * 8 * deck.nelements, **out}


def serve_run(cac_amd, decks, tab, n, steps, rates, slots, seed=77):
    """The serving workload: ``n`` replica requests (sizes log-uniform 2k-200k
    atom-eq, core/coarse 0.05-0.3, 100-600 K, +-2 % uniaxial strain) arriving
    as a Poisson stream, each running ``steps`` mixed steps on the resident
    server; one run per arrival rate in ``rates`` (per second), over the same
    replicas.  Requests are submitted at their arrival times; decks are built
    before the clock starts."""
    import numpy as np
    specs = decks.serving_stream(n=n, seed=seed, rate=1.0)
    dks = [decks.serving_deck(sp) for sp in specs]
    out = {}
    for rate in rates:
        arrive = [sp["arrival"] / rate for sp in specs]  # the unit-rate stream, rescaled
        srv = cac_amd.Server(slots=slots, eam_tables=tab)
        t0 = time.perf_counter()
        ids = []
        for t_arr, d in zip(arrive, dks):
            wait = t_arr - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(d, steps, DT, skin=SKIN, precision=cac_amd.MIXED))
        srv.drain()
        wall = time.perf_counter() - t0
        res = [srv.result(i) for i in ids]
        st = srv.stats()
        err = srv.L.cac_hip_server_last_error(srv.h).decode()
        srv.close()
        q = np.array([r["t_admit"] - r["t_submit"] for r in res])
        mean_step = np.array([r["device_step_ms_mean"] for r in res])
        ae = np.array([r["atom_equivalents"] for r in res], dtype=float)
        out[f"rate_{rate:g}_per_s"] = {
            "replicas": n, "steps_each": steps, "arrival_rate_per_s": rate, "slots": slots,
            "failed": int(sum(r["status"] != 0 for r in res)), "last_error": err or None,
            "atom_eq_mean": float(ae.mean()), "atom_eq_range": [float(ae.min()), float(ae.max())],
            # step intervals as the device completed them (an event after each
            # replica step on its stream); the host-side queueing intervals beside them
            "step_ms_p50": st["device_p50_ms"], "step_ms_p99": st["device_p99_ms"],
            "step_intervals_timed_on_device": st["device_steps"],
            "host_queue_interval_ms_p50": st["p50_ms"], "host_queue_interval_ms_p99": st["p99_ms"],
            "replica_mean_step_ms_p50": float(np.percentile(mean_step, 50)),
            "replica_mean_step_ms_p99": float(np.percentile(mean_step, 99)),
            "queue_wait_s_p50": float(np.percentile(q, 50)), "queue_wait_s_p99": float(np.percentile(q, 99)),
            "atom_steps_per_s": st["atom_steps_per_s"], "wall_s": wall, "peak_active": st["peak_active"],
            "peak_hbm_gib_estimate": st["peak_hbm_bytes"] / 2 ** 30}
    return out


def per_step_force_error(cac_amd, deck, tab, steps, sample=50):
    """Mixed-vs-FP64 force error along an FP64 trajectory: at each sampled
    step the FP64 model's positions go into a mixed model (which keeps its own
    lists, rebuilt when its trigger says so) and both forces on atoms and
    Gauss points are compared.  Per step: max |df| / max |f|, rms |df| /
    rms |f|, and per-point |df_i| / |f_i| (p50, p99, max) over the points
    whose |f_i| is at least 1 % of rms |f| (points near force balance have no
    meaningful relative error)."""
    import numpy as np
    m64 = model(cac_amd, deck, tab, cac_amd.FP64)
    mmx = model(cac_amd, deck, tab, cac_amd.MIXED)
    m64.setup()
    mmx.setup()
    at = set(np.linspace(1, steps, min(sample, steps)).round().astype(int).tolist())
    rows = []
    for k in range(1, steps + 1):
        m64.run(1, WL["dt"], thermo=False)
        if k not in at:
            continue
        x = m64.atoms()[0]
        el = m64.elements()
        mmx.sync_positions(x, el["nodex"])
        if mmx.decide():
            mmx.build_neighbor()
        mmx.compute(0, 0)
        fa, _, fg = mmx.forces()
        f64 = np.concatenate([m64.atoms()[2], el["gaussf"].reshape(-1, 3)])
        fmx = np.concatenate([fa, fg.reshape(-1, 3)])
        df = np.linalg.norm(fmx - f64, axis=1)
        fm = np.linalg.norm(f64, axis=1)
        rms = np.sqrt(np.mean(fm ** 2))
        keep = fm >= 0.01 * rms
        pp = df[keep] / fm[keep]
        rows.append((k, df.max() / fm.max(), np.sqrt(np.mean(df ** 2)) / rms, np.percentile(pp, 50),
                     np.percentile(pp, 99), pp.max(), int((~keep).sum())))
    m64.close()
    mmx.close()
    r = np.array(rows, dtype=float)
    return {"steps_compared": len(rows), "of_steps": steps,
            "max_rel": {"mean": float(r[:, 1].mean()), "worst": float(r[:, 1].max())},
            "rms_rel": {"mean": float(r[:, 2].mean()), "worst": float(r[:, 2].max())},
            "per_point_rel": {"p50": float(np.median(r[:, 3])), "p99": float(r[:, 4].max()),
                              "max": float(r[:, 5].max()),
                              "points_excluded_below_1pct_rms": int(r[:, 6].max())}}


def upkeep_costs(m, reps=5):
    """Extra device time of list upkeep, measured on a model at rest: a
    force computation alone (t0), after a forced re-prune of the inner list
    (t_prune) and after a full outer rebuild (t_build); each the median of
    `reps` synchronous calls, so host overheads cancel in the differences."""
    import statistics

    def timed(fn):
        ts = []
        for _ in range(reps):
            m.sync()
            a = time.perf_counter()
            fn()
            m.sync()
            ts.append(time.perf_counter() - a)
        return statistics.median(ts)

    _, _, delta, _ = m.inner_stats()
    t0 = timed(lambda: m.compute(0, 0))

    def prune():
        m.set_inner_skin(delta)
        m.compute(0, 0)

    def build():
        m.build_neighbor()
        m.compute(0, 0)

    t_prune = timed(prune)
    t_build = timed(build)
    return max(t_prune - t0, 0.0), max(t_build - t0, 0.0)


def force_bytes(P, inner_mean, live_sites):
    """Algorithmic HBM bytes of one mixed force launch (EAM or LJ).

    Per row: its site index, point index and inner-list length (12 B) and the
    FP64 force it writes (24 B); per inner-list entry its 4-byte word; per
    site the inner lists reference, its 16-byte FP32 record (x, y, z and F'
    for EAM) read once.
    """
    return P * (36.0 + 4.0 * inner_mean) + 16.0 * live_sites


def traffic_from_summary(path=TRAFFIC_SUMMARY):
    """The JSON block of a committed counter summary (HBM bytes per force launch measured with FETCH_SIZE /
    WRITE_SIZE, or VALU instructions per force launch), or None."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```json\s*(\{.*?\})\s*```", text, re.S)
    if not m:
        return None
    return json.loads(m.group(1))


def hip_copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of nbytes through the HIP runtime (a plain copy kernel)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemset failed")
    if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemcpy failed")
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cac_eam_x8", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=25)
    ap.add_argument("--cpu-steps-all", type=int, default=8,
                    help="NVE steps of each replica of the all-core CPU baseline")
    ap.add_argument("--error-steps", type=int, default=1,
                    help="1: compare mixed and FP64 forces along an FP64 trajectory of --steps steps (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--inner-skin", type=float, default=-1.0,
                    help="prune margin delta of the inner list (A); < 0 takes 0.3 * skin")
    ap.add_argument("--drift-steps", type=int, default=10000,
                    help="NVE steps of the FP64-vs-mixed drift comparison on the 32,000-atom EAM block (0: skip)")
    ap.add_argument("--serve-replicas", type=int, default=512,
                    help="replicas of the serving run (0: skip)")
    ap.add_argument("--serve-steps", type=int, default=250,
                    help="steps per replica: the CAC equations with 2x2x2 Gauss points on 9^3 elements are "
                         "linearly unstable (flutter; tests/test_cac_stability.py), and replicas at 100-600 K "
                         "deform an element past 2x after about 450 steps")
    ap.add_argument("--serve-rates", default="100,400",
                    help="Poisson arrival rates (replicas per second), comma-separated: a light load for latency "
                         "and a saturating one for capacity")
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-only", action="store_true", help="run only the serving workload")
    ap.add_argument("--handover", action="store_true",
                    help="run only the per-call drop-in (host positions in, forces out every step) on the workload")
    ap.add_argument("--profile-run", action="store_true",
                    help="mixed model only, --steps steps, then a 1 GiB copy (program of the rocprofv3 summaries)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; runs on several GPUs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import cac_amd
    from cac_amd import decks

    rates = [float(r) for r in args.serve_rates.split(",") if r]
    build_deck, desc, pair = WORKLOADS[args.workload]
    WL.update(pair=pair, skin=PAIR_UNITS[pair][0], dt=PAIR_UNITS[pair][1])
    dt = WL["dt"]
    force_name = "eam_force" if pair == "eam" else "lj_force"
    tab = decks.cu_eam_tables()
    if args.serve_only:
        print(json.dumps({"serving": serve_run(cac_amd, decks, tab, args.serve_replicas, args.serve_steps,
                                               rates, args.serve_slots)}))
        return
    deck = build_deck(decks)
    P = deck.atom_equivalents
    if args.handover:
        print(json.dumps({"handover": handover_run(cac_amd, deck, tab), "workload": args.workload}))
        return

    if args.profile_run:
        m = model(cac_amd, deck, tab, cac_amd.MIXED)
        if args.inner_skin >= 0:
            m.set_inner_skin(args.inner_skin)
        m.setup()
        print(f"profile-run: {args.workload} set up", flush=True)
        m.run(args.steps, dt, thermo=False)
        m.sync()
        _, nbar_in, _, live = m.inner_stats()
        print(f"profile-run: {args.steps} mixed steps done; inner mean {nbar_in:.2f}, live sites {live}, "
              f"force bytes/launch {force_bytes(P, nbar_in, live):.0f}", flush=True)
        hip_copy_calibration()
        print("profile-run: calibration copy of 1 GiB done", flush=True)
        return

    # CPU baselines first: the worker processes are forked before this
    # process touches the GPU
    cpu = None
    progress(f"{args.workload}: {deck.natoms} atoms + {deck.nelements} elements, {P} atom-eq")
    if not args.no_cpu_baseline:
        ncpu, workers = host_cores()
        t_setup, t_run = cpu_baseline(deck, tab, args.cpu_steps)
        progress(f"CPU baseline, one thread: {P * args.cpu_steps / t_run:.3g} atom-eq steps/s")
        agg, t_agg, _ = cpu_baseline_all_cores(deck, tab, args.cpu_steps_all, workers)
        progress(f"CPU baseline, {workers} processes: {agg:.3g} atom-eq steps/s")
        cpu = {"value": P * args.cpu_steps / t_run, "unit": "atom-eq steps/s", "cores": 1, "kind": "port",
               "sample": f"oracle (CPU restatement, one thread) running {args.cpu_steps} NVE steps of this deck "
                         f"after its setup; setup (neighbor build + first force, {t_setup:.1f} s) excluded",
               "all_cores": {"value": agg, "unit": "atom-eq steps/s", "cores": workers, "host_cpus": ncpu,
                             "sample": f"{workers} independent oracle replicas of this deck, one process each, "
                                       f"{args.cpu_steps_all} NVE steps each started together after their setup; "
                                       f"rate = replicas x atom-eq x steps / slowest replica ({t_agg:.1f} s)"}}

    # mixed-vs-FP64 force error per step along an FP64 trajectory (own models)
    step_err = per_step_force_error(cac_amd, deck, tab, args.steps) if args.error_steps else None
    if step_err:
        progress("per-step force error done")

    # FP64 and mixed models on the same deck
    m64 = model(cac_amd, deck, tab, cac_amd.FP64)
    mmx = model(cac_amd, deck, tab, cac_amd.MIXED)
    if args.inner_skin >= 0:
        m64.set_inner_skin(args.inner_skin)
        mmx.set_inner_skin(args.inner_skin)
    m64.setup()
    f64 = force_vector(m64)
    e64, _ = m64.energy()
    mmx.setup()
    fmx = force_vector(mmx)
    emx, _ = mmx.energy()
    df = np.linalg.norm(fmx - f64, axis=1)
    fmag = np.linalg.norm(f64, axis=1)
    max_rel_err = float(df.max() / fmag.max())
    rms_rel_err = float(np.sqrt(np.mean(df ** 2)) / np.sqrt(np.mean(fmag ** 2)))
    keep = fmag >= 0.01 * np.sqrt(np.mean(fmag ** 2))
    pp = df[keep] / fmag[keep]
    info = mmx.neigh_info()
    nbar = info["entries"] / info["rows"]
    progress(f"set-up of both models done; mixed vs FP64 force error {max_rel_err:.3g}")

    b0, ib0 = mmx.neigh_stats()[0], mmx.inner_stats()[0]
    # the force pass is timed live over the timed steps: one hipEvent pair per
    # launch on the model's stream, read back after the timed region
    mmx.run(args.warmup, dt, thermo=False)
    mmx.sync()
    b1, ib1 = mmx.neigh_stats()[0], mmx.inner_stats()[0]
    t_mx = timed_run(mmx, args.steps, 0, dt, force_timing=True)
    force_us, force_launches = mmx.kernel_time(force_name)
    mmx.enable_timing(False)
    nsteps_mx = args.steps + args.warmup
    rebuilds = (mmx.neigh_stats()[0] - b0) / nsteps_mx
    prunes = (mmx.inner_stats()[0] - ib0) / nsteps_mx
    n_build_timed = mmx.neigh_stats()[0] - b1
    n_prune_timed = mmx.inner_stats()[0] - ib1
    t_64 = timed_run(m64, args.steps, args.warmup, dt)
    value = P * args.steps / t_mx
    value64 = P * args.steps / t_64
    progress(f"timed runs: mixed {value:.4g}, FP64 {value64:.4g} atom-eq steps/s")

    # per-kernel durations from hipEvents on the context's stream (a separate
    # short run, so the events do not perturb the timed one)
    pair_names = ("eam_density", "eam_embed", "eam_force") if pair == "eam" else ("lj_force",)
    names = pair_names + ("site_refresh", "nodef", "nve", "neigh_build", "inner_build")
    mmx.enable_timing(True)
    mmx.run(args.profile_steps, dt, thermo=False)
    kt = {k: mmx.kernel_time(k) for k in names}
    mmx.enable_timing(False)
    m64.enable_timing(True)
    m64.run(args.profile_steps, dt, thermo=False)
    kt64 = {k: m64.kernel_time(k) for k in pair_names}
    m64.enable_timing(False)
    _, nbar_in, delta, live = mmx.inner_stats()

    # Steady-state rate, independent of --steps: the upkeep of the lists
    # (re-prune, outer rebuild) costs what a model at rest measures, and
    # recurs at the intervals a fixed window (the reference's 200-step run
    # after 20 warm-up steps, from the same set-up) shows; the timed window's
    # step without upkeep is its time less the upkeep it held.
    c_prune, c_build = upkeep_costs(mmx)
    progress(f"upkeep at rest: re-prune {1000 * c_prune:.3f} ms, rebuild {1000 * c_build:.3f} ms")
    mwin = model(cac_amd, deck, tab, cac_amd.MIXED)
    if args.inner_skin >= 0:
        mwin.set_inner_skin(args.inner_skin)
    mwin.setup()
    wb0, wp0 = mwin.neigh_stats()[0], mwin.inner_stats()[0]
    mwin.run(220, dt, thermo=False)
    mwin.sync()
    win_builds, win_prunes = mwin.neigh_stats()[0] - wb0, mwin.inner_stats()[0] - wp0
    mwin.close()
    t_plain = max((t_mx - n_prune_timed * c_prune - n_build_timed * c_build) / args.steps, 1e-9)
    t_amort = t_plain + (win_prunes / 220.0) * c_prune + (win_builds / 220.0) * c_build
    amortized = {
        "value": P / t_amort, "ms_per_step": 1000.0 * t_amort,
        "step_without_upkeep_ms": 1000.0 * t_plain,
        "reprune_ms": 1000.0 * c_prune, "rebuild_ms": 1000.0 * c_build,
        "window": "220 steps from set-up (20 warm-up + the reference's 200-step run)",
        "reprunes_per_step": win_prunes / 220.0, "rebuilds_per_step": win_builds / 220.0,
        "timed_window": {"reprunes": n_prune_timed, "rebuilds": n_build_timed},
        "note": "value is the timed --steps window (the contract); this is the steady-state rate that "
                "spreads the measured upkeep costs over the fixed window's intervals"}

    # dominant kernel: the mixed force pass, timed over the timed steps
    bytes_force = force_bytes(P, nbar_in, live)
    achieved = bytes_force / (force_us * 1e-6) / 1e9
    tr = traffic_from_summary() if args.workload == "cac_eam_x8" else None
    vs = traffic_from_summary(VALU_SUMMARY) if args.workload == "cac_eam_x8" else None
    valu = None
    if vs:
        # VALU instructions per inner-list entry are a property of the code
        # (counted once under rocprofv3); scaled to this run's lists and
        # divided by the live kernel time
        insts = vs["valu_insts_per_launch"] / vs["entries_per_launch"] * P * nbar_in
        rate = insts / (force_us * 1e-6)
        valu = {"insts_per_launch": insts, "achieved_insts_per_s": rate,
                "peak_insts_per_s": vs["peak_valu_insts_per_s"], "frac": rate / vs["peak_valu_insts_per_s"],
                "source": os.path.relpath(VALU_SUMMARY, ROOT)}
    # the whole step against HBM: 224 + 8 N (N = mean full-list entries) bytes
    # per atom-eq step for EAM (state streamed by the two list passes and the
    # integrator, one 8-byte list entry per neighbor and pass), 168 + 4 N for
    # LJ (one list pass); list upkeep separately: a rebuild writes and re-reads
    # the outer list and the site records, a re-prune reads the outer list and
    # writes the inner one
    b_step = 224.0 + 8.0 * nbar if pair == "eam" else 168.0 + 4.0 * nbar
    b_rebuild = P * 8.0 * nbar + 32.0 * info["sites"] + 8.0 * deck.represented_sites
    b_prune = P * (4.0 * nbar + 4.0 * nbar_in + 32.0) + 16.0 * info["sites"]
    step_roof = {"bytes_per_atom_eq_step": b_step, "achieved_GBs": value * b_step / 1e9,
                 "frac": value * b_step / 1e9 / HBM_PEAK_GBS,
                 "frac_amortized": amortized["value"] * b_step / 1e9 / HBM_PEAK_GBS,
                 "upkeep_bytes_per_atom_eq_step": (amortized["rebuilds_per_step"] * b_rebuild +
                                                   amortized["reprunes_per_step"] * b_prune) / P,
                 "rebuild_bytes": b_rebuild, "reprune_bytes": b_prune}

    drift = None
    if args.drift_steps > 0:
        drift = {}
        blk = decks.eam_block()
        cac2 = decks.cac_eam(box_cells=(6, 7, 7), core_cells=40, ncell=2, jitter=0.02, T=300.0, seed=2024)
        for key, dk, label, skin in (
                ("eam_block", blk, "EAM Cu FCC 20^3 periodic block, 32,000 atoms, 300 K, dt 1 fs", SKIN),
                ("cac_ncell2", cac2, f"CAC EAM deck of the benchmark geometry with 2^3-site elements "
                                     f"({cac2.natoms} atoms + {cac2.nelements} elements, every site a node: "
                                     f"the CAC path at its energy-conserving limit), 300 K, dt 1 fs", 1.0)):
            d64 = nve_drift(cac_amd, dk, tab, cac_amd.FP64, args.drift_steps, skin=skin)
            dmx = nve_drift(cac_amd, dk, tab, cac_amd.MIXED, args.drift_steps, skin=skin)
            drift[key] = {"deck": label, "steps": args.drift_steps, "fp64": d64, "mixed": dmx,
                          "ratio_mixed_over_fp64": abs(dmx["drift_eV_per_atom_ps"]) /
                          max(abs(d64["drift_eV_per_atom_ps"]), 1e-300)}
            progress(f"NVE drift on {key} done")
        # the benchmark deck itself over its stable window: both precisions
        # follow the CAC equations' flutter growth, so the energy change is
        # that growth, not integration error
        w = {}
        for name, prec in (("fp64", cac_amd.FP64), ("mixed", cac_amd.MIXED)):
            mm = model(cac_amd, deck, tab, prec)
            mm.setup()
            th = mm.run(200, dt, thermo_every=50)
            mm.close()
            w[name] = {"e0": float(th[0, 3]), "e_200": float(th[-1, 3]),
                       "rel_change": float((th[-1, 3] - th[0, 3]) / abs(th[0, 3]))}
        w["ratio_mixed_over_fp64"] = abs(w["mixed"]["rel_change"]) / max(abs(w["fp64"]["rel_change"]), 1e-300)
        drift[args.workload + "_200_steps"] = w

    serving = None
    if args.serve_replicas > 0:
        serving = serve_run(cac_amd, decks, tab, args.serve_replicas, args.serve_steps, rates, args.serve_slots)
        progress("serving done")

    line = {
        # the served p50/p99 step times are the "serving" object of the same line
        "metric": "atom-steps/s (mixed vs FP64 @ <1e-5 rel force err); served p50/p99 step ms",
        "value": value,
        "unit": "atom-eq steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_mx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms / fp64 positions and sums",
        "data": ("synthetic (seeded FCC Cu CAC deck, analytic EAM tabulated as setfl)" if pair == "eam"
                 else "synthetic (seeded FCC LJ block)"),
        "config": {"workload": args.workload, "description": desc, "atoms": deck.natoms,
                   "elements": deck.nelements, "atom_equivalents": P, "represented_sites": deck.represented_sites,
                   "pair_style": pair, "skin": WL["skin"], "dt": dt,
                   "units": "metal (A, ps)" if pair == "eam" else "lj (sigma, tau)",
                   "neighbor_mean_full": round(nbar, 2), "sites": info["sites"],
                   "inner_delta": delta, "inner_mean": round(nbar_in, 2), "inner_live_sites": live,
                   "outer_builds_per_step": round(rebuilds, 4), "inner_prunes_per_step": round(prunes, 4)},
        "amortized": amortized,
        "fp64": {"value": value64, "ms_per_step": 1000.0 * t_64 / args.steps,
                 "speedup_mixed_over_fp64": value / value64},
        "force_error_vs_fp64": {"max_rel": max_rel_err, "rms_rel": rms_rel_err,
                                "energy_rel": abs(emx - e64) / abs(e64),
                                "per_point_rel": {"p50": float(np.median(pp)), "p99": float(np.percentile(pp, 99)),
                                                  "max": float(pp.max()),
                                                  "points_excluded_below_1pct_rms": int((~keep).sum())},
                                "per_step": step_err},
        "kernels_us": {k: round(v[0], 3) for k, v in kt.items() if v[1]},
        "kernels_us_fp64": {k: round(v[0], 3) for k, v in kt64.items() if v[1]},
        "roofline": {"bound": "hbm", "kernel": f"{force_name} (mixed)", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "kernel_us": force_us, "kernel_launches": force_launches,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["bytes_per_launch"] if tr else None,
                     "algorithmic_bytes": bytes_force,
                     "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if tr else None,
                     "valu": valu, "step": step_roof},
        "nve_drift": drift,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
