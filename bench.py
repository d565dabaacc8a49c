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
under the reference's equations: it warms up, and its lists are rebuilt more
often as it does, so the per-step cost grows with --steps; the default
(20 warm-up + 200 timed steps) is the window `value` is quoted on.

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
    # name: (deck builder, description)
    "cac_eam_x8": (lambda d: d.cac_eam(box_cells=(14, 16, 16), core_cells=351, ncell=9, jitter=0.02, T=300.0,
                                       seed=2024),
                   "255,879 atoms + 13,985 Hex(9^3) elements, EAM, mixed vs FP64"),
    "cac_eam": (lambda d: d.cac_eam(box_cells=(7, 8, 8), core_cells=44, ncell=9, jitter=0.02, T=300.0, seed=2024),
                "32,076 atoms + 1,748 Hex(9^3) elements, EAM, mixed vs FP64"),
    "cac_dislocation": (lambda d: d.cac_dislocation(),
                        "2,000,493 atoms around an edge dislocation dipole + 40,097 Hex(11^3) elements, "
                        "55.4 M represented sites, EAM, mixed vs FP64"),
}
SKIN, DT = 2.0, 0.001
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


def cpu_baseline(deck, tab, nsteps):
    """The oracle (one-thread C restatement of the reference's algorithm) on the same deck."""
    from oracle import oracle as O
    o = O.Oracle(deck, "eam", skin=SKIN, eam_tables=tab)
    t0 = time.perf_counter()
    o.setup(eflag=0, vflag=0)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.run(nsteps, DT)
    t_run = time.perf_counter() - t0
    o.close()
    return t_setup, t_run


def nve_drift(cac_amd, deck, tab, precision, steps, every=100):
    """Total-energy trace of an NVE run: drift = slope of a linear fit of
    E_total(t) per atom-equivalent (eV / atom / ps) and the largest excursion
    |E(t) - E(0)| / |E(0)|."""
    import numpy as np
    m = cac_amd.Model(deck, "eam", skin=SKIN, eam_tables=tab, precision=precision)
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
        m = cac_amd.Model(deck, "eam", skin=SKIN, eam_tables=tab, precision=prec)
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
        mean_step = np.array([r["step_ms_mean"] for r in res])
        ae = np.array([r["atom_equivalents"] for r in res], dtype=float)
        out[f"rate_{rate:g}_per_s"] = {
            "replicas": n, "steps_each": steps, "arrival_rate_per_s": rate, "slots": slots,
            "failed": int(sum(r["status"] != 0 for r in res)), "last_error": err or None,
            "atom_eq_mean": float(ae.mean()), "atom_eq_range": [float(ae.min()), float(ae.max())],
            "step_ms_p50": st["p50_ms"], "step_ms_p99": st["p99_ms"],
            "replica_mean_step_ms_p50": float(np.percentile(mean_step, 50)),
            "replica_mean_step_ms_p99": float(np.percentile(mean_step, 99)),
            "queue_wait_s_p50": float(np.percentile(q, 50)), "queue_wait_s_p99": float(np.percentile(q, 99)),
            "atom_steps_per_s": st["atom_steps_per_s"], "wall_s": wall, "peak_active": st["peak_active"],
            "peak_hbm_gib_estimate": st["peak_hbm_bytes"] / 2 ** 30}
    return out


def force_bytes(P, inner_mean, live_sites):
    """Algorithmic HBM bytes of one mixed EAM force launch.

    Per row: its site index, point index and inner-list length (12 B) and the
    FP64 force it writes (24 B); per inner-list entry its 4-byte word; per
    site the inner lists reference, its FP32 record (x, y, z, F') read once.
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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--inner-skin", type=float, default=-1.0,
                    help="prune margin delta of the inner list (A); < 0 takes 0.3 * skin")
    ap.add_argument("--drift-steps", type=int, default=10000,
                    help="NVE steps of the FP64-vs-mixed drift comparison on the 32,000-atom EAM block (0: skip)")
    ap.add_argument("--serve-replicas", type=int, default=512,
                    help="replicas of the serving run (0: skip)")
    ap.add_argument("--serve-steps", type=int, default=250,
                    help="steps per replica: CAC with 2x2x2 Gauss points on 9^3 elements is unstable past "
                         "about 450 steps (element deformed more than 2x, in the oracle as on the GPU)")
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
    build_deck, desc = WORKLOADS[args.workload]
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
        m = cac_amd.Model(deck, "eam", skin=SKIN, eam_tables=tab, precision=cac_amd.MIXED)
        if args.inner_skin >= 0:
            m.set_inner_skin(args.inner_skin)
        m.setup()
        print(f"profile-run: {args.workload} set up", flush=True)
        m.run(args.steps, DT, thermo=False)
        m.sync()
        _, nbar_in, _, live = m.inner_stats()
        print(f"profile-run: {args.steps} mixed steps done; inner mean {nbar_in:.2f}, live sites {live}, "
              f"force bytes/launch {force_bytes(P, nbar_in, live):.0f}", flush=True)
        hip_copy_calibration()
        print("profile-run: calibration copy of 1 GiB done", flush=True)
        return

    # FP64 and mixed models on the same deck
    m64 = cac_amd.Model(deck, "eam", skin=SKIN, eam_tables=tab, precision=cac_amd.FP64)
    mmx = cac_amd.Model(deck, "eam", skin=SKIN, eam_tables=tab, precision=cac_amd.MIXED)
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
    info = mmx.neigh_info()
    nbar = info["entries"] / info["rows"]

    b0, ib0 = mmx.neigh_stats()[0], mmx.inner_stats()[0]
    # the force pass is timed live over the timed steps: one hipEvent pair per
    # launch on the model's stream, read back after the timed region
    t_mx = timed_run(mmx, args.steps, args.warmup, DT, force_timing=True)
    force_us, force_launches = mmx.kernel_time("eam_force")
    mmx.enable_timing(False)
    nsteps_mx = args.steps + args.warmup
    rebuilds = (mmx.neigh_stats()[0] - b0) / nsteps_mx
    prunes = (mmx.inner_stats()[0] - ib0) / nsteps_mx
    t_64 = timed_run(m64, args.steps, args.warmup, DT)
    value = P * args.steps / t_mx
    value64 = P * args.steps / t_64

    # per-kernel durations from hipEvents on the context's stream (a separate
    # short run, so the events do not perturb the timed one)
    names = ("eam_density", "eam_embed", "eam_force", "site_refresh", "nodef", "nve", "neigh_build", "inner_build")
    mmx.enable_timing(True)
    mmx.run(args.profile_steps, DT, thermo=False)
    kt = {k: mmx.kernel_time(k) for k in names}
    mmx.enable_timing(False)
    m64.enable_timing(True)
    m64.run(args.profile_steps, DT, thermo=False)
    kt64 = {k: m64.kernel_time(k) for k in ("eam_density", "eam_embed", "eam_force")}
    m64.enable_timing(False)
    _, nbar_in, delta, live = mmx.inner_stats()

    # dominant kernel: the mixed EAM force pass, timed over the timed steps
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

    drift = None
    if args.drift_steps > 0:
        blk = decks.eam_block()
        d64 = nve_drift(cac_amd, blk, tab, cac_amd.FP64, args.drift_steps)
        dmx = nve_drift(cac_amd, blk, tab, cac_amd.MIXED, args.drift_steps)
        drift = {"deck": "EAM Cu FCC 20^3 periodic block, 32,000 atoms, 300 K, dt 1 fs", "steps": args.drift_steps,
                 "fp64": d64, "mixed": dmx,
                 "ratio_mixed_over_fp64": abs(dmx["drift_eV_per_atom_ps"]) / max(abs(d64["drift_eV_per_atom_ps"]),
                                                                                   1e-300)}

    serving = None
    if args.serve_replicas > 0:
        serving = serve_run(cac_amd, decks, tab, args.serve_replicas, args.serve_steps, rates, args.serve_slots)

    cpu = None
    if not args.no_cpu_baseline:
        t_setup, t_run = cpu_baseline(deck, tab, args.cpu_steps)
        cpu = {"value": P * args.cpu_steps / t_run, "unit": "atom-eq steps/s", "cores": 1, "kind": "port",
               "sample": f"oracle (CPU restatement, one thread) running {args.cpu_steps} NVE steps of this deck "
                         f"after its setup; setup (neighbor build + first force, {t_setup:.1f} s) excluded"}

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
        "data": "synthetic (seeded FCC Cu CAC deck, analytic EAM tabulated as setfl)",
        "config": {"workload": args.workload, "description": desc, "atoms": deck.natoms,
                   "elements": deck.nelements, "atom_equivalents": P, "represented_sites": deck.represented_sites,
                   "skin": SKIN, "dt_ps": DT, "neighbor_mean_full": round(nbar, 2), "sites": info["sites"],
                   "inner_delta": delta, "inner_mean": round(nbar_in, 2), "inner_live_sites": live,
                   "outer_builds_per_step": round(rebuilds, 4), "inner_prunes_per_step": round(prunes, 4)},
        "fp64": {"value": value64, "ms_per_step": 1000.0 * t_64 / args.steps,
                 "speedup_mixed_over_fp64": value / value64},
        "force_error_vs_fp64": {"max_rel": max_rel_err, "rms_rel": rms_rel_err,
                                "energy_rel": abs(emx - e64) / abs(e64)},
        "kernels_us": {k: round(v[0], 3) for k, v in kt.items() if v[1]},
        "kernels_us_fp64": {k: round(v[0], 3) for k, v in kt64.items() if v[1]},
        "roofline": {"bound": "hbm", "kernel": "eam_force (mixed)", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "kernel_us": force_us, "kernel_launches": force_launches,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["bytes_per_launch"] if tr else None,
                     "algorithmic_bytes": bytes_force,
                     "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if tr else None,
                     "valu": valu},
        "nve_drift": drift,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
