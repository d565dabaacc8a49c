#!/usr/bin/env python3
"""Benchmark of the CAC force + neighbor hot path on one MI355X.

Workload (default): the 8x-volume mixed-precision EAM CAC configuration:
255,879 core atoms + 13,985 Hex elements of 9^3 sites (2x2x2 Gauss points),
367,759 atom-equivalents, 10.45 M represented sites, Cu-like analytic EAM
tabulated as setfl (Nr = Nrho = 2000), skin 2.0 A, 300 K, dt = 1 fs.

A step is one velocity-Verlet step of the whole model: NVE half kick + drift,
neighbor decision (and rebuild when due), site refresh, EAM density /
embedding / force passes, nodal force assembly, NVE half kick.  The metric is
atom-equivalent steps per second (real atoms + Gauss points) in mixed
precision; the FP64 path is run on the same deck for the force error and the
speedup.  Inputs are resident in HBM before the timed region.

Prints one JSON line.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (box cells, core cells, ncell, description)
    "cac_eam_x8": ((14, 16, 16), 351, 9, "255,879 atoms + 13,985 Hex(9^3) elements, EAM, mixed vs FP64"),
    "cac_eam": ((7, 8, 8), 44, 9, "32,076 atoms + 1,748 Hex(9^3) elements, EAM"),
}


def force_vector(m):
    import numpy as np
    a = m.atoms()[2]
    e = m.elements()["gaussf"].reshape(-1, 3)
    return np.concatenate([a, e])


def timed_run(m, steps, warmup, dt):
    m.run(warmup, dt, thermo=False)
    m.sync()
    t0 = time.perf_counter()
    m.run(steps, dt, thermo=False)
    m.sync()
    return time.perf_counter() - t0


def cpu_baseline(deck, tab, skin, dt, nsteps):
    from oracle import oracle as O
    o = O.Oracle(deck, "eam", skin=skin, eam_tables=tab)
    t0 = time.perf_counter()
    o.setup(eflag=0, vflag=0)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.run(nsteps, dt)
    t_run = time.perf_counter() - t0
    o.close()
    return t_setup, t_run


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cac_eam_x8", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--inner-skin", type=float, default=-1.0,
                    help="prune margin delta of the inner list (A); < 0 takes 0.25 * skin")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import cac_amd
    from cac_amd import decks

    box, core, nc, desc = WORKLOADS[args.workload]
    skin, dt = 2.0, 0.001
    tab = decks.cu_eam_tables()
    deck = decks.cac_eam(box_cells=box, core_cells=core, ncell=nc, jitter=0.02, T=300.0, seed=2024)
    P = deck.atom_equivalents

    # FP64 and mixed models on the same deck
    m64 = cac_amd.Model(deck, "eam", skin=skin, eam_tables=tab, precision=cac_amd.FP64)
    m64.setup()
    f64 = force_vector(m64)
    e64, _ = m64.energy()
    mmx = cac_amd.Model(deck, "eam", skin=skin, eam_tables=tab, precision=cac_amd.MIXED)
    if args.inner_skin >= 0:
        m64.set_inner_skin(args.inner_skin)
        mmx.set_inner_skin(args.inner_skin)
    mmx.setup()
    fmx = force_vector(mmx)
    emx, _ = mmx.energy()
    df = np.linalg.norm(fmx - f64, axis=1)
    fmag = np.linalg.norm(f64, axis=1)
    max_rel_err = float(df.max() / fmag.max())
    rms_rel_err = float(np.sqrt(np.mean(df ** 2)) / np.sqrt(np.mean(fmag ** 2)))
    info = mmx.neigh_info()
    nbar = info["entries"] / info["rows"]
    _, nbar_in, delta = mmx.inner_stats()

    b0, ib0 = mmx.neigh_stats()[0], mmx.inner_stats()[0]
    t_mx = timed_run(mmx, args.steps, args.warmup, dt)
    t_64 = timed_run(m64, args.steps, args.warmup, dt)
    nsteps_mx = args.steps + args.warmup
    rebuilds = (mmx.neigh_stats()[0] - b0) / nsteps_mx
    prunes = (mmx.inner_stats()[0] - ib0) / nsteps_mx
    value = P * args.steps / t_mx
    value64 = P * args.steps / t_64

    # per-kernel durations from hipEvents on the context stream (separate short run)
    mmx.enable_timing(True)
    mmx.run(args.profile_steps, dt, thermo=False)
    names = ("eam_density", "eam_embed", "eam_force", "site_refresh", "nodef", "nve", "neigh_build", "inner_build")
    kt = {k: mmx.kernel_time(k) for k in names}
    mmx.enable_timing(False)
    m64.enable_timing(True)
    m64.run(args.profile_steps, dt, thermo=False)
    kt64 = {k: m64.kernel_time(k) for k in ("eam_density", "eam_embed", "eam_force")}
    m64.enable_timing(False)

    # dominant kernel: EAM force pass (mixed).  Algorithmic bytes per launch:
    # per row its site index, point index and inner-list length (12) and the
    # FP64 force write (24); 4 bytes per inner-list entry; each site record
    # (x, y, z, F' as FP32, 16 bytes) read once.
    force_us = kt["eam_force"][0]
    bytes_force = P * (36.0 + 4.0 * nbar_in) + 16.0 * info["sites"]
    achieved = bytes_force / (force_us * 1e-6) / 1e9
    peak = 8000.0

    cpu = None
    if not args.no_cpu_baseline:
        t_setup, t_run = cpu_baseline(deck, tab, skin, dt, args.cpu_steps)
        cpu = {"value": P * args.cpu_steps / t_run, "unit": "atom-eq steps/s", "cores": 1, "kind": "port",
               "sample": f"oracle (CPU restatement, one thread) running {args.cpu_steps} NVE steps of this deck "
                         f"after setup; setup (neighbor build + first force) {t_setup:.1f} s excluded"}

    line = {
        "metric": "atom-steps/s (mixed vs FP64 @ <1e-5 rel force err)",
        "value": value,
        "unit": "atom-eq steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_mx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms / fp64 accumulation",
        "data": "synthetic (seeded FCC Cu CAC deck, analytic EAM tabulated as setfl)",
        "config": {"workload": args.workload, "description": desc, "atoms": deck.natoms,
                   "elements": deck.nelements, "atom_equivalents": P, "represented_sites": deck.represented_sites,
                   "skin": skin, "dt_ps": dt, "neighbor_mean_full": round(nbar, 2), "sites": info["sites"],
                   "inner_delta": delta, "inner_mean": round(nbar_in, 2),
                   "outer_builds_per_step": round(rebuilds, 4), "inner_prunes_per_step": round(prunes, 4)},
        "fp64": {"value": value64, "ms_per_step": 1000.0 * t_64 / args.steps,
                 "speedup_mixed_over_fp64": value / value64},
        "force_error_vs_fp64": {"max_rel": max_rel_err, "rms_rel": rms_rel_err,
                                "energy_rel": abs(emx - e64) / abs(e64)},
        "kernels_us": {k: round(v[0], 3) for k, v in kt.items() if v[1]},
        "kernels_us_fp64": {k: round(v[0], 3) for k, v in kt64.items()},
        "roofline": {"bound": "hbm", "kernel": "eam_force (mixed)", "achieved": achieved, "peak": peak,
                     "unit": "GB/s", "frac": achieved / peak, "traffic": None,
                     "bytes_per_launch": bytes_force},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
