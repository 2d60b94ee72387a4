#!/usr/bin/env python3
"""Benchmark of the SIS MD step on one MI355X.

Default workload: 200 chains of 47-nt CAG repeat (9,400 beads) in a periodic
2,499.1 A cube (1 mM nucleotides), WCA + Debye-Hueckel (150 mM, 50 A cutoff)
+ triplet base pairing with the limit rule (one pair per nucleotide, MC
every step), GJF-2GJ Langevin at 310 K, dt = 0.05, displacement-triggered
neighbour-list rebuild with a 5 A skin, FP64.  One "step" is one pass of
the job_md loop body over all beads.  Synthetic seeded coordinates.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TAU_FS = 48.888  # reduced time unit (g/mol, A, kcal/mol)


def bytes_per_bead_step(p_wca, p_ele, p_bp):
    """Algorithmic HBM bytes per bead-step of the reference algorithm:
    x, v, beta read + write (144), bonded index lists (36), two int32 per
    WCA / DH half-list pair, seven int32 + two FP64 per BP candidate."""
    return 144.0 + 36.0 + 8.0 * p_wca + 8.0 * p_ele + 44.0 * p_bp


def load_traffic(workload, mode):
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        return d.get(f"{workload}:{mode}")
    except Exception:
        return None


def cpu_baseline(sysd, steps, threads):
    """Reference Fortran (oracle/_ref, OpenMP build) timed on this host:
    job_md for `steps` steps minus a 0-step run (set-up, initial lists and
    energies), i.e. the time-loop cost only."""
    from oracle import refdrv
    if not refdrv.available(omp=True):
        return None
    t1 = refdrv.run(sysd, [("job_md_timed", steps, "t")], omp=True, threads=threads, timeout=900)["t"]
    t0 = refdrv.run(sysd, [("job_md_timed", 0, "t")], omp=True, threads=threads, timeout=900)["t"]
    dt = max(t1 - t0, 1e-9)
    n = int(np.sum(sysd["nmp_chain"])) * int(sysd["nrep"])
    return dict(value=n * steps / dt, unit="bead-steps/s", cores=threads, kind="reference",
                sample=f"{steps} job_md steps of the same system (set-up excluded), OMP_NUM_THREADS={threads}",
                seconds=round(dt, 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4"])
    ap.add_argument("--mode", default="auto", choices=["auto", "phases", "persistent", "batched"])
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(replica / spatial sharding over several GPUs is a design note only).", file=sys.stderr)
        sys.exit(2)

    from sis_amd import configs
    from sis_amd.engine import Engine

    if args.workload == "c2":
        sysd = configs.condensate_c2()
        wname = "C2: 200 x 47-nt CAG, 9,400 beads, PBC 2,499.1 A (1 mM), 150 mM, 310 K, FP64, single job"
    elif args.workload == "c4":
        sysd = configs.condensate_c4()
        wname = "C4: 2,000 x 47-nt CAG, 94,000 beads, PBC 4,273.5 A (2 mM), 150 mM, 310 K"
    else:
        sysd = configs.remd_c3()
        wname = "C3: 256 T-replicas x 128 nt, 290-380 K, exchange every 1,000 steps"
    nmp = int(np.sum(sysd["nmp_chain"]))
    nrep = int(sysd["nrep"])
    print(f"[bench] workload {args.workload}: {nmp} beads x {nrep} replicas; mode {args.mode}", flush=True)

    eng = Engine(sysd, precision=args.precision, parity=False, mode=args.mode, seed=2024,
                 exchange_every=int(sysd.get("exchange_every", 0)))
    eng.init_velocities()
    eng.neighbor_list(0)
    t0 = time.perf_counter()
    eng.step(args.warmup)
    print(f"[bench] warmup {args.warmup} steps {time.perf_counter() - t0:.2f}s", flush=True)
    info = eng.info()
    t0 = time.perf_counter()
    reb = eng.step(args.steps)          # returns after the device work completes
    wall = time.perf_counter() - t0
    dev_ms = eng.last_ms(0)
    dom_ms = eng.last_ms(3)
    dom_n = int(eng.last_ms(4))
    info2 = eng.info()
    x, v, a = eng.get_state()
    assert np.all(np.isfinite(x)), "non-finite coordinates"
    bead_steps = nmp * nrep * args.steps
    value = bead_steps / wall
    p_wca = info2["nwca"] / nmp
    p_ele = info2["nele"] / nmp
    p_bp = info2["nbp"] / nmp
    B = bytes_per_bead_step(p_wca, p_ele, p_bp)
    steps_per_launch = args.steps / max(dom_n, 1)
    achieved = None
    if dom_ms > 0:
        achieved = B * nmp * nrep * steps_per_launch / (dom_ms * 1e-3) / 1e9
    mode_used = {1: "phases", 2: "persistent", 3: "batched"}.get(info2["mode"], str(info2["mode"]))
    ns_day = (args.steps / wall) * 0.05 * TAU_FS * 86400 * 1e-6
    out = {
        "metric": "bead-steps/s batched FP32 vs plain FP64 at force tol; p50/p99 job latency",
        "value": round(value, 1),
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-pair/f64-state",
        "data": "synthetic (seeded random coils)",
        "config": {"workload": wname, "beads": nmp, "replicas": nrep, "mode": mode_used,
                   "rebuilds": int(reb), "p_wca": round(p_wca, 3), "p_ele": round(p_ele, 3),
                   "p_bp": round(p_bp, 3), "bytes_per_bead_step": round(B, 1),
                   "ns_per_day": round(ns_day, 2), "device_ms": round(dev_ms, 3)},
        "roofline": None,
        "cpu_baseline": None,
    }
    if achieved is not None:
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                           "traffic": load_traffic(args.workload, mode_used),
                           "kernel": {"phases": "k_force_int", "persistent": "k_persistent_steps",
                                      "batched": "k_batched_steps"}.get(mode_used, mode_used),
                           "kernel_ms": round(dom_ms, 5), "launches": dom_n,
                           "bead_steps_per_launch": int(nmp * nrep * steps_per_launch)}
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        try:
            out["cpu_baseline"] = cpu_baseline(sysd, args.cpu_steps, threads)
        except Exception as exc:  # keep the GPU line even if the CPU leg fails
            print(f"[bench] cpu baseline failed: {exc}", file=sys.stderr)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
