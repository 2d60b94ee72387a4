#!/usr/bin/env python3
"""Benchmark of the SIS MD step on one MI355X.

Default workload (the metric's "batched FP32" case): 256 temperature
replicas of one 128-nt chain, geometric ladder 290-380 K, non-periodic,
exchange every 1,000 steps; WCA + Debye-Hueckel (150 mM, 50 A cutoff) +
triplet base pairing with the limit rule (MC every step); GJF-2GJ Langevin,
dt = 0.05; displacement-triggered rebuild with a 5 A skin.  Batched form
(one workgroup per replica), FP32 pair arithmetic with FP64 state and
accumulation.  One "step" is one pass of the job_md loop body over every
replica.  Synthetic seeded coordinates.

Beside the headline value the same JSON line carries:
  plain_fp64   - the same chain as one job per launch sequence, FP64,
                 per-phase form (the baseline the batched FP32 form is
                 compared with), and the ratio;
  fp32_force   - per-bead force error of FP32 pair arithmetic vs FP64 on
                 the same lists (|dF| / max(|F|, 1e-3 kcal/mol/A));
  serving      - p50/p99 latency and throughput of a Poisson stream of
                 jobs through the job pool (steps scaled down, stated);
  roofline     - algorithmic HBM bytes of the dominant kernel / its
                 measured launch time;
  cpu_baseline - the reference Fortran (OpenMP build) on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TAU_FS = 48.888  # reduced time unit (g/mol, A, kcal/mol)
METRIC = "bead-steps/s batched FP32 vs plain FP64 at force tol; p50/p99 job latency"


def bytes_per_bead_step(p_wca, p_ele, p_bp):
    """Algorithmic HBM bytes per bead-step of the reference algorithm:
    x, v, beta read + write (144), bonded index lists (36), two int32 per
    WCA / DH half-list pair, seven int32 + two FP64 per BP candidate."""
    return 144.0 + 36.0 + 8.0 * p_wca + 8.0 * p_ele + 44.0 * p_bp


def force_kernel_bytes(p_wca, p_ele):
    """Algorithmic bytes of the per-phase pair/integrator kernel per bead:
    state read + write and the WCA / DH pair indices."""
    return 144.0 + 8.0 * p_wca + 8.0 * p_ele


def load_traffic(workload, mode):
    """Measured HBM bytes per bead-step of the dominant kernel (FETCH_SIZE +
    WRITE_SIZE, calibrated; profiles/traffic.json names the summary each
    figure comes from), or None."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        e = json.loads(f.read_text()).get(f"{workload}:{mode}")
        return float(e["bytes_per_bead_step"]) if e else None
    except Exception:
        return None


def cpu_baseline(sysd, steps, threads, t_setup=None):
    """Reference Fortran (oracle/_ref, OpenMP build) timed on this host:
    job_md for `steps` steps minus a 0-step run (set-up, initial lists and
    energies), i.e. the time-loop cost only.  t_setup: that 0-step time
    when already measured."""
    from oracle import refdrv
    if not refdrv.available(omp=True):
        return None
    nex = int(sysd.get("exchange_every", 0))
    t1 = refdrv.run(sysd, [("job_md_timed", steps, "t")], omp=True, threads=threads, timeout=900,
                    nstep_exchange=nex)["t"]
    if t_setup is None:
        t_setup = refdrv.run(sysd, [("job_md_timed", 0, "t")], omp=True, threads=threads, timeout=900,
                             nstep_exchange=nex)["t"]
    dt = max(t1 - t_setup, 1e-9)
    n = int(np.sum(sysd["nmp_chain"])) * int(sysd["nrep"])
    return dict(value=round(n * steps / dt, 1), unit="bead-steps/s", cores=threads, kind="reference",
                sample=f"{steps} job_md steps of the same workload (set-up excluded), OMP_NUM_THREADS={threads}",
                seconds=round(dt, 3), setup_s=round(t_setup, 3))


def host_cores():
    """(cores this process may use, cores of one socket among them): from
    the affinity mask and /proc/cpuinfo's physical ids; at most 16, the
    host share of one GPU."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    sock = {}
    try:
        cur = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("processor"):
                cur = int(line.split(":")[1])
            elif line.startswith("physical id") and cur is not None:
                sock[cur] = int(line.split(":")[1])
    except OSError:
        pass
    per = {}
    for c in cpus:
        per.setdefault(sock.get(c, 0), []).append(c)
    allc = min(16, len(cpus))
    one = min(allc, max(len(v) for v in per.values()) if per else allc)
    return allc, one


def cpu_baseline_legs(sysd, steps):
    """The reference at 1 core, one socket and all cores of this host's
    share (the 1-core leg runs 1/8 of the steps); the headline figure is
    the all-core leg."""
    allc, sock = host_cores()
    legs, t_setup = [], None
    for th in sorted({1, sock, allc}):
        st = max(1, steps // 8) if th == 1 else steps
        r = cpu_baseline(sysd, st, th, t_setup)
        if r is None:
            return None
        t_setup = r["setup_s"]
        legs.append(dict(threads=th, value=r["value"], steps=st, seconds=r["seconds"]))
        print(f"[bench] cpu baseline {th} thread(s): {r['value']:.3e} bead-steps/s", flush=True)
    out = dict(r)
    out["legs"] = legs
    out["host"] = f"{allc} cores usable, {sock} per socket"
    out["mpi"] = ("not measured: no MPI in this image, so the reference's replica-parallel MPI build "
                  "(PAR_MPI, one replica set per rank) cannot be built or run")
    return out


def single_replica(sysd, r=0):
    s = dict(sysd)
    s["nrep"] = 1
    s["temps"] = np.asarray(sysd["temps"])[r: r + 1]
    s["xyz"] = np.asarray(sysd["xyz"]).reshape(int(sysd["nrep"]), -1, 3)[r: r + 1]
    s["params"] = dict(sysd["params"], tempK=float(s["temps"][0]))
    return s


def plain_fp64_rate(sysd1, steps, warmup):
    """One job, FP64, per-phase form (one launch per phase per step)."""
    from sis_amd.engine import Engine
    eng = Engine(sysd1, precision=0, parity=False, mode="phases", seed=7)
    eng.init_velocities()
    eng.neighbor_list(0)
    eng.step(warmup)
    t0 = time.perf_counter()
    eng.step(steps)
    wall = time.perf_counter() - t0
    n = int(np.sum(sysd1["nmp_chain"]))
    eng.close()
    return dict(value=round(n * steps / wall, 1), unit="bead-steps/s", steps=steps,
                ms_per_step=round(1e3 * wall / steps, 4), form="phases", dtype="f64")


def fp32_force_error(sysd, mode="phases"):
    """FP32 vs FP64 pair arithmetic on the same lists, every replica, limit
    MC off so that every formed pair contributes.  mode "batched": the
    forces come from the batched kernel itself (its force-only launch),
    whose FP64 form the tests pin against the oracle."""
    from sis_amd.engine import Engine
    s = dict(sysd, params=dict(sysd["params"], nstep_bp_mc=0))
    e64 = Engine(s, precision=0, mode=mode)
    e32 = Engine(s, precision=1, mode=mode)
    worst, sq, n, gnum, gden = 0.0, 0.0, 0, 0.0, 0.0
    for e in (e64, e32):
        e.neighbor_list(0)
    for r in range(1, int(s["nrep"]) + 1):
        f64 = e64.force(r, 0)
        f32 = e32.force(r, 0)
        d = np.linalg.norm(f32 - f64, axis=1)
        rel = d / np.maximum(np.linalg.norm(f64, axis=1), 1e-3)
        worst = max(worst, float(rel.max()))
        sq += float((rel ** 2).sum())
        n += rel.size
        gnum += float((d ** 2).sum())
        gden += float((f64 ** 2).sum())
    e64.close()
    e32.close()
    return dict(max=float(f"{worst:.3e}"), rms=float(f"{np.sqrt(sq / n):.3e}"),
                global_l2=float(f"{np.sqrt(gnum / max(gden, 1e-300)):.3e}"), tol=1e-5,
                floor_kcal_mol_A=1e-3, kernel={"batched": "k_batched", "phases": "k_force_lanes"}[mode])


def precision_report(sysd, steps, nsamples=5, seed=2025):
    """The precision config's report: an FP32-pair run and an FP64 run of
    the same system from the same seed (same noise), per-phase form (the
    form C4 runs), for
    `steps` steps; at nsamples evenly spaced times in the second half,
    per-chain Rg and paired fraction (2,000 values each) and E_total.
    Returns the KS tests of the pooled per-chain distributions and the
    relative difference of the mean total energies."""
    from scipy.stats import ks_2samp
    from sis_amd import observables
    from sis_amd.engine import Engine
    engs = [Engine(sysd, precision=p, mode="phases", seed=seed) for p in (0, 1)]
    for e in engs:
        e.init_velocities()
        e.neighbor_list(0)
    half = steps // 2
    every = max(1, half // nsamples)
    samples = [[], []]
    t0 = time.perf_counter()
    for i, e in enumerate(engs):
        e.step(half)
        for k in range(nsamples):
            e.step(every)
            samples[i].append(observables.chain_sample(e))
            print(f"[bench] c4 report: precision {i} sample {k + 1}/{nsamples} "
                  f"({time.perf_counter() - t0:.0f}s)", flush=True)
    for e in engs:
        e.close()
    out = dict(steps=half + every * nsamples, samples=nsamples, seed=seed, form="phases")
    for key in ("rg", "bp_fraction"):
        a = np.concatenate([q[key] for q in samples[0]])
        b = np.concatenate([q[key] for q in samples[1]])
        r = ks_2samp(a, b)
        out[key] = dict(D=round(float(r.statistic), 4), p=round(float(r.pvalue), 4),
                        mean_f64=round(float(a.mean()), 4), mean_f32=round(float(b.mean()), 4), n=int(a.size))
    e64 = np.array([q["e_total"] for q in samples[0]])
    e32 = np.array([q["e_total"] for q in samples[1]])
    out["e_total"] = dict(mean_f64=round(float(e64.mean()), 3), mean_f32=round(float(e32.mean()), 3),
                          rel_diff=float(f"{abs(e32.mean() - e64.mean()) / abs(e64.mean()):.3e}"))
    return out


def ensemble_agreement(nrep=64, n=128, tempK=330.0, steps=(2000, 4000, 6000), seed=3):
    """Ensembles of FP32 and FP64 pair arithmetic: nrep independent copies
    of the C3 chain at one temperature, batched form, same start and
    different noise; at each sample time, two-sample KS tests over the
    replicas on Rg, E_total and the paired fraction.  Reports the smallest
    p-value and every test."""
    from sis_amd import configs, observables
    from sis_amd.engine import Engine
    s = configs.remd_c3(nrep=nrep, n=n, seed=seed)
    s["temps"] = np.full(nrep, tempK)
    s["params"] = dict(s["params"], tempK=tempK)
    engs = [Engine(s, precision=p, mode="batched", seed=100 + p) for p in (0, 1)]
    for e in engs:
        e.init_velocities()
        e.neighbor_list(0)
    tests, done = {}, 0
    for t in steps:
        samp = []
        for e in engs:
            e.step(t - done)
            samp.append(observables.sample(e))
        done = t
        tests[str(t)] = observables.ks_compare(samp[0], samp[1])
    for e in engs:
        e.close()
    pmin = min(v["p"] for d in tests.values() for v in d.values())
    return dict(replicas=nrep, tempK=tempK, steps=list(steps), min_p=pmin, tests=tests)


def _draw_job(args):
    seed, k, max_chains = args
    from sis_amd import configs
    return configs.serving_job(np.random.default_rng((seed, k)), max_chains=max_chains)


def draw_serving_jobs(njobs, max_chains, seed=2026):
    """Jobs drawn like the serving config (1-max_chains chains of 20-400 nt,
    40-70% GC, 280-360 K, 50-500 mM, 0.5-5 mM, log-uniform 1e4-1e6 steps);
    job k from its own stream (seed, k), in worker processes started before
    this process touches the GPU."""
    import multiprocessing as mp
    nproc = max(1, min(16, os.cpu_count() or 1))
    with mp.get_context("spawn").Pool(nproc) as wp:
        return wp.map(_draw_job, [(seed, k, max_chains) for k in range(njobs)])


def serving_run(drawn, step_scale, rate, max_chains, seed=2026, timeout=150.0):
    """Poisson arrivals at `rate` of the drawn jobs, steps times step_scale,
    alternately FP64 and FP32, through the job pool.  Prints progress every
    few seconds; stops waiting after `timeout` s (reported as incomplete)."""
    from sis_amd.serving import DONE, JobPool
    njobs = len(drawn)
    jobs = [(s, max(1, int(round(s["nstep"] * step_scale))), k % 2) for k, s in enumerate(drawn)]
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=njobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    nb = [int(np.sum(s["nmp_chain"])) for s, _, _ in jobs]
    print(f"[bench] serving: {njobs} jobs, beads {min(nb)}-{max(nb)} (mean {int(np.mean(nb))}), steps "
          f"{min(j[1] for j in jobs)}-{max(j[1] for j in jobs)}, rate {rate}/s", flush=True)
    lat, ids = [], []
    incomplete = 0
    with JobPool() as pool:
        t0 = time.perf_counter()
        k = 0
        pend = {}
        t_print = 0.0
        while k < njobs or pend:
            now = time.perf_counter() - t0
            if now - t_print > 3.0:
                t_print = now
                st_ = pool.stats()
                print(f"[bench] serving t={now:.1f}s submitted {k} done {int(st_['done'])} running "
                      f"{int(st_['running'])} queued {int(st_['queued'])} launches {int(st_['launches'])} "
                      f"bead-steps {st_['bead_steps']:.3e}", flush=True)
            if now > timeout:
                incomplete = len(pend) + (njobs - k)
                break
            while k < njobs and arrivals[k] <= now:
                s, nst, pr = jobs[k]
                jid = pool.submit(s, nst, seed=1000 + k, precision=pr)
                pend[jid] = (k, nst * int(np.sum(s["nmp_chain"])))
                ids.append(jid)
                k += 1
            for jid in list(pend):
                st, tm, _ = pool.poll(jid)
                if st == DONE:
                    lat.append(tm[2] - tm[0])
                    pend.pop(jid)
                    pool.forget(jid)
            time.sleep(0.002)
        wall = time.perf_counter() - t0
        stt = pool.stats()
    lat = np.array(lat) if lat else np.array([np.nan])
    return dict(jobs=njobs, incomplete=incomplete, step_scale=step_scale, arrival_rate_per_s=rate,
                max_chains=max_chains,
                p50_s=round(float(np.percentile(lat, 50)), 4), p99_s=round(float(np.percentile(lat, 99)), 4),
                bead_steps_per_s=round(stt["bead_steps"] / wall, 1), wall_s=round(wall, 2),
                gpu_busy_frac=round(stt["gpu_s"] / wall, 3), launches=int(stt["launches"]),
                mix="half FP32, half FP64 (alternating)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="c3", choices=["c2", "c3", "c4", "c5"])
    ap.add_argument("--report-steps", type=int, default=200000,
                    help="c4: steps of the FP32-vs-FP64 distribution report")
    ap.add_argument("--mode", default="auto", choices=["auto", "phases", "persistent", "batched"])
    ap.add_argument("--precision", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip plain FP64, FP32 error and serving legs")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (defaults: 256 threads, input skin)")
    ap.add_argument("--serving-jobs", type=int, default=48)
    ap.add_argument("--serving-scale", type=float, default=2e-3)
    ap.add_argument("--serving-rate", type=float, default=40.0)
    ap.add_argument("--serving-max-chains", type=int, default=64)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device 4 times (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        nbytes = 256 << 20
        x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(4):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                          "launches": 4}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(replica / spatial sharding over several GPUs is a design note only).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c5":
        serving_main(args)
        return

    drawn = None
    if not args.no_extras and args.serving_jobs > 0 and args.workload == "c3":
        t0 = time.perf_counter()
        drawn = draw_serving_jobs(args.serving_jobs, args.serving_max_chains)
        print(f"[bench] drew {len(drawn)} serving jobs in {time.perf_counter() - t0:.1f}s", flush=True)

    from sis_amd import configs
    from sis_amd.engine import Engine

    if args.workload == "c2":
        sysd = configs.condensate_c2()
        wname = "C2: 200 x 47-nt CAG, 9,400 beads, PBC 2,499.1 A (1 mM), 150 mM, 310 K, single job"
        steps, warm, prec, cpu_steps = 4000, 400, 0, 300
    elif args.workload == "c4":
        sysd = configs.condensate_c4()
        wname = "C4: 2,000 x 47-nt CAG, 94,000 beads, PBC 4,273.5 A (2 mM), 150 mM, 310 K"
        steps, warm, prec, cpu_steps = 1000, 100, 1, 20
    else:
        sysd = configs.remd_c3()
        wname = "C3: 256 T-replicas x 128 nt (50% GC), 290-380 K, non-periodic, exchange every 1,000 steps"
        steps, warm, prec, cpu_steps = 20000, 2000, 1, 5000
    steps = args.steps if args.steps is not None else steps
    warm = args.warmup if args.warmup is not None else warm
    prec = args.precision if args.precision is not None else prec
    cpu_steps = args.cpu_steps if args.cpu_steps is not None else cpu_steps
    nmp = int(np.sum(sysd["nmp_chain"]))
    nrep = int(sysd["nrep"])
    print(f"[bench] workload {args.workload}: {nmp} beads x {nrep} replicas; mode {args.mode}; "
          f"precision {prec}", flush=True)

    eng = Engine(sysd, precision=prec, parity=False, mode=args.mode, seed=2024,
                 exchange_every=int(sysd.get("exchange_every", 0)))
    eng.init_velocities()
    eng.neighbor_list(0)
    tune = None
    if not args.no_tune and args.workload in ("c2", "c3", "c4"):
        t0 = time.perf_counter()
        tune = eng.autotune((3.0, 5.0, 8.0), nsteps=200 if args.workload == "c3" else 100)
        print(f"[bench] autotune {time.perf_counter() - t0:.1f}s: shape {tune['shape']} skin {tune['skin']} A; "
              + ", ".join(f"{t['shape']}/{t['skin']:g}: {t['ms_per_step']:.4f} ms" for t in tune["trials"]), flush=True)
    t0 = time.perf_counter()
    eng.step(warm)
    print(f"[bench] warmup {warm} steps {time.perf_counter() - t0:.2f}s", flush=True)
    t0 = time.perf_counter()
    reb = eng.step(steps)          # returns after the device work completes
    wall = time.perf_counter() - t0
    dev_ms = eng.last_ms(0)
    dom_ms = eng.last_ms(3)
    dom_n = int(eng.last_ms(4))
    info2 = eng.info()
    x, v, a = eng.get_state()
    assert np.all(np.isfinite(x)), "non-finite coordinates"
    mode_used0 = {1: "phases", 2: "persistent", 3: "batched"}.get(info2["mode"], "")
    if mode_used0 == "phases" and dom_n == 0:
        # replayed graphs carry no per-kernel events: time the pair /
        # integrator kernel on host-launched steps of the same state
        pe = Engine(sysd, precision=prec, parity=False, mode="phases", seed=2024, no_graph=True,
                    exchange_every=int(sysd.get("exchange_every", 0)))
        pe.set_state(x, v, a)
        pe.neighbor_list(0)
        pe.step(min(steps, 50))
        dom_ms, dom_n = pe.last_ms(3), int(pe.last_ms(4))
        pe.close()
    eng.close()
    print(f"[bench] timed {steps} steps {wall:.3f}s", flush=True)
    bead_steps = nmp * nrep * steps
    value = bead_steps / wall
    p_wca = info2["nwca"] / nmp
    p_ele = info2["nele"] / nmp
    p_bp = info2["nbp"] / nmp
    B = bytes_per_bead_step(p_wca, p_ele, p_bp)
    mode_used = {1: "phases", 2: "persistent", 3: "batched"}.get(info2["mode"], str(info2["mode"]))
    # one k_force_lanes launch per step in the per-phase form; the batched
    # and persistent kernels run whole segments per launch
    steps_per_launch = 1 if mode_used == "phases" else steps / max(dom_n, 1)
    if mode_used == "phases":
        kname, kbytes = "k_force_lanes", force_kernel_bytes(p_wca, p_ele)
    else:
        kname, kbytes = {"persistent": "k_persistent", "batched": "k_batched"}[mode_used], B
    ns_day = (steps / wall) * 0.05 * TAU_FS * 86400 * 1e-6
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(1e3 * wall / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else "f32-pair/f64-state",
        "data": "synthetic (seeded random coils)",
        "config": {"workload": wname, "beads": nmp, "replicas": nrep, "form": mode_used,
                   "rebuilds": int(reb), "p_wca": round(p_wca, 3), "p_ele": round(p_ele, 3),
                   "p_bp": round(p_bp, 3), "bytes_per_bead_step": round(B, 1),
                   "ns_per_day_per_trajectory": round(ns_day, 2), "device_ms": round(dev_ms, 3),
                   "autotune": None if tune is None else {"shape": tune["shape"], "skin_A": tune["skin"],
                                                         "trials": tune["trials"]},
                   "device_exchanges": info2.get("device_exchanges", 0)},
        "roofline": None,
        "cpu_baseline": None,
    }
    if dom_ms > 0:
        achieved = kbytes * nmp * nrep * steps_per_launch / (dom_ms * 1e-3) / 1e9
        tb = load_traffic(args.workload, mode_used)
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                           # counter bytes per bead-step x this run's bead-steps per launch
                           "traffic": None if tb is None else int(tb * nmp * nrep * steps_per_launch),
                           "traffic_bytes_per_bead_step": tb,
                           "kernel": kname, "kernel_ms": round(dom_ms, 5), "launches": dom_n,
                           "bytes_per_bead_step": round(kbytes, 1),
                           "bead_steps_per_launch": int(nmp * nrep * steps_per_launch)}
    if not args.no_extras:
        if nrep > 1 and mode_used == "batched" and prec == 1:
            # the same batched form with FP64 pair arithmetic: the effect of
            # FP32 alone
            s64 = dict(sysd, params=dict(sysd["params"], nl_margin=tune["skin"])) if tune else sysd
            bf = Engine(s64, precision=0, parity=False, mode="batched", seed=2024,
                        exchange_every=int(sysd.get("exchange_every", 0)), block=tune["shape"] if tune else 0)
            bf.init_velocities()
            bf.neighbor_list(0)
            bf.step(1000)
            n64 = min(steps, 5000)
            t0 = time.perf_counter()
            bf.step(n64)
            w64 = time.perf_counter() - t0
            bf.close()
            out["batched_fp64"] = dict(value=round(nmp * nrep * n64 / w64, 1), unit="bead-steps/s", steps=n64,
                                       ms_per_step=round(1e3 * w64 / n64, 5), form="batched", dtype="f64")
            print(f"[bench] batched fp64 {out['batched_fp64']['value']:.3e}", flush=True)
        if nrep > 1:
            s1 = single_replica(sysd)
            pl = plain_fp64_rate(s1, steps=min(steps, 2000), warmup=200)
            pl["workload"] = "replica 1 of the same workload alone"
            out["plain_fp64"] = pl
            out["ratio_vs_plain_fp64"] = round(value / pl["value"], 2)
            print(f"[bench] plain fp64 {pl['value']:.3e}", flush=True)
        out["fp32_force"] = fp32_force_error(sysd, mode="batched" if mode_used == "batched" else "phases")
        print(f"[bench] fp32 force error {out['fp32_force']}", flush=True)
        if args.workload == "c4":
            out["precision_report"] = precision_report(sysd, args.report_steps)
            print(f"[bench] c4 precision report {out['precision_report']}", flush=True)
        if args.workload == "c3":
            out["fp32_ensemble"] = ensemble_agreement()
            print(f"[bench] fp32 vs fp64 ensembles: min KS p {out['fp32_ensemble']['min_p']}", flush=True)
        if drawn:
            out["serving"] = serving_run(drawn, args.serving_scale, args.serving_rate, args.serving_max_chains)
            print(f"[bench] serving {out['serving']}", flush=True)
    if not args.no_cpu and args.workload == "c4" and args.cpu_steps is None:
        # the reference's set-up alone (an all-pairs neighbour list over
        # 94,000 beads) exceeds a bounded sample; run it with --cpu-steps
        out["cpu_baseline_note"] = ("not run by default for C4: the reference's all-pairs neighbour list over "
                                    "94,000 beads alone exceeds a bounded sample (use --cpu-steps)")
    elif not args.no_cpu:
        # the reference loops over replicas one after another, so its rate per
        # bead does not depend on the replica count: the 256-replica workload
        # is sampled with 32 of its replicas (same chain, same 290-380 K span;
        # setting up 256 jump-ahead MT streams alone takes minutes)
        csys, cnote = sysd, ""
        if nrep > 32:
            csys = configs.remd_c3(nrep=32)
            cnote = "; sample: 32 of the 256 replicas (same chain and 290-380 K ladder span)"
        try:
            out["cpu_baseline"] = cpu_baseline_legs(csys, cpu_steps)
            if out["cpu_baseline"] and cnote:
                out["cpu_baseline"]["sample"] += cnote
        except Exception as exc:  # keep the GPU line even if the CPU leg fails
            print(f"[bench] cpu baseline failed: {exc}", file=sys.stderr)
    print(json.dumps(out), flush=True)


def serving_main(args):
    """The serving config as its own line: args.serving_jobs jobs (2,000 by
    default here) drawn as the config states, Poisson arrivals at a fixed
    rate, steps scaled by args.serving_scale (stated in the line), half
    FP32 and half FP64; value = aggregate bead-steps/s, with p50/p99."""
    njobs = args.serving_jobs if args.serving_jobs != 48 else 2000
    t0 = time.perf_counter()
    drawn = draw_serving_jobs(njobs, args.serving_max_chains)
    print(f"[bench] drew {len(drawn)} serving jobs in {time.perf_counter() - t0:.1f}s", flush=True)
    sv = serving_run(drawn, args.serving_scale, args.serving_rate, args.serving_max_chains, timeout=480.0)
    out = {
        "metric": METRIC, "value": sv["bead_steps_per_s"], "unit": "bead-steps/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "mixed f64 / f32-pair", "data": "synthetic (seeded random jobs)",
        "config": {"workload": f"C5: {njobs} jobs, Poisson {args.serving_rate}/s, 1-{args.serving_max_chains} "
                               f"chains x 20-400 nt, steps x {args.serving_scale}"},
        "serving": sv, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
