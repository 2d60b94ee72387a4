"""Serving form (job pool, sis_pool_* in include/sis_gpu.h) on the GPU.

Tolerances: a job's result alone and in a mix of other jobs is compared
bit for bit (its arithmetic and random streams depend only on its own
inputs); the pool against the per-phase form of the same single-chain job
agrees to 1e-8 A after 20 steps (lane count per bead, hence FP64 summation
order, differs between the two forms)."""
import time

import numpy as np
import pytest

from sis_amd import configs
from sis_amd.engine import Engine, SisGpuError
from sis_amd.serving import DONE, FAILED, JobPool

pytestmark = pytest.mark.gpu


def _wait(pool, ids, timeout=240.0):
    t0 = time.time()
    out = {}
    while len(out) < len(ids):
        for j in ids:
            if j in out:
                continue
            try:
                st, tm, x = pool.poll(j, want_xyz=True)
            except SisGpuError as e:
                out[j] = (FAILED, None, str(e))
                continue
            if st == DONE:
                out[j] = (st, tm, x)
        assert time.time() - t0 < timeout, "pool did not finish"
        time.sleep(0.01)
    return out


def _job(seed, nch, length, box, tK=300.0, ion=0.15):
    rng = np.random.default_rng(seed)
    seqs = [configs.random_seq(rng, length, 0.55) for _ in range(nch)]
    x = configs._place_chains(rng, seqs, box, clash=6.0)
    p = configs.default_params(tempK=tK, ionic_strength=ion, nl_margin=5.0, pbc=1, box=[box] * 3)
    return configs._mk("pool_job", seqs, x, p)


def test_job_same_alone_and_in_the_mix():
    """One small job (one workgroup) and one large job (> 2,048 beads: the
    per-phase path over the whole GPU), each run alone and then among other
    jobs of both kinds submitted around it: bit-identical results."""
    x_job = _job(1, 4, 40, 160.0)
    big_job = _job(6, 3, 700, 420.0, tK=320.0)
    others = [(_job(2, 2, 90, 200.0, tK=340.0, ion=0.3), 0), (_job(3, 6, 25, 150.0, tK=290.0), 1),
              (_job(4, 1, 300, 300.0, ion=0.05), 0), (_job(7, 2, 1100, 450.0), 1)]
    print("[case] pool alone", flush=True)
    alone = {}
    for name, job, nst in (("x", x_job, 400), ("big", big_job, 60)):
        with JobPool(quantum_bead_steps=20000) as pool:
            a = pool.submit(job, nst, seed=77)
            alone[name] = _wait(pool, [a])[a]
        assert alone[name][0] == DONE
    print("[case] pool mix", flush=True)
    with JobPool(quantum_bead_steps=20000) as pool:
        ids = [pool.submit(others[0][0], 300, seed=5, precision=others[0][1])]
        time.sleep(0.05)
        b = pool.submit(x_job, 400, seed=77)
        g = pool.submit(big_job, 60, seed=77)
        ids += [pool.submit(s, 50 if len(s["seq"]) > 2048 else 250, seed=9 + k, precision=pr)
                for k, (s, pr) in enumerate(others[1:])]
        res = _wait(pool, ids + [b, g])
        st = pool.stats()
    assert all(r[0] == DONE for r in res.values())
    assert np.array_equal(res[b][2], alone["x"][2]), np.abs(res[b][2] - alone["x"][2]).max()
    assert np.array_equal(res[g][2], alone["big"][2]), np.abs(res[g][2] - alone["big"][2]).max()
    assert st["done"] == 6 and st["failed"] == 0
    nmp = [int(np.sum(s["nmp_chain"])) for s, _ in others]
    expect = 300 * nmp[0] + 250 * (nmp[1] + nmp[2]) + 50 * nmp[3] + 400 * int(np.sum(x_job["nmp_chain"])) \
        + 60 * int(np.sum(big_job["nmp_chain"]))
    assert st["bead_steps"] == pytest.approx(expect)


def test_pool_matches_per_phase_form():
    """A single-chain job (candidate order equal in both forms): pool vs
    Engine in the per-phase form, same seed, 20 steps."""
    s = _job(11, 1, 120, 180.0)
    print("[case] pool vs phases", flush=True)
    with JobPool() as pool:
        j = pool.submit(s, 20, seed=123)
        r = _wait(pool, [j])[j]
    eng = Engine(s, parity=False, mode="phases", seed=123)
    eng.neighbor_list(0)
    eng.init_velocities()
    eng.step(20)
    x = eng.get_state()[0][0]
    assert r[0] == DONE
    assert np.abs(r[2] - x).max() < 1e-8


def test_pool_large_component_job_matches_per_phase_form():
    """A 2,000-nt chain (70% GC) run by one workgroup of the shared launch
    forms limit-MC components of more than 128 pairs within 60 steps: the
    workgroup resolves them with its whole scratch, and the job follows the
    per-phase form."""
    rng = np.random.default_rng(8)
    seq = configs.random_seq(rng, 2000, 0.7)
    s = configs._mk("pool_chain", [seq], configs.random_coil(rng, 2000),
                    configs.default_params(tempK=300.0, nl_margin=5.0))
    print("[case] pool large components", flush=True)
    with JobPool() as pool:
        j = pool.submit(s, 60, seed=321)
        r = _wait(pool, [j])[j]
    eng = Engine(s, parity=False, mode="phases", seed=321)
    eng.neighbor_list(0)
    eng.init_velocities()
    eng.step(60)
    b = eng.bp(1)
    assert int((b["ene"] < 0).sum()) > 300
    assert r[0] == DONE
    assert np.abs(r[2] - eng.get_state()[0][0]).max() < 1e-6


def test_pool_mixed_precision_batch_and_bad_job():
    """Eight jobs of different sizes, FP32 and FP64 mixed, plus one job the
    GPU backend refuses (base-pair model 3): the refused one fails with its
    message, the others finish with finite coordinates."""
    rng = np.random.default_rng(8)
    jobs = []
    for k in range(8):
        nch = int(rng.integers(1, 6))
        jobs.append((_job(20 + k, nch, int(rng.integers(20, 120)), 220.0), k % 2))
    bad = _job(99, 2, 30, 120.0)
    bad["params"] = dict(bad["params"], bp_model=3)
    print("[case] pool mixed batch", flush=True)
    with JobPool(quantum_bead_steps=50000) as pool:
        ids = [pool.submit(s, 300, seed=100 + k, precision=pr) for k, (s, pr) in enumerate(jobs)]
        jb = pool.submit(bad, 100, seed=1)
        res = _wait(pool, ids + [jb])
        st = pool.stats()
    assert res[jb][0] == FAILED and "model" in res[jb][2]
    for j in ids:
        assert res[j][0] == DONE and np.all(np.isfinite(res[j][2]))
    assert st["done"] == 8 and st["failed"] == 1
