"""The performance forms against the oracle, through the C ABI: the batched
kernel (k_batched, one workgroup per replica), the persistent kernel
(k_persistent, one cooperative launch per rebuild interval) and the job
pool's shared launch (k_jobs), in FP64 and FP32 pair arithmetic.

The oracle draws the same counter-based random numbers as the engine's
performance form (Philox4x32-10 keyed by seed, replica, step and bead,
restated in oracle/sis_oracle.c and pinned by its known-answer vectors), so
initial velocities, noise, rebuilds and exchange decisions are the same
calculation on both sides.  The limit MC is off (nstep_MC = 0): its
performance form resolves conflicts per component with its own draws, and
its distribution is tested in test_gpu_parity.py.

Bars: FP64 positions and velocities within 1e-9 (A, A/tau) of the oracle
after 100-200 steps with in-kernel rebuilds and exchanges, labels equal;
FP32 pair arithmetic within 1e-5; forces from the batched kernel itself,
per-bead relative error |dF| / max(|F_ref|, 1e-3 kcal/mol/A): FP64 <= 1e-10,
FP32 <= 1e-5 (the north star's bound)."""
import time

import numpy as np
import pytest

from sis_amd import configs
from sis_amd.engine import Engine
from sis_amd.serving import DONE, JobPool
from oracle import corc

pytestmark = pytest.mark.gpu

SEED = 4242
TOL = {0: 1e-9, 1: 1e-5}


def _oracle_run(s, seed, nstep, nex):
    o = corc.Oracle(s)
    o.set_counter_rng(seed)
    o.init_velocities()
    nreb, _ = o.md(nstep, nex)
    return o, nreb


def _c3(nrep, margin=1.0):
    s = configs.remd_c3(nrep=nrep, n=128, seed=3, margin=margin)
    s["params"] = dict(s["params"], nstep_bp_mc=0)
    return s


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nrep,nstep,nex,mode", [(16, 200, 50, "batched"), (256, 100, 50, "batched"),
                                                 (16, 200, 50, "phases")])
def test_batched_trajectory_matches_oracle(prec, nrep, nstep, nex, mode):
    """C3's chain (128 nt, 290-380 K ladder) with a 1 A skin: in-kernel
    rebuilds and exchanges every 50 steps; 256 replicas is C3's own ladder
    at the reference's per-dimension maximum.  mode "phases": the per-phase
    kernels replayed from captured graphs, rebuilds between replays."""
    s = _c3(nrep)
    print(f"[case] {mode} prec {prec} R {nrep}", flush=True)
    e = Engine(s, mode=mode, precision=prec, seed=SEED, exchange_every=nex)
    e.init_velocities()
    e.neighbor_list(0)
    e.step(nstep)
    assert e.info()["mode"] == {"batched": 3, "phases": 1}[mode]
    o, nreb = _oracle_run(s, SEED, nstep, nex)
    assert nreb > 0 and e.info()["rebuilds"] > 0
    x, v, _ = e.get_state()
    xo, vo, _ = o.state()
    assert np.abs(x - xo).max() < TOL[prec]
    assert np.abs(v - vo).max() < TOL[prec]
    assert np.array_equal(e.labels(), o.labels())
    assert not np.array_equal(o.labels(), np.arange(1, nrep + 1))   # exchanges were accepted


@pytest.mark.parametrize("prec", [0, 1])
def test_batched_kernel_forces_match_oracle(prec):
    """Forces of the batched kernel's own arithmetic (force-only launch) on
    replicas 1 and 5 of C3's chain."""
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    print(f"[case] batched forces prec {prec}", flush=True)
    e = Engine(s, mode="batched", precision=prec, seed=SEED)
    e.neighbor_list(0)
    o = corc.Oracle(s)
    for r in range(1, 9):
        o.neighbor_list(r)
    for r in (1, 5):
        fe, fo = e.force(r, 0), o.force(r, 0)
        rel = np.linalg.norm(fe - fo, axis=1) / np.maximum(np.linalg.norm(fo, axis=1), 1e-3)
        assert rel.max() <= (1e-10 if prec == 0 else 1e-5), rel.max()


def test_batched_kernel_limit_mc_step():
    """One force evaluation of the batched kernel with the limit MC: the
    energy of every pair formed by energy equals the oracle's (the MC keeps
    it for the pairs it deletes), the surviving pairs are a subset
    of those formed by energy, and no nucleotide is left above
    max_bp_per_nt (src/force_bp_limit_triplet.F90:96-330)."""
    s = configs.remd_c3(nrep=4, n=128, seed=3)
    print("[case] batched limit MC step", flush=True)
    e = Engine(s, mode="batched", precision=0, seed=SEED)
    e.neighbor_list(0)
    # the oracle without the MC: every candidate below the energy cutoff is
    # formed (src/force_bp_triplet.F90 with nstep_MC = 0)
    o = corc.Oracle(dict(s, params=dict(s["params"], nstep_bp_mc=0)))
    for r in range(1, 5):
        o.neighbor_list(r)
    mp = s["params"]["max_bp_per_nt"]
    deleted = 0
    for r in range(1, 5):
        e.force(r, 1)
        o.force(r, 0)
        be = e.bp(r)
        st_o, en_o = o.bp_state(r)
        lst = e.lists(r)["bp"]
        assert np.array_equal(lst, o.lists(r)["bp"])
        fo = st_o.astype(bool)
        assert np.allclose(be["ene"][fo], en_o[fo], rtol=1e-10, atol=1e-12)   # kept for deleted pairs too
        st_e = be["status"].astype(bool)
        assert np.all(~st_e | fo)
        cnt = np.zeros(s["xyz"].shape[1] + 1, int)
        for (i, j) in lst[st_e][:, :2]:
            cnt[i] += 1
            cnt[j] += 1
        assert cnt.max() <= mp
        deleted += int(st_o.astype(bool).sum() - st_e.sum())
    assert deleted > 0   # the limit had work to do


@pytest.mark.parametrize("prec", [0, 1])
def test_persistent_trajectory_matches_oracle(prec):
    """One periodic condensate (40 x 47-nt CAG, 1,880 beads, 1 mM, cell
    list) in the persistent kernel, 1 A skin so that rebuilds end launches."""
    s = configs.cag_system(40, 47, 1.0, 310.0, margin=1.0)
    s["params"] = dict(s["params"], nstep_bp_mc=0)
    print(f"[case] persistent prec {prec}", flush=True)
    e = Engine(s, mode="persistent", precision=prec, seed=SEED)
    e.init_velocities()
    e.neighbor_list(0)
    e.step(80)
    assert e.info()["mode"] == 2
    o, nreb = _oracle_run(s, SEED, 80, 0)
    assert nreb > 0 and e.info()["rebuilds"] > 0
    x, v, _ = e.get_state()
    xo, vo, _ = o.state()
    assert np.abs(x - xo).max() < TOL[prec]
    assert np.abs(v - vo).max() < TOL[prec]


def _job(seed, nch, length, box, tK=300.0, ion=0.15):
    rng = np.random.default_rng(seed)
    seqs = [configs.random_seq(rng, length, 0.55) for _ in range(nch)]
    x = configs._place_chains(rng, seqs, box, clash=6.0)
    p = configs.default_params(tempK=tK, ionic_strength=ion, nl_margin=1.0, pbc=1, box=[box] * 3, nstep_bp_mc=0)
    return configs._mk("pool_job", seqs, x, p)


def test_pool_jobs_match_oracle():
    """Jobs of the serving pool's shared launch (k_jobs), one per
    precision, each against the oracle seeded with the job's seed."""
    jobs = [(_job(1, 4, 40, 160.0), 200, 11, 0), (_job(2, 2, 90, 200.0, tK=340.0, ion=0.3), 150, 12, 1)]
    print("[case] pool jobs vs oracle", flush=True)
    with JobPool(quantum_bead_steps=20000) as pool:
        ids = [pool.submit(j, n, seed=sd, precision=pr) for j, n, sd, pr in jobs]
        res = {}
        t0 = time.time()
        while len(res) < len(ids):
            assert time.time() - t0 < 120, "pool did not finish"
            for i in ids:
                if i not in res:
                    st, tm, x = pool.poll(i, want_xyz=True)
                    if st == DONE:
                        res[i] = (x, tm)
            time.sleep(0.01)
    for (j, n, sd, pr), i in zip(jobs, ids):
        o, _ = _oracle_run(j, sd, n, 0)
        x, tm = res[i]
        assert tm[4] > 0   # rebuilds inside the job
        assert np.abs(x - o.state()[0][0]).max() < TOL[pr]


@pytest.fixture(scope="module")
def c4_system():
    s = configs.condensate_c4()
    s["params"] = dict(s["params"], nstep_bp_mc=0)
    return s


def test_c4_full_size_lists_and_forces(c4_system):
    """C4 at full size (2,000 x 47-nt CAG, 94,000 beads, 4,273.5 A box):
    the cell-list rebuild gives the oracle's lists (WCA / DH pair sets,
    base-pair candidates in the reference's order); FP64 forces within
    1e-10 of the oracle's; FP32 pair arithmetic within the 1e-5 per-bead
    bound (|dF| / max(|F_ref|, 1e-3 kcal/mol/A))."""
    s = c4_system
    print("[case] C4 lists and forces", flush=True)
    o = corc.Oracle(s)
    o.neighbor_list(1)
    fo = o.force(1, 0)
    Lo = o.lists(1)
    e = Engine(s, mode="phases", precision=0, seed=SEED)
    e.neighbor_list(0)
    L = e.lists(1)
    for k in ("wca", "ele"):
        a, b = L[k], Lo[k]
        assert np.array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])]), k
    assert np.array_equal(L["bp"], Lo["bp"]) and np.array_equal(L["bpcoef"], Lo["bpcoef"])
    f64 = e.force(1, 0)
    assert np.abs(f64 - fo).max() / np.abs(fo).max() < 1e-10
    e32 = Engine(s, mode="phases", precision=1, seed=SEED)
    e32.neighbor_list(0)
    f32 = e32.force(1, 0)
    rel = np.linalg.norm(f32 - fo, axis=1) / np.maximum(np.linalg.norm(fo, axis=1), 1e-3)
    print(f"C4 FP32 force error vs oracle: max {rel.max():.3e} rms {np.sqrt((rel ** 2).mean()):.3e}", flush=True)
    assert rel.max() <= 1e-5


@pytest.mark.parametrize("prec", [0, 1])
def test_c4_persistent_trajectory_matches_oracle(c4_system, prec):
    """C4 in the persistent kernel for 20 steps against the oracle on the
    same counter-RNG draws."""
    s = c4_system
    print(f"[case] C4 persistent prec {prec}", flush=True)
    e = Engine(s, mode="persistent", precision=prec, seed=SEED)
    e.init_velocities()
    e.neighbor_list(0)
    e.step(20)
    assert e.info()["mode"] == 2
    o, _ = _oracle_run(s, SEED, 20, 0)
    x, v, _ = e.get_state()
    xo, vo, _ = o.state()
    assert np.abs(x - xo).max() < TOL[prec]
    assert np.abs(v - vo).max() < TOL[prec]


@pytest.mark.parametrize("ladder", ["temperature", "ionic", "grid"])
@pytest.mark.parametrize("block", [256, 128])
def test_device_exchange_equals_host_exchange(ladder, block):
    """Batched replica exchange done on the device (energies at both
    conditions in the batched kernel with the limit MC, Metropolis sweep
    and rescale on the device, segments replayed as a HIP graph) against
    the host's sweep between segments: same labels and the same
    trajectories, on a temperature and on an ionic-strength ladder, MC
    every step, 1 A skin (in-kernel rebuilds); and on a grid of 4
    temperatures x 4 ionic strengths (the pair table of each dimension in
    turn)."""
    s = configs.remd_c3(nrep=16, n=128, seed=3, margin=1.0)
    if ladder == "ionic":
        s["temps"] = np.full(16, 300.0)
    if ladder == "grid":
        s["temps"] = np.repeat([290.0, 310.0, 335.0, 365.0], 4)
    print(f"[case] device vs host exchange {ladder} block {block}", flush=True)
    res = []
    for host in (True, False):
        e = Engine(s, mode="batched", precision=0, seed=SEED, exchange_every=50, block=block, host_exchange=host)
        if ladder == "ionic":
            e.set_rep_conditions(None, np.geomspace(0.02, 0.8, 16))
        if ladder == "grid":
            e.set_rep_dims([4, 4])
            e.set_rep_conditions(None, np.tile([0.02, 0.08, 0.2, 0.8], 4))
        e.init_velocities()
        e.neighbor_list(0)
        e.step(300)
        info = e.info()
        assert info["device_exchanges"] == (0 if host else 6) and info["block"] == block
        res.append((e.get_state(), e.labels(), info["rebuilds"]))
    (xa, va, aa), la, ra = res[0]
    (xb, vb, ab), lb, rb = res[1]
    assert np.array_equal(la, lb)
    assert not np.array_equal(la, np.arange(1, 17))
    assert np.abs(xa - xb).max() < 1e-10 and np.abs(va - vb).max() < 1e-10
    assert ra == rb


def test_autotuner_sweeps_and_restores():
    """sis_gpu_autotune on a batched ladder: 3 workgroup sizes x 3 skins
    timed, the fastest kept, the state (coordinates, velocities, step,
    labels) restored; the run then matches a fresh context created with
    the chosen block and skin."""
    s = configs.remd_c3(nrep=16, n=128, seed=3)
    print("[case] autotuner", flush=True)
    e = Engine(s, mode="batched", precision=1, seed=SEED, exchange_every=50)
    e.init_velocities()
    e.neighbor_list(0)
    x0, v0, a0 = e.get_state()
    rep = e.autotune((3.0, 5.0, 8.0), nsteps=100)
    assert len(rep["trials"]) == 9 and all(t["ms_per_step"] > 0 for t in rep["trials"])
    assert rep["shape"] in (128, 256, 512) and rep["skin"] in (3.0, 5.0, 8.0)
    # the narrowest shape whose fastest skin is within 3% of the fastest
    # trial (times are reported rounded to 1e-5 ms)
    best = min(t["ms_per_step"] for t in rep["trials"])
    fastest = {sh: min(t["ms_per_step"] for t in rep["trials"] if t["shape"] == sh) for sh in (128, 256, 512)}
    assert fastest[rep["shape"]] <= best * 1.03 + 2e-5
    assert all(fastest[sh] > best * 1.03 - 2e-5 for sh in (128, 256, 512) if sh < rep["shape"])
    chosen = [t for t in rep["trials"] if t["shape"] == rep["shape"] and t["skin"] == rep["skin"]]
    assert len(chosen) == 1 and chosen[0]["ms_per_step"] <= fastest[rep["shape"]] + 2e-5
    x1, v1, a1 = e.get_state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1) and np.array_equal(a0, a1)
    assert e.info()["istep"] == 0
    e.step(100)
    s2 = dict(s, params=dict(s["params"], nl_margin=rep["skin"]))
    f = Engine(s2, mode="batched", precision=1, seed=SEED, exchange_every=50, block=rep["shape"])
    f.set_state(x0, v0, a0)
    f.neighbor_list(0)
    f.step(100)
    assert np.abs(e.get_state()[0] - f.get_state()[0]).max() < 1e-9
    assert np.array_equal(e.labels(), f.labels())


def test_autotuner_per_phase_graph_form():
    """sis_gpu_autotune on the per-phase form with device-driven batches of
    steps (a periodic condensate on the cell list): batches replayed from a
    graph (shape 0) or issued directly (shape 1) at every skin, the state
    restored; the run then matches a fresh context with the chosen skin,
    whichever way its steps are issued."""
    s = configs.small_pbc(6, 40, 150.0, seed=8)
    print("[case] autotuner per-phase", flush=True)
    e = Engine(s, mode="phases", precision=1, seed=SEED)
    e.init_velocities()
    e.neighbor_list(0)
    x0, v0, a0 = e.get_state()
    rep = e.autotune((2.0, 4.0), nsteps=64)
    assert [(t["shape"], t["skin"]) for t in rep["trials"]] == [(0, 2.0), (0, 4.0), (1, 2.0), (1, 4.0)]
    assert rep["shape"] in (0, 1) and rep["skin"] in (2.0, 4.0)
    x1, v1, a1 = e.get_state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1) and np.array_equal(a0, a1)
    e.step(64)
    s2 = dict(s, params=dict(s["params"], nl_margin=rep["skin"]))
    f = Engine(s2, mode="phases", precision=1, seed=SEED)
    f.set_state(x0, v0, a0)
    f.neighbor_list(0)
    f.step(64)
    assert np.abs(e.get_state()[0] - f.get_state()[0]).max() < 1e-9


@pytest.mark.parametrize("case", ["chain_mc1", "ladder_mc3", "pbc_cells"])
def test_graph_steps_equal_host_steps(case):
    """The per-phase form replayed from captured graphs (device step count,
    MC flags and skin test, halting for rebuilds) gives the same state as
    the same kernels launched step by step from the host: one chain with MC
    every step (0.3 A skin), a 4-replica ladder with MC every 3 steps and
    exchanges, and a periodic condensate on the cell list (1 A skins), so
    that rebuilds fall inside the replays."""
    if case == "chain_mc1":
        s, nex = configs.remd_c3(nrep=1, n=128, seed=3, margin=0.3), 0
    elif case == "ladder_mc3":
        s, nex = configs.remd_c3(nrep=4, n=128, seed=3, margin=1.0), 25
        s["params"] = dict(s["params"], nstep_bp_mc=3)
    else:
        s, nex = configs.cag_system(40, 47, 1.0, 310.0, margin=1.0), 0
    print(f"[case] graph steps {case}", flush=True)
    out = []
    for no_graph in (True, False):
        e = Engine(s, mode="phases", precision=1, seed=SEED, exchange_every=nex, no_graph=no_graph)
        e.init_velocities()
        e.neighbor_list(0)
        e.step(37)
        e.step(163)
        out.append((e.get_state(), e.labels() if s["nrep"] > 1 else None, e.info()["rebuilds"], e.bp(1)["status"]))
    (xa, va, aa), la, ra, ba = out[0]
    (xb, vb, ab), lb, rb, bb = out[1]
    assert ra == rb and ra > 2
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(aa, ab)
    assert np.array_equal(ba, bb)
    if la is not None:
        assert np.array_equal(la, lb)
