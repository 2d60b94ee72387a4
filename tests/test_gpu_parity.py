"""GPU engine vs the oracle and the reference's golden fixtures, through the
C ABI (sis_amd/libsis_gpu.so).

Bars: interaction lists bit-exact (same pair sets; BP candidates in the
reference's order); FP64 forces / energies within 1e-10 relative of the
oracle; base-pair status after the limit MC identical on the same MT words;
trajectories driven by the same MT words within 1e-7 A (positions) and
1e-6 A/tau (velocities) of the reference after the fixture's 40-60 steps
(FP64 summation order differs, so differences grow with the step count); FP32 pair arithmetic within 1e-5 per-bead
relative force error (|dF| / max(|F_ref|, 1e-3 kcal/mol/A))."""
from pathlib import Path

import numpy as np
import pytest

from sis_amd import configs
from sis_amd.engine import Engine
from oracle import corc
from tests.golden.make_golden import force_mc, load

pytestmark = pytest.mark.gpu

FIXTURES = ["hairpin64", "small_pbc", "mixed", "remd4", "remd4mc3", "remd4bd", "remd2ext", "hairpin_ss", "hairpin_anneal", "pbc_resize", "cag47_m1", "remd2rgs", "remd4ion", "remd4rg", "remd4dcf",
            "t2s2hp_sisff", "t2s2hp_ion8", "t2s2hp_temp8", "t2s2hp_ion8c3", "t2s2hp_temp8c2", "remd2x3"]


def _sorted_rows(a):
    return a[np.lexsort(a.T[::-1])] if len(a) else a


def _pair_sets_equal(eng_pairs, ref_pairs):
    return np.array_equal(_sorted_rows(eng_pairs), _sorted_rows(ref_pairs))


def _engine(sysd, d=None, **kw):
    eng = Engine(sysd, parity=True, **kw)
    if "remd_dims" in sysd:
        eng.set_rep_dims(sysd["remd_dims"])
    if "remd_ion" in sysd:
        eng.set_rep_conditions(None, sysd["remd_ion"])
    if "external" in sysd:
        eng.set_external(sysd["external"])
    if "rg_schedule" in sysd:
        eng.set_rg_schedule(*sysd["rg_schedule"])
    if "remd_rg" in sysd:
        eng.set_rep_rg(sysd["remd_rg"])
    if "remd_dcf" in sysd:
        eng.set_rep_dcf(sysd["remd_dcf"], sysd["dcf_dir"])
    if "ss" in sysd:
        eng.set_ss(sysd["ss"])
    if "anneal" in sysd:
        eng.set_anneal(*sysd["anneal"])
    if "resize" in sysd:
        eng.set_box_resize(*sysd["resize"])
    if d is not None:
        for r in range(1, sysd["nrep"] + 1):
            eng.set_rng_tape(r, d["tapes"][r - 1])
        if sysd["nrep"] > 1:
            eng.set_rng_tape(-1, d["tape_rep"])
    return eng


@pytest.mark.parametrize("name", FIXTURES)
def test_lists_match_reference(name):
    sysd, d = load(name)
    print(f"[case] lists {name}", flush=True)
    eng = _engine(sysd, d)
    eng.neighbor_list(0)
    L = eng.lists(1)
    assert _pair_sets_equal(L["wca"], d["wca"])
    assert _pair_sets_equal(L["ele"], d["ele"])
    assert np.array_equal(L["bp"], d["bp"])
    assert np.array_equal(L["bpcoef"], d["bpcoef"])


@pytest.mark.parametrize("name", FIXTURES)
def test_force_energy_bp_state(name):
    sysd, d = load(name)
    print(f"[case] force {name}", flush=True)
    eng = _engine(sysd, d)
    eng.neighbor_list(0)
    f = eng.force(1, force_mc(sysd))
    fr = d["force"]
    assert np.abs(f - fr).max() / np.abs(fr).max() < 1e-10
    b = eng.bp(1)
    assert np.array_equal(b["status"], d["bp_status"].astype(np.uint8))
    assert np.allclose(b["ene"], d["ene_bp"], rtol=1e-9, atol=1e-300)
    e = eng.energy_sumup(1, float(sysd["temps"][0]), 0, 1)
    assert np.allclose(e[1:], d["energies"][1:], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("name", FIXTURES)
def test_trajectory_matches_reference(name):
    """Same MT words as the reference run: initial velocities, initial
    energies with the BP MC, nstep steps (rebuilds, MC, exchanges)."""
    sysd, d = load(name)
    print(f"[case] trajectory {name}", flush=True)
    nex = int(d["nstep_exchange"])
    eng = _engine(sysd, d, exchange_every=nex)
    eng.init_velocities()
    eng.neighbor_list(0)
    for r in range(1, sysd["nrep"] + 1):
        eng.energy_sumup(r, -1.0, 1, 0)
    eng.step(int(d["nstep"]))
    x, v, a = eng.get_state()
    assert np.abs(x - d["md_xyz"]).max() < 1e-7
    assert np.abs(v - d["md_velos"]).max() < 1e-6
    assert np.abs(a - d["md_accels"]).max() < 1e-12   # noise terms: same words, same values
    if sysd["nrep"] > 1:
        assert np.array_equal(eng.labels(), d["md_rep2lab"][: sysd["nrep"]])


def test_twz_lines_match_reference(tmp_path):
    """Tweezers' force-ramp lines (.twz) written during the remd2ext run
    (every 10 steps, from the positions at the start of the step, both
    replicas) against the reference's own lines of the same run: header and
    layout identical (i13, i3, g15.8 fields), values within 1e-6 relative
    (the trajectories agree to 1e-7 A)."""
    sysd, d = load("remd2ext")
    print("[case] twz remd2ext", flush=True)
    eng = _engine(sysd, d, exchange_every=int(d["nstep_exchange"]))
    path = tmp_path / "dev.twz"
    eng.twz_open(path, 10)
    eng.init_velocities()
    eng.neighbor_list(0)
    for r in range(1, sysd["nrep"] + 1):
        eng.energy_sumup(r, -1.0, 1, 0)
    eng.step(int(d["nstep"]))
    eng.twz_close()
    got = path.read_text().splitlines()
    want = str(d["twz_text"]).splitlines()
    assert len(got) == len(want) == 1 + 3 * sysd["nrep"]
    assert got[0] == want[0]
    for g, w in zip(got[1:], want[1:]):
        assert len(g) == len(w)
        gs, ws = g.split(), w.split()
        assert gs[:2] == ws[:2]
        assert np.allclose(np.array(gs[2:], float), np.array(ws[2:], float), rtol=1e-6, atol=1e-6)


def test_condensate_full_size_lists_and_forces():
    """C2 at full size (200 x 47 nt, periodic 2,499 A box): cell-list
    rebuild gives the oracle's pair sets bit-exactly; FP64 forces agree."""
    s = configs.condensate_c2()
    print("[case] C2 lists", flush=True)
    eng = Engine(s, parity=True)
    tape = corc.mt_words(77, 2_000_000)
    eng.set_rng_tape(1, tape)
    o = corc.Oracle(s)
    o.set_tape(1, tape)
    eng.neighbor_list(0)
    o.neighbor_list(1)
    Lg, Lo = eng.lists(1), o.lists(1)
    assert _pair_sets_equal(Lg["wca"], Lo["wca"])
    assert _pair_sets_equal(Lg["ele"], Lo["ele"])
    assert np.array_equal(Lg["bp"], Lo["bp"])
    fg, fo = eng.force(1, 1), o.force(1, 1)
    assert np.abs(fg - fo).max() / np.abs(fo).max() < 1e-10
    assert np.array_equal(eng.bp(1)["status"], o.bp_state(1)[0])


def test_fp32_pair_arithmetic_error():
    """Precision form: FP32 pair math, FP64 state and accumulation."""
    s = configs.cag_system(60, 47, 2.0, 310.0, 5.0, seed=9)
    print("[case] fp32", flush=True)
    e64 = Engine(s, precision=0)
    e32 = Engine(s, precision=1)
    o = corc.Oracle(s)
    o.neighbor_list(1)
    fo = o.force(1, 0)
    for e in (e64, e32):
        e.neighbor_list(0)
    f64 = e64.force(1, 0)
    f32 = e32.force(1, 0)
    den = np.maximum(np.linalg.norm(fo, axis=1), 1e-3)
    rel32 = np.linalg.norm(f32 - fo, axis=1) / den
    rel64 = np.linalg.norm(f64 - fo, axis=1) / den
    print(f"fp32 max {rel32.max():.2e} rms {np.sqrt((rel32**2).mean()):.2e}; fp64 max {rel64.max():.2e}")
    assert rel64.max() < 1e-10
    assert rel32.max() < 1e-5


def test_performance_mode_limit_rule():
    """Counter-RNG limit MC per component: afterwards no nucleotide keeps
    more than max_bp_per_nt pairs, and only pairs formed before the MC
    survive (formed set taken from a run without the limit rule)."""
    s = configs.condensate_c2()
    print("[case] perf MC", flush=True)
    s0 = dict(s)
    s0["params"] = dict(s["params"], nstep_bp_mc=0)
    eng0 = Engine(s0, parity=False, seed=99)
    eng0.neighbor_list(0)
    eng0.force(1, 0)
    pre = eng0.bp(1)["status"].copy()
    eng = Engine(s, parity=False, seed=99)
    eng.neighbor_list(0)
    eng.force(1, 1)
    b = eng.bp(1)
    st = b["status"]
    assert np.array_equal(b["bp"], eng0.bp(1)["bp"])
    assert np.all(st <= pre)
    cnt = np.zeros(s["xyz"].shape[1] + 1, int)
    for row in b["bp"][st == 1]:
        cnt[row[0]] += 1
        cnt[row[1]] += 1
    assert cnt.max() <= 1


def test_performance_mode_limit_distribution():
    """The per-component limit MC samples the same distribution as the
    reference's serial MC: survivor counts and per-candidate survival
    frequencies over 160 independent draws on one C2 configuration."""
    s = configs.condensate_c2()
    print("[case] perf MC distribution", flush=True)
    K = 160
    o = corc.Oracle(s)
    o.neighbor_list(1)
    ser = []
    for k in range(K):
        o.set_tape(1, corc.mt_words(1000 + k, 120_000))
        o.force(1, 1)
        ser.append(o.bp_state(1)[0].astype(float))
    ser = np.array(ser)
    eng = Engine(s, parity=False)
    eng.neighbor_list(0)
    par = []
    for k in range(K):
        eng.set_rng(5000 + k)
        eng.force(1, 1)
        par.append(eng.bp(1)["status"].astype(float))
    par = np.array(par)
    a, b = ser.sum(1), par.sum(1)
    se = np.sqrt(a.var() / K + b.var() / K)
    print(f"survivors serial {a.mean():.2f}+-{a.std():.2f}  components {b.mean():.2f}+-{b.std():.2f}")
    assert abs(a.mean() - b.mean()) < 4.0 * se + 1e-9
    pa, pb = ser.mean(0), par.mean(0)
    mask = (pa + pb) > 0
    assert np.corrcoef(pa[mask], pb[mask])[0, 1] > 0.95
    assert np.abs(pa - pb).max() < 0.25


def test_performance_mode_dynamics_stable():
    s = configs.condensate_c2()
    print("[case] perf steps", flush=True)
    eng = Engine(s, parity=False, seed=7)
    eng.init_velocities()
    eng.neighbor_list(0)
    eng.step(200)
    x, v, a = eng.get_state()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    mass = np.array([{0: 328.212, 1: 305.164, 2: 344.212, 3: 304.182}[q] for q in s["seq"]])
    kb = 1.380649e-23 * 6.02214076e23 / 4184.0
    T = (mass[:, None] * v[0] ** 2).sum() / (3 * len(mass) * kb)
    assert 250.0 < T < 380.0
    e = eng.energy_sumup(1, -1.0, 0, 1)
    assert np.all(np.isfinite(e))


@pytest.mark.parametrize("precision,mc_every", [(0, 1), (1, 1), (0, 3)])
def test_persistent_form_matches_phase_form(precision, mc_every):
    """The persistent single-launch step (grid barriers between phases)
    gives the same trajectory as one launch per phase: same counter-RNG
    streams, same summation order.  mc_every > 1 covers the steps between
    limit-MC steps, which evaluate only the pairs the last MC kept."""
    s = configs.condensate_c2()
    s["params"] = dict(s["params"], nstep_bp_mc=mc_every)
    print(f"[case] persistent vs phases precision={precision} nstep_bp_mc={mc_every}", flush=True)
    out = []
    for mode in ("phases", "persistent"):
        eng = Engine(s, parity=False, mode=mode, seed=31, precision=precision)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(150)
        out.append(eng.get_state())
        assert eng.info()["mode"] == (1 if mode == "phases" else 2)
    for a, b in zip(out[0], out[1]):
        assert np.abs(a - b).max() < 1e-9


@pytest.mark.parametrize("mc_every", [1, 3])
def test_batched_form_matches_phase_form(mc_every):
    """Batched form (one workgroup per replica, in-workgroup rebuilds and
    limit MC, host exchange at the boundaries) follows the per-phase form
    on 8 temperature replicas of a 128-nt chain with exchanges every 10
    steps: same exchange decisions, same trajectories to rounding."""
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    s["params"] = dict(s["params"], nstep_bp_mc=mc_every)
    print(f"[case] batched vs phases nstep_bp_mc={mc_every}", flush=True)
    res = []
    for mode in ("phases", "batched"):
        eng = Engine(s, parity=False, mode=mode, seed=77, exchange_every=10)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(40)
        res.append((eng.get_state(), eng.labels(), eng.info()["mode"]))
    assert res[0][2] == 1 and res[1][2] == 3
    assert np.array_equal(res[0][1], res[1][1])
    for a, b in zip(res[0][0], res[1][0]):
        assert np.abs(a - b).max() < 1e-6


@pytest.mark.parametrize("form", ["batched", "persistent"])
def test_brownian_forms_match_phase_form(form):
    """Ermak-McCammon steps (src/job_md.F90:446-456) in the batched form
    (8 replicas of a 128-nt chain, exchanges every 10 steps) and the
    persistent form (one 2,000-nt condensate replica) follow the per-phase
    form on the same counter-RNG streams.  Velocities are not integrated:
    without exchanges (which rescale them) they stay as initialised."""
    if form == "batched":
        s = configs.remd_c3(nrep=8, n=128, seed=3)
        kw = dict(exchange_every=10)
    else:
        s = configs.condensate_c2()
        kw = {}
    s["params"] = dict(s["params"], integrator=2, dt=0.005)
    print(f"[case] brownian {form} vs phases", flush=True)
    res = []
    for mode in ("phases", form):
        eng = Engine(s, parity=False, mode=mode, seed=41, **kw)
        eng.init_velocities()
        v0 = eng.get_state()[1]
        eng.neighbor_list(0)
        eng.step(60)
        st = eng.get_state()
        if not kw:
            assert np.array_equal(st[1], v0)
        assert np.all(np.isfinite(st[0]))
        res.append((st, eng.labels()))
    assert np.array_equal(res[0][1], res[1][1])
    assert np.abs(res[0][0][0] - res[1][0][0]).max() < 1e-6


def test_box_resize_persistent_form(tmp_path):
    """Box resizing every 20 steps (x shrinks 1 A until 2,496 A, z grows)
    on the 2,000-nt condensate: the persistent form follows the per-phase
    form, and the last DCD frame carries the box the schedule gives
    (original + change x count, clamped at the target).  The frame of step
    100 is saved before that step's resize (the save block precedes the
    resize block, src/job_md.F90:485-596): it has the box of 4 resizes."""
    from sis_amd.engine import read_dcd_file
    s = configs.condensate_c2()
    b0 = np.array(s["params"]["box"], float)
    ch, tg = np.array([-1.0, 0.0, 0.5]), b0 + np.array([-3.0, 0.0, 10.0])
    print("[case] box resize persistent vs phases", flush=True)
    res = []
    for mode in ("phases", "persistent"):
        eng = Engine(s, parity=False, mode=mode, seed=29)
        eng.set_box_resize(20, ch, tg)
        eng.init_velocities()
        eng.neighbor_list(0)
        path = tmp_path / f"{mode}.dcd"
        eng.dcd_open([path], nstep_save=20)
        eng.step(100)
        eng.dcd_close()
        frames, box = read_dcd_file(path, int(np.sum(s["nmp_chain"])))
        assert len(frames) == 5
        want = np.minimum(np.maximum(b0 + ch * 4, np.minimum(b0, tg)), np.maximum(b0, tg))
        assert np.allclose(box, want, rtol=0, atol=1e-9), (box, want)
        res.append(eng.get_state())
    for a, b in zip(res[0], res[1]):
        assert np.abs(a - b).max() < 1e-6


def test_batched_form_512_threads():
    """The batched kernel at 512 threads (8 waves) per replica, an option of
    sis_options.block, follows the per-phase form like the default 256."""
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    print("[case] batched 512 threads vs phases", flush=True)
    res = []
    for mode, blk in (("phases", 0), ("batched", 512)):
        eng = Engine(s, parity=False, mode=mode, seed=77, exchange_every=10, block=blk)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(40)
        res.append((eng.get_state(), eng.labels(), eng.info()["block"]))
    assert res[1][2] == 512 and np.array_equal(res[0][1], res[1][1])
    for a, b in zip(res[0][0], res[1][0]):
        assert np.abs(a - b).max() < 1e-6


def test_ion_exchange_batched_form():
    """Replica exchange over the ionic strength (8 replicas of a 128-nt
    chain at one temperature, 0.02-0.6 M): the batched form follows the
    per-phase form, with the same exchanges."""
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    s["temps"] = np.full(8, float(s["params"]["tempK"]))
    ion = np.geomspace(0.02, 0.6, 8)
    print("[case] ion exchange batched vs phases", flush=True)
    res = []
    for mode in ("phases", "batched"):
        eng = Engine(s, parity=False, mode=mode, seed=31, exchange_every=10)
        eng.set_rep_conditions(None, ion)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(60)
        res.append((eng.get_state(), eng.labels()))
    assert np.array_equal(res[0][1], res[1][1]) and not np.array_equal(res[0][1], np.arange(1, 9))
    for a, b in zip(res[0][0], res[1][0]):
        assert np.abs(a - b).max() < 1e-6


def test_anneal_batched_form():
    """An annealing schedule (temperatures switched at steps 20 and 40,
    force MC at the input kT) in the batched form follows the per-phase
    form on 4 replicas of a 128-nt chain without exchange."""
    s = configs.remd_c3(nrep=4, n=128, seed=3)
    print("[case] anneal batched vs phases", flush=True)
    res = []
    for mode in ("phases", "batched"):
        eng = Engine(s, parity=False, mode=mode, seed=23, exchange_every=0)
        eng.set_anneal([0, 20, 40], [360.0, 330.0, 300.0])
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(30)
        eng.step(30)
        assert eng.info()["mode"] == (1 if mode == "phases" else 3)
        res.append(eng.get_state())
    for a, b in zip(res[0], res[1]):
        assert np.abs(a - b).max() < 1e-6


def test_ss_restriction_batched_form():
    """A secondary-structure restriction in the batched form (8 replicas of a
    128-nt chain): the run follows the per-phase form, and every candidate
    pair is one the list allows."""
    from tests.golden.make_golden import ss_pairs_for
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    pairs = ss_pairs_for(s, seed=2, keep=0.5)
    allowed = {(min(a, b), max(a, b)) for a, b in pairs}
    print("[case] ss restriction batched vs phases", flush=True)
    res = []
    for mode in ("phases", "batched"):
        eng = Engine(s, parity=False, mode=mode, seed=19, exchange_every=10)
        eng.set_ss(pairs)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(60)
        assert eng.info()["mode"] == (1 if mode == "phases" else 3)
        for r in range(1, 9):
            bp = eng.bp(r)["bp"]
            assert len(bp) > 0 and all((int(i), int(j)) in allowed for i, j in bp[:, :2])
        res.append((eng.get_state(), eng.labels()))
    assert np.array_equal(res[0][1], res[1][1])
    for a, b in zip(res[0][0], res[1][0]):
        assert np.abs(a - b).max() < 1e-6


def test_external_terms_in_performance_mode():
    """External terms (stage, tweezers, Rg bias, restraints) with the counter
    RNG: a context created for the batched form steps in the per-phase
    form (the terms are not in the batched kernel), so its run equals a
    per-phase context's exactly; the run stays above the stage wall and its
    external energies stay finite."""
    from tests.golden.make_golden import remd_external
    s = remd_external()
    print("[case] external terms, performance mode", flush=True)
    res = []
    for mode in ("phases", "batched"):
        eng = Engine(s, parity=False, mode=mode, seed=5, exchange_every=10)
        eng.set_external(s["external"])
        assert eng.info()["mode"] == 1
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(200)
        x = eng.get_state()[0]
        e = eng.energy_sumup(1, float(s["temps"][0]), 0, 1)
        assert np.all(np.isfinite(x)) and x[..., 2].min() > 8.5
        assert np.all(np.isfinite(e)) and np.all(e[7:] != 0.0)
        res.append((x, eng.labels()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def _largest_formed_component(eng, irep=1):
    """Pairs in the largest connected set of candidates with negative
    energy (formed before the limit rule; pairs sharing a nucleotide are
    connected), from the engine's last evaluation."""
    b = eng.bp(irep)
    mp = b["bp"][b["ene"] < 0][:, :2]
    par = {}

    def find(x):
        while par.setdefault(x, x) != x:
            par[x] = par[par[x]]
            x = par[x]
        return x
    for a, c in mp:
        ra, rc = find(int(a)), find(int(c))
        if ra != rc:
            par[ra] = rc
    cnt = {}
    for a, _ in mp:
        r = find(int(a))
        cnt[r] = cnt.get(r, 0) + 1
    return max(cnt.values()) if cnt else 0


def _limit_holds(eng, irep=1):
    """After the limit MC no nucleotide has more formed pairs than allowed."""
    b = eng.bp(irep)
    ends = b["bp"][b["status"] > 0][:, :2].ravel()
    return ends.size == 0 or np.bincount(ends).max() <= eng.sysd["params"].get("max_bp_per_nt", 1)


@pytest.mark.parametrize("form", ["persistent", "batched"])
def test_large_components_forms_match_phases(form):
    """Limit MC on components above one wavefront's capacity (more than 128
    formed pairs; a 1,024-nt chain forms components of about 150): the
    large form (one wavefront with the workgroup's whole scratch) in the
    persistent and batched kernels against the per-phase form."""
    nrep = 1 if form == "persistent" else 4
    s = configs.remd_c3(nrep=nrep, n=1024, seed=3)
    print(f"[case] large components {form}", flush=True)
    res = []
    for mode in ("phases", form):
        eng = Engine(s, parity=False, mode=mode, seed=5, exchange_every=0 if nrep == 1 else 10)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(60)
        assert eng.info()["mode"] == {"phases": 1, "persistent": 2, "batched": 3}[mode]
        big = max(_largest_formed_component(eng, k + 1) for k in range(nrep))
        assert 128 < big <= 256, big
        assert all(_limit_holds(eng, k + 1) for k in range(nrep))
        res.append((eng.get_state(), eng.labels()))
    assert np.array_equal(res[0][1], res[1][1])
    tol = 1e-9 if form == "persistent" else 1e-6
    for a, b in zip(res[0][0], res[1][0]):
        assert np.abs(a - b).max() < tol


def test_components_above_bucket_capacity():
    """Components of more than 256 formed pairs (one replica of eight
    1,024-nt chains reaches about 300 after 200 steps) are resolved by the
    unbounded wavefront path (global scratch) in the per-phase and batched
    forms: the limit holds, and both forms give the same exchanges and
    trajectories."""
    s = configs.remd_c3(nrep=8, n=1024, seed=3)
    print("[case] components above the bucket capacity", flush=True)
    res = []
    for mode in ("phases", "batched"):
        eng = Engine(s, parity=False, mode=mode, seed=9, exchange_every=50)
        eng.init_velocities()
        eng.neighbor_list(0)
        big = 0
        for _ in range(6):
            eng.step(50)
            big = max([big] + [_largest_formed_component(eng, k + 1) for k in range(8)])
            assert all(_limit_holds(eng, k + 1) for k in range(8))
        assert big > 256, big
        res.append((eng.get_state(), eng.labels()))
    assert np.array_equal(res[0][1], res[1][1])
    for a, b in zip(res[0][0], res[1][0]):
        assert np.abs(a - b).max() < 1e-6


def test_components_above_bucket_capacity_persistent():
    """The same unbounded path inside the persistent kernel (wave 0 of
    block 0 between two grid barriers): a 2,048-nt chain (60% GC) whose
    largest formed component holds about 320 pairs; persistent and
    per-phase forms agree."""
    rng = np.random.default_rng(11)
    seq = configs.random_seq(rng, 2048, 0.6)
    s = configs._mk("chain2048", [seq], configs.random_coil(rng, 2048),
                    configs.default_params(tempK=300.0, nl_margin=5.0))
    print("[case] components above the bucket capacity, persistent", flush=True)
    out = []
    for mode in ("phases", "persistent"):
        eng = Engine(s, parity=False, mode=mode, seed=9)
        eng.init_velocities()
        eng.neighbor_list(0)
        eng.step(100)
        assert eng.info()["mode"] == (1 if mode == "phases" else 2)
        assert _largest_formed_component(eng) > 256
        assert _limit_holds(eng)
        out.append(eng.get_state()[0])
    assert np.abs(out[0] - out[1]).max() < 1e-9


def test_fp32_and_fp64_ensembles_agree():
    """64 independent copies of the 128-nt chain at 330 K, batched form:
    FP32 and FP64 pair arithmetic give the same distributions of Rg, total
    potential energy and paired fraction at 2,000, 4,000 and 6,000 steps
    (two-sample KS over replicas, every p-value above 0.001)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", Path(__file__).resolve().parents[1] / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    print("[case] fp32 vs fp64 ensembles", flush=True)
    res = bench.ensemble_agreement()
    print(res, flush=True)
    assert res["min_p"] > 1e-3, res


def test_kinetic_energy_matches_oracle():
    """sis_gpu_kinetic (energy_kinetic) against the oracle on random
    velocities of three replicas, rel <= 1e-12 (tree vs sequential sum)."""
    s = configs.remd_c3(nrep=3, n=50, seed=5)
    v = np.random.default_rng(12).normal(0.0, 2.0, size=(3, 50, 3))
    print("[case] kinetic energy", flush=True)
    eng = Engine(s, parity=False, mode="phases")
    eng.set_state(velos=v)
    o = corc.Oracle(s)
    o.set_state(velos=v)
    ek = eng.kinetic(0)
    for r in range(3):
        want = o.kinetic(r + 1)
        assert abs(ek[r] - want) <= 1e-12 * want
        assert abs(eng.kinetic(r + 1) - want) <= 1e-12 * want


def test_batched_dynamics_stable_fp32():
    s = configs.remd_c3(nrep=32, n=128, seed=4)
    print("[case] batched fp32", flush=True)
    eng = Engine(s, parity=False, mode="batched", seed=5, exchange_every=100, precision=1)
    eng.init_velocities()
    eng.neighbor_list(0)
    eng.step(300)
    x, v, a = eng.get_state()
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))
    assert eng.info()["istep"] == 300


def test_fortran_host_matches_c_abi(tmp_path):
    """The Fortran host program (sis_amd/fortran/sis_gpu_host_demo, bound
    through module sis_gpu_iface) gives the same forces, energies and
    trajectory as the same calls made through ctypes."""
    import subprocess
    from sis_amd.abi import params_to_struct
    exe = Path(__file__).resolve().parents[1] / "sis_amd" / "fortran" / "sis_gpu_host_demo"
    if not exe.exists():
        pytest.skip("Fortran host program not built (flang absent at build time)")
    s = configs.small_pbc()
    nsteps, seed = 50, 1234
    nmp = int(np.sum(s["nmp_chain"]))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(bytes(params_to_struct(s["params"])))
        f.write(np.array([s["nchains"], nmp, 1, 0], np.int32).tobytes())
        for k in ("nmp_chain", "moltype", "seq"):
            f.write(np.ascontiguousarray(s[k], np.int32).tobytes())
        f.write(np.ascontiguousarray(s["xyz"], np.float64).tobytes())
        f.write(np.array([nsteps, seed], np.int64).tobytes())
    print("[case] fortran host demo", flush=True)
    r = subprocess.run([str(exe), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    nd = 3 * nmp + 11 + 3 * nmp
    d = np.frombuffer(raw[: 8 * nd], np.float64)
    f_ft, e_ft, x_ft = d[: 3 * nmp].reshape(nmp, 3), d[3 * nmp: 3 * nmp + 11], d[3 * nmp + 11:].reshape(nmp, 3)
    nreb_ft = int(np.frombuffer(raw[8 * nd: 8 * nd + 8], np.int64)[0])

    eng = Engine(s, parity=False, mode="auto", seed=seed)
    eng.neighbor_list(0)
    f_py = eng.force(1, 1)
    e_py = eng.energy_sumup(1, -1.0, 1, 1)
    eng.init_velocities()
    nreb_py = eng.step(nsteps)
    x_py = eng.get_state()[0][0]
    assert np.abs(f_ft - f_py).max() <= 1e-12 * np.abs(f_py).max()
    assert np.allclose(e_ft, e_py, rtol=1e-12, atol=1e-12)
    assert nreb_ft == nreb_py
    assert np.abs(x_ft - x_py).max() < 1e-9


@pytest.mark.parametrize("delta", [1e-6, 1e-5, 1e-4, 1e-3])
def test_fp32_nearly_collinear_base_pair(delta):
    """A G-C candidate whose angle (i-1)-i-j is pi - delta: FP32 pair
    arithmetic must stay finite and within the FP32 force tolerance of
    FP64 (acos and the reference's radicand leave their domains at FP32
    rounding here).  Limit MC off, so the candidate contributes."""
    zig = lambda k: np.array([k * 5.9 * np.sin(np.pi / 3), (k % 2) * 5.9 * np.cos(np.pi / 3), 0.0])
    A = np.array([zig(k) for k in range(6)])
    u = (A[3] - A[2]) / np.linalg.norm(A[3] - A[2])
    c, s_ = np.cos(delta), np.sin(delta)
    w = np.array([c * u[0] - s_ * u[1], s_ * u[0] + c * u[1], 0.0])   # u turned by delta in the plane
    j = A[3] + 13.7 * w
    ez, ep = np.array([0.0, 0.0, 1.0]), np.cross(w, [0.0, 0.0, 1.0])
    Bc = np.array([j + (k - 3) * 5.9 * np.sin(np.pi / 3) * ez + ((k % 2) - 1) * 2.95 * ep for k in range(6)])
    p = configs.default_params(tempK=300.0, nl_margin=5.0, nstep_bp_mc=0)
    s = configs._mk("collinear", ["GGGGGG", "CCCCCC"], np.vstack([A, Bc]), p)
    print(f"[case] fp32 collinear delta={delta}", flush=True)
    f = []
    for prec in (0, 1):
        e = Engine(s, precision=prec)
        e.neighbor_list(0)
        assert e.info()["nbp"] >= 1
        f.append(e.force(1, 0))
    assert np.all(np.isfinite(f[0])) and np.all(np.isfinite(f[1]))
    rel = np.linalg.norm(f[1] - f[0], axis=1) / np.maximum(np.linalg.norm(f[0], axis=1), 1e-3)
    assert rel.max() < 1e-5


def test_new_temperatures_keep_each_replicas_ionic_strength():
    """After exchanges on the ionic-strength ladder (remd4ion) the labels are
    permuted; new temperatures reset the labels to the identity, and each
    replica keeps the ionic strength its label carried: its energies equal
    those of a fresh context given those temperatures and ionic strengths
    directly.  A non-positive value is refused before anything changes."""
    sysd, d = load("remd4ion")
    print("[case] set_rep_conditions after ion exchanges", flush=True)
    eng = _engine(sysd, d, exchange_every=int(d["nstep_exchange"]))
    eng.init_velocities()
    eng.neighbor_list(0)
    for r in range(1, sysd["nrep"] + 1):
        eng.energy_sumup(r, -1.0, 1, 0)
    eng.step(int(d["nstep"]))
    lab = eng.labels()
    assert not np.array_equal(lab, np.arange(1, sysd["nrep"] + 1))   # exchanges happened
    ion = np.asarray(sysd["remd_ion"])[lab - 1]
    x, v, a = eng.get_state()
    with pytest.raises(Exception):
        eng.set_rep_conditions([300.0, -1.0, 300.0, 300.0])
    assert np.array_equal(eng.labels(), lab)
    newT = np.array([301.0, 302.0, 303.0, 304.0])
    eng.set_rep_conditions(newT)
    assert np.array_equal(eng.labels(), np.arange(1, sysd["nrep"] + 1))
    eng.neighbor_list(0)
    ref = Engine(dict(sysd, temps=newT), parity=True)
    ref.set_rep_conditions(newT, ion)
    ref.set_state(x, v, a)
    ref.neighbor_list(0)
    for r in range(1, sysd["nrep"] + 1):
        e1 = eng.energy_sumup(r, -1.0, 0, 0)
        e2 = ref.energy_sumup(r, -1.0, 0, 0)
        assert np.allclose(e1, e2, rtol=1e-12, atol=1e-12), r
