"""Device-fed trajectory output: frames written during sis_gpu_step (pinned
staging buffers, background writer) are byte-identical to the host writer
applied to state snapshots of the same trajectory taken between shorter
sis_gpu_step calls."""
from pathlib import Path

import numpy as np
import pytest

from sis_amd import configs
from sis_amd.engine import Engine, write_dcd_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["batched", "pbc_phases"])
def test_dcd_frames_during_steps(tmp_path, case):
    if case == "batched":
        s = configs.remd_c3(nrep=4, n=128, seed=3)
        kw = dict(mode="batched", precision=1, exchange_every=50)
        fix_com, box = 0, None
    else:
        s = configs.small_pbc()
        kw = dict(mode="phases", precision=0)
        fix_com, box = 1, s["params"]["box"]
    R = int(s["nrep"])
    print(f"[case] dcd {case}", flush=True)
    eng = Engine(s, seed=41, **kw)
    eng.init_velocities()
    eng.neighbor_list(0)
    paths = [tmp_path / f"dev_{r}.dcd" for r in range(R)]
    eng.dcd_open(paths, fix_com_origin=fix_com, nstep_save=25)
    eng.dcd_write()
    eng.step(60)
    eng.step(40)
    eng.dcd_close()
    ref = Engine(s, seed=41, **kw)
    ref.init_velocities()
    ref.neighbor_list(0)
    frames = [ref.get_state()[0]]
    for _ in range(4):
        ref.step(25)
        frames.append(ref.get_state()[0])
    frames = np.array(frames)   # (5, R, N, 3)
    for r in range(R):
        want = tmp_path / f"host_{r}.dcd"
        write_dcd_file(want, frames[:, r], box=box, fix_com_origin=fix_com)
        assert paths[r].read_bytes() == want.read_bytes(), r


@pytest.mark.parametrize("mode", ["batched", "phases"])
def test_restart_continues_the_run(mode):
    """A run stopped after two exchanges and restarted in a new context from
    its state, step count and labels (the restart-file blocks) follows the
    uninterrupted run: same exchanges, trajectories equal to rounding (the
    restarted context rebuilds its lists at the restart, which changes the
    neighbour order and hence the FP64 summation order)."""
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    print(f"[case] restart {mode}", flush=True)
    kw = dict(mode=mode, seed=77, exchange_every=20, precision=0)
    a = Engine(s, **kw)
    a.init_velocities()
    a.neighbor_list(0)
    a.step(80)
    xa, va, aa = a.get_state()
    b = Engine(s, **kw)
    b.init_velocities()
    b.neighbor_list(0)
    b.step(40)
    x, v, acc = b.get_state()
    lab, istep = b.labels(), b.info()["istep"]
    c = Engine(s, **kw)
    c.set_labels(lab)
    c.set_state(x, v, acc)
    c.set_step(istep)
    c.step(40)
    xc, vc, _ = c.get_state()
    assert c.info()["istep"] == 80
    assert np.array_equal(a.labels(), c.labels())
    assert np.abs(xa - xc).max() < 1e-6
    assert np.abs(va - vc).max() < 1e-5


@pytest.mark.parametrize("mode", ["batched", "phases"])
def test_restart_files_continue_the_run(tmp_path, mode):
    """Restart files written by sis_gpu_step every 40 steps (after that
    step's exchange, src/job_md.F90:664-666) hold the state, labels and step
    of the run at step 40 exactly; a new context that reads them
    (sis_gpu_rst_read) starts from that state and, after 60 more steps,
    follows the uninterrupted run (same exchanges; trajectories equal to
    rounding, since the lists are rebuilt at the restart).  The restart
    falls after an even number of exchanges: the exchange pair table is not
    in the file and restarts from its first half (src/var_replica.F90:85),
    as the uninterrupted run's does after an even count."""
    from sis_amd.engine import read_rst_file
    s = configs.remd_c3(nrep=8, n=128, seed=3)
    R, N = int(s["nrep"]), int(np.sum(s["nmp_chain"]))
    print(f"[case] restart files {mode}", flush=True)
    kw = dict(mode=mode, seed=77, exchange_every=20, precision=0)
    a = Engine(s, **kw)
    a.init_velocities()
    a.neighbor_list(0)
    a.step(100)
    xa, va, _ = a.get_state()
    b = Engine(s, **kw)
    b.init_velocities()
    b.neighbor_list(0)
    paths = [tmp_path / f"run_{r + 1:04d}.rst" for r in range(R)]
    b.rst_open(paths, nstep_save_rst=40)
    b.step(25)
    b.step(15)
    xb, vb, ab = b.get_state()
    lab = b.labels()
    b.step(30)    # no multiple of 40 in 41..70: the files stay those of step 40
    b.rst_close()
    for r in range(R):
        f = read_rst_file(paths[r], N, R)
        assert f["istep"] == 40 and f["grep"] == r + 1 and f["seed"] == 77
        assert np.array_equal(f["xyz"], xb[r]) and np.array_equal(f["velos"], vb[r])
        assert np.array_equal(f["accels"], ab[r])
        assert f["blocks"] == ([1, 3, 4, 5, 6, 11] if r == 0 else [1, 3, 4, 5, 11])
    assert np.array_equal(read_rst_file(paths[0], N, R)["rep2lab"], lab)
    c = Engine(s, **kw)
    c.rst_read(paths)
    assert c.info()["istep"] == 40 and np.array_equal(c.labels(), lab)
    xc0, vc0, ac0 = c.get_state()
    assert np.array_equal(xc0, xb) and np.array_equal(vc0, vb) and np.array_equal(ac0, ab)
    c.step(60)
    xc, vc, _ = c.get_state()
    assert np.array_equal(a.labels(), c.labels())
    assert np.abs(xa - xc).max() < 1e-6
    assert np.abs(va - vc).max() < 1e-5


def test_restart_files_carry_anneal_and_box(tmp_path):
    """An annealed run with box resizing, stopped at step 50 and restarted
    from its restart files: the ANNEAL block (entry 2 of the 0/30/70
    schedule) and the PBC block (the resized box) bring the new context to
    the uninterrupted run's conditions, which it then follows to step 100
    (src/job_md.F90:207-230, src/pbc.F90:35-45)."""
    from sis_amd.engine import read_rst_file
    s = configs.small_pbc(4, 30, 100.0, seed=5)
    N = int(np.sum(s["nmp_chain"]))
    sched = ([0, 30, 70], [330.0, 310.0, 290.0])
    print("[case] restart with anneal and box resize", flush=True)

    def fresh():
        e = Engine(s, mode="phases", seed=13, precision=0)
        e.set_anneal(*sched)
        e.set_box_resize(25, [-0.5, 0.0, 0.5])
        return e

    a = fresh()
    a.init_velocities()
    a.neighbor_list(0)
    a.step(100)
    b = fresh()
    b.init_velocities()
    b.neighbor_list(0)
    path = tmp_path / "run_0001.rst"
    b.rst_open([path], nstep_save_rst=50)
    b.step(50)
    b.rst_close()
    f = read_rst_file(path, N, 1)
    assert 2 in f["blocks"] and 10 in f["blocks"] and f["istep"] == 50
    assert np.allclose(f["box"], [99.0, 100.0, 101.0])   # resized at steps 25 and 50
    c = Engine(s, mode="phases", seed=13, precision=0)
    c.rst_read([path])
    c.set_anneal(*sched)                         # entry 2 from the ANNEAL block
    c.set_box_resize(25, [-0.5, 0.0, 0.5])       # counts from the restart's box
    c.step(50)
    xa, xc = a.get_state()[0], c.get_state()[0]
    assert np.abs(xa - xc).max() < 1e-6


def test_restart_file_with_bad_block_is_refused(tmp_path):
    """A restart file whose STEP block has the wrong size, or whose REPLICA
    block labels are not a permutation, is refused, and the context keeps
    its state, step and labels (nothing is applied before every file has
    been read and checked)."""
    from sis_amd.engine import SisGpuError
    import struct
    s = configs.remd_c3(nrep=2, n=30, seed=3)
    print("[case] restart file with a bad block", flush=True)
    e = Engine(s, mode="phases", seed=3, precision=0)
    e.init_velocities()
    e.neighbor_list(0)
    e.step(20)
    paths = [tmp_path / "r_0001.rst", tmp_path / "r_0002.rst"]
    e.rst_open(paths)
    e.rst_write()
    e.rst_close()
    x0, v0, _ = e.get_state()
    lab0 = e.labels().copy()
    good = [p.read_bytes() for p in paths]

    def blocks(b):
        out, p = [], 0
        while p < len(b):
            i, n = struct.unpack_from("<ii", b, p)
            out.append((i, b[p + 8:p + 8 + n]))
            p += 8 + n
        return out

    def join(bl):
        return b"".join(struct.pack("<ii", i, len(v)) + v for i, v in bl)

    bad_step = join([(i, v + b"\0\0\0\0") if i == 1 else (i, v) for i, v in blocks(good[0])])
    rep = [(i, v) for i, v in blocks(good[0])]
    bad_lab = join([(i, struct.pack("<iiiii", 2, 1, 1, 2, 1)) if i == 6 else (i, v) for i, v in rep])
    for bad in (bad_step, bad_lab):
        paths[0].write_bytes(bad)
        with pytest.raises(SisGpuError):
            e.rst_read(paths)
        x1, v1, _ = e.get_state()
        assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
        assert np.array_equal(e.labels(), lab0) and e.info()["istep"] == 20


@pytest.mark.parametrize("case", ["batched", "pbc_phases"])
def test_bp_records_during_steps(tmp_path, case):
    """Base-pair records (.bpall: every formed pair; the E <= -kT filter of
    .bp is checked on the host writer) written at the save steps equal the
    formed pairs, in list order, read back from an identical run between
    shorter sis_gpu_step calls; the first set follows an energy
    evaluation, as the initial write_bp does."""
    from tests.test_output import _read_bp
    if case == "batched":
        s = configs.remd_c3(nrep=4, n=128, seed=3)
        kw = dict(mode="batched", precision=1)
    else:
        s = configs.small_pbc()
        kw = dict(mode="phases", precision=0)
    R = int(s["nrep"])
    print(f"[case] bp output {case}", flush=True)
    eng = Engine(s, seed=43, **kw)
    eng.init_velocities()
    eng.neighbor_list(0)
    for r in range(R):
        eng.energy_sumup(r + 1)
    paths = [tmp_path / f"dev_{r}.bp" for r in range(R)]
    eng.bp_open(paths, kind=1, nstep_save=25)
    eng.bp_write()
    eng.step(60)
    eng.step(40)
    eng.bp_close()
    ref = Engine(s, seed=43, **kw)
    ref.init_velocities()
    ref.neighbor_list(0)
    for r in range(R):
        ref.energy_sumup(r + 1)
    want = [[] for _ in range(R)]
    for k in range(5):
        if k:
            ref.step(25)
        for r in range(R):
            b = ref.bp(r + 1)
            keep = b["status"] > 0
            want[r].append([(int(i), int(j), float(np.float32(e)))
                            for (i, j), e in zip(b["bp"][keep][:, :2], b["ene"][keep])])
    total = 0
    for r in range(R):
        got = _read_bp(paths[r])
        assert got == want[r], r
        total += sum(len(x) for x in got)
    assert total > 0


@pytest.mark.parametrize("case", ["batched", "pbc_phases"])
def test_out_lines_during_steps(tmp_path, case):
    """The initial line (sis_gpu_out_write: energies with a fresh limit MC)
    and the energy lines written at the save steps equal the host line
    writer applied to the kinetic and potential energies (step's base-pair
    state, sis_gpu_energy with bp_energy_valid) of an identical run read
    between shorter sis_gpu_step calls."""
    from sis_amd.engine import write_out_file
    if case == "batched":
        s = configs.remd_c3(nrep=4, n=128, seed=3)
        kw = dict(mode="batched", precision=1)
    else:
        s = configs.small_pbc()
        kw = dict(mode="phases", precision=0)
    R = int(s["nrep"])
    print(f"[case] out lines {case}", flush=True)
    eng = Engine(s, seed=47, **kw)
    eng.init_velocities()
    eng.neighbor_list(0)
    paths = [tmp_path / f"dev_{r}.out" for r in range(R)]
    eng.out_open(paths, nstep_save=25)
    eng.out_write()
    eng.step(60)
    eng.step(40)
    eng.out_close()
    ref = Engine(s, seed=47, **kw)
    ref.init_velocities()
    ref.neighbor_list(0)
    rows = [[] for _ in range(R)]
    ek0, lab0 = ref.kinetic(0), ref.labels()
    for r in range(R):
        e = ref.energy_sumup(r + 1, bp_mc=1, bp_energy_valid=0)
        rows[r].append((0, lab0[r], float(s["temps"][lab0[r] - 1]), ek0[r], e))
    for k in range(4):
        ref.step(25)
        ek = ref.kinetic(0)
        lab = ref.labels()
        for r in range(R):
            e = ref.energy_sumup(r + 1, bp_mc=1, bp_energy_valid=1)
            rows[r].append((25 * (k + 1), lab[r], float(s["temps"][lab[r] - 1]), ek[r], e))
    ele = bool(s["params"].get("ele_on", 1))
    for r in range(R):
        want = tmp_path / f"host_{r}.out"
        st, lb, tk, ek, en = zip(*rows[r])
        write_out_file(want, st, lb, tk, ek, np.array(en), replica=R > 1, ele=ele)
        assert paths[r].read_text() == want.read_text(), r


def test_cag47_dcd_energies_on_device():
    """job_dcd on the device over the 100 OpenMM frames of CAG x 47 (NHT22,
    model-1 base pairs): per-frame energies equal the oracle's (1e-9) and
    the OpenMM columns within the tolerances of the oracle's own test."""
    from oracle import corc
    from tests.golden.make_golden import CAG47_DIR
    from tests.test_oracle import CAG47_COLS, cag47_frames_and_reference
    s, frames, ref = cag47_frames_and_reference()
    print("[case] cag47 dcd energies", flush=True)
    eng = Engine(s, mode="phases", precision=0)
    E = eng.dcd_energies(CAG47_DIR / "CAG47.dcd")
    assert E.shape == (100, 11)
    o = corc.Oracle(s)
    o.analysis_lists(1)
    for k in (0, 37, 99):
        o.set_xyz(frames[k][None])
        eo = o.energy(1, 295.15, 0, 0)
        assert np.allclose(E[k], eo, rtol=1e-9, atol=1e-9), k
    for name, ci, ei, rtol, atol in CAG47_COLS:
        assert np.allclose(E[:, ei], ref[:, ci], rtol=rtol, atol=atol), name
    # every column, E_bp included, against the reference build's job_dcd
    want = np.load(Path(__file__).parent / "golden" / "cag47_dcd.npz")["energies"]
    assert np.allclose(E, want, rtol=1e-9, atol=1e-9)


def test_dcd_split_on_device(tmp_path):
    """job_dcd with the limit MC on a duplex of two chains and a circular
    chain (the dcd_split fixture, written by the reference build): per-frame
    energies within 1e-10 of the reference's on the same MT words; the
    base-pair candidates are the reference's list_bp entries within the
    neighbour cutoff, in list_bp's (imp, jmp) order; the pairs formed after
    the last frame's limit MC are the reference's."""
    from sis_amd.engine import write_dcd_file
    from tests.golden.make_golden import load
    sysd, d = load("dcd_split")
    print("[case] dcd_split energies", flush=True)
    eng = Engine(sysd, parity=True, mode="phases")
    eng.set_rng_tape(1, d["tapes"][0])
    p = tmp_path / "split.dcd"
    write_dcd_file(p, d["frames"].astype(np.float64))
    E = eng.dcd_energies(p)
    assert E.shape == (len(d["frames"]), 11)
    assert np.allclose(E, d["energies"], rtol=1e-10, atol=1e-10)
    b = eng.bp(1)
    ref_rows = [tuple(r) for r in d["bp"]]
    pos = [ref_rows.index(tuple(r)) for r in b["bp"]]
    assert pos == sorted(pos)   # a subsequence of list_bp's order
    assert np.allclose(b["bpcoef"], d["bpcoef"][pos])
    formed = {ref_rows[i][:2] for i in np.flatnonzero(d["bp_status"][-1])}
    assert {tuple(r[:2]) for r in b["bp"][b["status"] > 0]} == formed and formed
    # list_exv pairs the circular chain's ends (beads 28 and 43), the MD
    # job's neighbour list does not; those lists come back with the next step
    ends = (28, 43)
    assert ends in {tuple(r) for r in eng.pairs(1, 0)}
    eng.neighbor_list(0)
    assert ends not in {tuple(r) for r in eng.pairs(1, 0)}
    run = Engine(sysd, seed=3, mode="phases")
    run.analysis_lists()
    assert ends in {tuple(r) for r in run.pairs(1, 0)}
    run.init_velocities()
    run.step(1)
    assert ends not in {tuple(r) for r in run.pairs(1, 0)}


def test_dcd_energy_analysis(tmp_path):
    """job_dcd: energies of the frames of a DCD file equal those of the same
    (float32-rounded) coordinates set by hand, and the .out lines are those
    of the host writer (frame number, T, Ekin = 0)."""
    from sis_amd.engine import write_dcd_file, write_out_file
    s = configs.small_pbc()
    print("[case] dcd energies", flush=True)
    eng = Engine(s, seed=5, mode="phases")
    eng.init_velocities()
    eng.neighbor_list(0)
    frames = []
    for _ in range(4):
        eng.step(30)
        frames.append(eng.get_state()[0][0])
    frames = np.array(frames)
    p = tmp_path / "t.dcd"
    write_dcd_file(p, frames, box=s["params"]["box"])
    ana = Engine(s, seed=5, mode="phases")
    e = ana.dcd_energies(p, out_path=tmp_path / "t.out")
    assert e.shape == (4, 11)
    chk = Engine(s, seed=5, mode="phases")
    want = []
    for f in frames.astype(np.float32).astype(np.float64):
        chk.set_state(f[None])
        chk.analysis_lists()
        want.append(chk.energy_sumup(1))
    assert np.array_equal(e, np.array(want))
    host = tmp_path / "h.out"
    write_out_file(host, np.arange(1, 5), np.ones(4), np.full(4, s["params"]["tempK"]), np.zeros(4), e,
                   replica=False, ele=True)
    assert (tmp_path / "t.out").read_text() == host.read_text()


def test_check_force_on_device(tmp_path):
    """job_check_force on the device (h = 1e-3, limit MC off as in the
    oracle's own finite-difference pin): central differences of the total
    energy agree with the forces to 1e-4 relative (floor 1 kcal/mol/A),
    and the report has the reference's two lines per bead."""
    s = configs.small_pbc(3, 20, 90.0, seed=3)
    s["params"] = dict(s["params"], nstep_bp_mc=0)
    print("[case] check_force", flush=True)
    eng = Engine(s, mode="phases")
    eng.neighbor_list(0)
    fe, ff = eng.check_force(1, 1e-3, out_path=tmp_path / "cf.out")
    rel = np.abs(fe - ff) / np.maximum(1.0, np.abs(ff))
    assert rel.max() < 1e-4, rel.max()
    lines = (tmp_path / "cf.out").read_text().splitlines()
    assert len(lines) == 2 * ff.shape[0]
    assert lines[0].startswith("imp=     1  f_energy") and lines[1].startswith("f_energy - f_force     1")
