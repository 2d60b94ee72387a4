"""Write the golden fixtures in tests/golden/ from the reference SIS Fortran
(oracle/_ref, built from /root/reference by oracle/Makefile).

Each fixture holds the inputs (system, MT word tapes of the reference's own
mt_stream streams) and the reference's outputs: interaction lists, forces,
energies, base-pair state after the limit MC, and the state after a short
job_md run.  Only the tape prefix that the run consumes is stored (counted
by replaying the run in the C restatement).

Run:  python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from sis_amd import configs  # noqa: E402
from sis_amd.abi import MOLT_CIRCRNA  # noqa: E402
from oracle import corc, refdrv  # noqa: E402


def mixed_chains(seed=11):
    """A circular 16-nt chain, then linear chains of 24 and 40 nt packed in
    a 64 A region (unequal lengths exercise the bp_mp(6) index rule of
    src/neighbor_list.F90:212), non-periodic.  The first seed from `seed`
    whose base-pair candidates include that rule is used."""
    for sd in range(seed, seed + 200):
        rng = np.random.default_rng(sd)
        seqs = [configs.random_seq(rng, 16, 0.6), configs.random_seq(rng, 24, 0.6),
                configs.random_seq(rng, 40, 0.6)]
        lin = configs._place_chains(rng, seqs[1:], 64.0, clash=9.0)
        n = len(seqs[0])
        R = 5.9 / (2 * np.sin(np.pi / n))
        th = 2 * np.pi * np.arange(n) / n
        ring = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(n)], axis=1)
        ring = ring + np.array([-40.0, 32.0, 32.0])
        p = configs.default_params(tempK=300.0, nl_margin=6.0)
        sysd = configs._mk("mixed", seqs, np.vstack([ring, lin]), p)
        sysd["moltype"][0] = MOLT_CIRCRNA
        o = corc.Oracle(sysd)
        o.neighbor_list(1)
        bp = o.lists(1)["bp"]
        if np.any(bp[:, 5] != bp[:, 1] + 1):
            return sysd
    raise RuntimeError("no seed exercises the bp_mp(6) rule")


def remd_small(nrep=4, seed=5):
    sysd = configs.remd_c3(nrep=nrep, n=24, seed=seed)
    sysd["name"] = "remd4"
    return sysd


def remd_mc3(nrep=4, seed=9):
    """Limit MC every 3 steps with a 0.04 A skin: rebuilds land on non-MC
    steps, exercising the replica loop's shared MC flag
    (src/job_md.F90:364-401)."""
    sysd = configs.remd_c3(nrep=nrep, n=40, seed=seed, margin=0.04)
    sysd["params"] = dict(sysd["params"], nstep_bp_mc=3)
    sysd["name"] = "remd4mc3"
    return sysd


def remd_bd(nrep=4, seed=4):
    """Ermak-McCammon Brownian dynamics (src/job_md.F90:446-456) on a
    temperature ladder, dt 0.005 ps, a 1.5 A skin: rebuilds and exchanges
    inside the run."""
    sysd = configs.remd_c3(nrep=nrep, n=30, seed=seed, margin=1.5)
    sysd["params"] = dict(sysd["params"], integrator=2, dt=0.005)
    sysd["name"] = "remd4bd"
    return sysd


def remd_external(nrep=2, seed=6):
    """Every external term of src/force.F90:62-75 on a two-replica ladder
    lifted above the stage wall: stage, two constant-force tweezer pairs, a
    force ramp, the Rg bias (harmonic, per-replica target) and two sigmoid
    restraints (one two-sided, one to a fixed bead), near their bounds so
    that their forces are not small."""
    sysd = configs.remd_c3(nrep=nrep, n=40, seed=seed)
    x = np.asarray(sysd["xyz"], float).copy()
    x[..., 2] += 20.0 - x[..., 2].min()
    sysd["xyz"] = x
    n = 40
    x0 = x.reshape(nrep, n, 3)[0]
    rng = np.random.default_rng(seed)
    d = lambda i, j: float(np.linalg.norm(x0[i - 1] - x0[j - 1]))  # noqa: E731
    sysd["external"] = dict(
        stage=1, stage_sigma=10.0, stage_eps=1.2,
        dcf_pairs=np.array([[1, n], [3, n - 4]]), dcf_forces=rng.normal(0, 2.0, size=(nrep, 2, 3)),
        fr_pairs=np.array([[2, n - 1]]), fr_k=np.array([[0.5, 0.7]]),
        fr_init=(x0[[1, n - 2]] + rng.normal(0, 3.0, size=(2, 3)))[None],
        fr_velo=rng.normal(0, 0.01, size=(1, 2, 3)),
        rg=1, rg_k=0.8, rg_0=np.linspace(5.0, 12.0, nrep),
        rest_id=np.array([[5, n - 10, 1], [7, n - 15, 2]]),
        rest_rcut=np.array([d(5, n - 10) + 5.0, d(7, n - 15) + 5.0]),
        rest_para=np.array([[2.0, d(5, n - 10) - 1.0, 3.0], [1.5, d(7, n - 15) + 0.5, 2.0]]))
    sysd["name"] = "remd2ext"
    return sysd


def ss_pairs_for(sysd, seed=3, keep=0.4):
    """A base-pair list for a secondary-structure restriction: a random
    share of the pairs init_bp would allow, plus pairs read_ss drops (chain
    ends, too short a loop, non-canonical, repeated, reversed order)."""
    import ctypes as C
    o = corc.Oracle(sysd)
    n = o.nmp
    ok = []
    for i in range(n):
        for j in range(i + 1, n):
            bt, b3 = C.c_int32(), C.c_int32()
            corc.lib().orc_bp_pair(o.h, i, j, C.byref(bt), C.byref(b3))
            if b3.value:
                ok.append((i + 1, j + 1))
    rng = np.random.default_rng(seed)
    pick = [p for p in ok if rng.random() < keep]
    junk = [(1, n // 2), (n // 2, n), (10, 11), (5, 5 + 3)] + [(b, a) for a, b in pick[:5]] + pick[:3]
    seq = np.asarray(sysd["seq"])
    for a in range(2, n - 1):
        for b in range(a + 5, n):
            if {int(seq[a - 1]), int(seq[b - 1])} in ({0, 0}, {2, 0}, {3, 1}):   # AA, GA, CU
                junk.append((a, b))
                break
        if len(junk) > 16:
            break
    return np.array(pick + junk, np.int32)


def hairpin_ss():
    """hairpin64 restricted to a base-pair list (src/read_ss.F90 with a BPL
    file, src/init_bp.F90:177-209)."""
    sysd = configs.hairpin64(1)
    sysd["ss"] = ss_pairs_for(sysd)
    sysd["name"] = "hairpin_ss"
    return sysd


def hairpin_anneal():
    """hairpin64 annealed 300 -> 290 -> 280 K (steps 0, 100, 200) with the
    input temperature at 900 K: the force's limit MC keeps the input kT
    (src/force_bp_limit_triplet.F90:61-67, kT set once by read_input), the
    electrostatics the first entry's (src/main.F90:90-97); over 300 steps
    the trajectory depends on both."""
    sysd = configs.hairpin64(1)
    sysd["params"] = dict(sysd["params"], tempK=900.0)
    sysd["temps"] = np.array([900.0])
    sysd["anneal"] = (np.array([0, 100, 200], np.int64), np.array([300.0, 290.0, 280.0]))
    sysd["name"] = "hairpin_anneal"
    return sysd


def pbc_resize():
    """small_pbc with [PBC_box] resizing every 10 steps: x and y shrink,
    z grows, each until its target (x reaches it at step 40, y after the
    run, z at step 40) (src/pbc.F90:79-138, src/job_md.F90:576-596)."""
    sysd = configs.small_pbc(4, 30, 100.0, seed=5)
    sysd["resize"] = (10, np.array([-0.5, -0.4, 0.3]), np.array([98.0, 97.0, 101.0]))
    sysd["name"] = "pbc_resize"
    return sysd


CAG47_DIR = HERE / "cag47"   # data of the reference's check/test_DCD_CAG47 (OpenMM-written DCD)


def cag47_system(frame=50, model=1):
    """CAG x 47 (141 nt) from frame `frame` of CAG47.dcd with the NHT22
    force field (para/rna_NHT22.ff): base-pair model 1 without limit MC."""
    from sis_amd.engine import read_dcd_file
    from sis_amd.forcefield import nht22_params
    seq = "".join(l.strip() for l in open(CAG47_DIR / "CAG47.fasta") if not l.startswith(">"))
    frames, _ = read_dcd_file(CAG47_DIR / "CAG47.dcd", len(seq))
    sysd = configs._mk("CAG47", [seq], frames[frame], nht22_params(bp_model=model))
    sysd["name"] = "cag47_m1"
    return sysd


def remd_rg_schedule():
    """remd2ext with a timed Rg bias (read_timed_bias_rg; k and target
    switch at steps 10 and 20, src/job_md.F90:614-627)."""
    sysd = remd_external()
    sysd["rg_schedule"] = (np.array([0, 10, 20], np.int64), np.array([0.8, 1.5, 0.3]),
                           np.array([12.0, 9.0, 14.0]))
    # the static part of the fixture sees the terms before job_md applies
    # the schedule's first entry: start from that entry
    sysd["external"] = dict(sysd["external"], rg_k=0.8, rg_0=np.array([12.0, 12.0]))
    sysd["name"] = "remd2rgs"
    return sysd


def remd_ion():
    """Replica exchange over the ionic strength (REPT%ION: 0.05-0.4 M at one
    temperature; src/energy_replica.F90:90-113, src/replica_exchange.F90)."""
    sysd = configs.remd_c3(nrep=4, n=40, seed=9)
    sysd["temps"] = np.full(4, float(sysd["params"]["tempK"]))
    sysd["remd_ion"] = np.array([0.05, 0.1, 0.2, 0.4])
    sysd["name"] = "remd4ion"
    return sysd


def remd_rg():
    """Replica exchange over the Rg-bias target (REPT%RG: 6-12 A at one
    temperature; src/energy_replica.F90:107-109,143-145,
    src/job_md.F90:564-566) with the other external terms of remd2ext."""
    sysd = remd_external(nrep=4)
    sysd["temps"] = np.full(4, float(sysd["params"]["tempK"]))
    sysd["remd_rg"] = np.array([6.0, 8.0, 10.0, 12.0])
    sysd["external"] = dict(sysd["external"], rg_0=sysd["remd_rg"].copy())
    sysd["name"] = "remd4rg"
    return sysd


def remd_force():
    """Replica exchange over the tweezers' constant force (f-REMD,
    REPT%TWZDCF: one pulled pair, 2-8 kcal/mol/A along one direction, at one
    temperature; src/init_tweezers.F90:24-40, src/energy_replica.F90:103-105)."""
    sysd = remd_external(nrep=4)
    sysd["temps"] = np.full(4, float(sysd["params"]["tempK"]))
    ext = dict(sysd["external"])
    ext["dcf_pairs"] = ext["dcf_pairs"][:1]
    sysd["remd_dcf"] = np.array([2.0, 4.0, 6.0, 8.0])
    sysd["dcf_dir"] = np.array([0.3, -0.5, 0.8])
    u = sysd["dcf_dir"] / np.linalg.norm(sysd["dcf_dir"])
    ext["dcf_forces"] = sysd["remd_dcf"][:, None, None] * u[None, None, :]
    sysd["external"] = ext
    sysd["name"] = "remd4dcf"
    return sysd


def remd_grid():
    """Two-dimensional replica exchange, temperature x ionic strength
    (src/init_replica.F90:66-93 label layout, the pair table of
    src/var_replica.F90:162-259: sets T-odd, ION-odd, T-even, ION-even in
    turn): 2 temperatures x 3 ionic strengths, exchange every 10 steps."""
    temps_t = np.array([300.0, 304.0])
    ion = np.array([0.05, 0.15, 0.4])
    sysd = configs.remd_c3(nrep=6, n=30, seed=12)
    sysd["temps"] = np.repeat(temps_t, len(ion))
    sysd["remd_ion"] = np.tile(ion, len(temps_t))
    sysd["remd_grid"] = dict(temps=temps_t, ion=ion)
    sysd["remd_dims"] = np.array([len(temps_t), len(ion)], np.int32)
    sysd["name"] = "remd2x3"
    return sysd


REF_DIR = Path("/root/reference")   # the reference's own example inputs (read here only, when writing fixtures)


def _t2s2hp(coords="folded"):
    """The reference's T2S2HP hairpin (27 nt, check/T2S2HP.fasta) at the
    coordinates of check/T2S2HP_folded.pdb or check/T2S2HP_random.xyz."""
    seq = "".join(l.strip() for l in open(REF_DIR / "check/T2S2HP.fasta") if not l.startswith(">"))
    if coords == "folded":
        x = configs.read_pdb_coords(REF_DIR / "check/T2S2HP_folded.pdb")
    else:
        x = configs.read_xyz_coords(REF_DIR / "check/T2S2HP_random.xyz")
    assert len(x) == len(seq)
    return seq, x


def t2s2hp_sisff():
    """The reference's template force field sis.ff (harmonic angle, cosine
    dihedral, sis.ff:8-18) on T2S2HP from its folded structure, with the
    electrostatics of check/input_T2S2HP.toml:33-38 (cutoff type 2, 10 Debye
    lengths, 0.15 M, 300 K, 10 A margin, viscosity 1e-5 Pa s).  sis.ff's
    negative U0 make it a model-1/3 file (E = U0 exp(-u)); base-pair model 3
    without the limit MC (src/force_bp.F90)."""
    from sis_amd.forcefield import read_ff
    seq, x = _t2s2hp("folded")
    p = read_ff(REF_DIR / "sis.ff", tempK=300.0, ionic_strength=0.15, ele_cutoff_type=2, ele_cutoff=10.0,
                nl_margin=10.0, viscosity_pas=1e-5, dt=0.05, bp_model=3, bp_min_loop=3, nstep_bp_mc=0)
    sysd = configs._mk("t2s2hp_sisff", [seq], x, p)
    return sysd


def t2s2hp_ion8(cutoff=10.0, name="t2s2hp_ion8"):
    """Replica exchange over the ionic strength as in
    check/T2S2HP_REMD_C/input_REMD.toml: 8 replicas at 0.02-0.80 M, 300 K,
    DH cutoff type 2 at 10 Debye lengths, exchange every 10 steps, base-pair
    model 5 without the limit MC (nstep_MC = 0), from T2S2HP_folded.pdb.
    Force field check/para.ff (NNend = true, para.ff:24; that input's
    htv23-19.ff is not in the reference).  The margin is 1 A instead of the
    file's 10 so that the 100 steps hold rebuilds between exchanges: each
    exchange changes lambdaD at once and the cutoff only at the replica's
    next rebuild (src/job_md.F90:556, src/neighbor_list.F90:80-82)."""
    from sis_amd.forcefield import read_ff
    seq, x = _t2s2hp("folded")
    p = read_ff(REF_DIR / "check/para.ff", tempK=300.0, ionic_strength=0.15, ele_cutoff_type=2,
                ele_cutoff=cutoff, nl_margin=1.0, viscosity_pas=1e-5, dt=0.05, bp_model=5, bp_min_loop=3,
                max_bp_per_nt=1, nstep_bp_mc=0)
    sysd = configs._mk(name, [seq], x, p, nrep=8, temps=np.full(8, 300.0))
    sysd["remd_ion"] = np.array([0.02, 0.05, 0.10, 0.15, 0.20, 0.30, 0.50, 0.80])
    return sysd


def t2s2hp_temp8(cutoff=10.0, name="t2s2hp_temp8"):
    """Replica exchange over the temperature as in
    check/T2S2HP_REMD_T/input_REMD.toml: 8 replicas at 293.15-363.15 K,
    0.15 M, DH cutoff type 2 at 10 Debye lengths, exchange every 10 steps,
    from check/T2S2HP_random.xyz with check/sis_dih_ReB_NN.ff, the force
    field of check/input_T2S2HP.toml.  The limit MC runs every step
    (input_T2S2HP.toml's nstep_MC = 1) and the margin is 1 A, so that
    exchanges fall between rebuilds."""
    from sis_amd.forcefield import read_ff
    seq, x = _t2s2hp("random")
    p = read_ff(REF_DIR / "check/sis_dih_ReB_NN.ff", tempK=300.0, ionic_strength=0.15, ele_cutoff_type=2,
                ele_cutoff=cutoff, nl_margin=1.0, viscosity_pas=1e-5, dt=0.05, bp_model=5, bp_min_loop=3,
                max_bp_per_nt=1, nstep_bp_mc=1)
    temps = np.array([293.15, 303.15, 313.15, 323.15, 333.15, 343.15, 353.15, 363.15])
    sysd = configs._mk(name, [seq], x, p, nrep=8, temps=temps)
    return sysd


SYSTEMS = {
    "hairpin64": (lambda: configs.hairpin64(1), 60, 0),
    "small_pbc": (lambda: configs.small_pbc(4, 30, 100.0, seed=5), 60, 0),
    "mixed": (mixed_chains, 40, 0),
    "remd4": (remd_small, 40, 10),
    "remd4mc3": (remd_mc3, 40, 10),
    "remd4bd": (remd_bd, 40, 10),
    "remd2ext": (remd_external, 30, 10),
    "hairpin_ss": (hairpin_ss, 60, 0),
    "hairpin_anneal": (hairpin_anneal, 300, 0),
    "pbc_resize": (pbc_resize, 60, 0),
    "cag47_m1": (cag47_system, 40, 0),
    "remd2rgs": (remd_rg_schedule, 30, 10),
    "remd4ion": (remd_ion, 40, 10),
    "remd4rg": (remd_rg, 40, 10),
    "remd4dcf": (remd_force, 40, 10),
    "t2s2hp_sisff": (t2s2hp_sisff, 200, 0),
    "t2s2hp_ion8": (t2s2hp_ion8, 100, 10),
    "t2s2hp_temp8": (t2s2hp_temp8, 100, 10),
    # the same ladders with the cutoff at 3 and 2 Debye lengths, where the
    # cutoff's timing moves the trajectory by far more than the tolerances
    "t2s2hp_ion8c3": (lambda: t2s2hp_ion8(3.0, "t2s2hp_ion8c3"), 100, 10),
    "t2s2hp_temp8c2": (lambda: t2s2hp_temp8(2.0, "t2s2hp_temp8c2"), 100, 10),
    "remd2x3": (remd_grid, 80, 10),
}
LIST_PARAMS = ("nn_dH", "nn_dS", "nnend_dH", "nnend_dS")


def force_mc(sysd) -> int:
    """Whether the fixture's single force evaluation runs the limit MC: as
    job_md would (flg_bp_MC only with nstep_MC > 0, src/job_md.F90:270-273,
    364-367); models 1/3 run without it."""
    p = sysd["params"]
    return 1 if p["bp_model"] == 5 and p["nstep_bp_mc"] > 0 else 0


def make(name, builder, nstep, nex, seed=4321):
    s = builder()
    nrep = int(s["nrep"])
    big = 3_000_000
    ops = [("tape", r, big, f"tape{r}") for r in range(1, nrep + 1)]
    if nrep > 1:
        ops.append(("tape", -1, 4096, "tape_rep"))
    ext = s.get("external")
    pre = [("external", ext, 0)] if ext else []
    if s.get("rg_schedule") is not None:
        pre.append(("rg_schedule",) + tuple(s["rg_schedule"]))
    if s.get("remd_dcf") is not None:
        pre.append(("dcf_remd", s["dcf_dir"]))
    if s.get("ss") is not None:
        pre.append(("ss", s["ss"]))
    if s.get("anneal") is not None:
        pre.append(("anneal", s["anneal"][0], s["anneal"][1]))
    if s.get("resize") is not None:
        pre.append(("resize",) + tuple(s["resize"]))
    ops = pre + ops + [("nl", r) for r in range(1, nrep + 1)]
    fmc = force_mc(s)
    ops += [("lists", 1, "L"), ("force", 1, fmc, "F"), ("bp_state", 1, "B"),
            ("energy", 1, float(s["temps"][0]), 0, 1, "E"), ("ele", 1, "ele"), ("bp_tables", "T")]
    ref0 = refdrv.run(s, ops, seed=seed, nstep_exchange=nex)
    twz = None
    if ext and ext.get("fr_pairs") is not None:   # the reference's .twz lines of the run, every 10 steps
        import tempfile
        twz = Path(tempfile.mkdtemp()) / "run.twz"
        pre1 = pre + [("twz", twz, 10)]
    else:
        pre1 = pre
    ref1 = refdrv.run(s, pre1 + [("job_md", nstep, "R")], seed=seed, nstep_exchange=nex)
    # count the words the static part + the run consume, using the oracle
    o = corc.Oracle(s)
    if s.get("remd_dims") is not None:
        o.set_rep_dims(s["remd_dims"])
    if s.get("remd_ion") is not None:
        o.set_ionic(s["remd_ion"])
    if ext:
        o.set_external(ext)
    if s.get("rg_schedule") is not None:
        o.set_rg_schedule(*s["rg_schedule"])
    if s.get("remd_rg") is not None:
        o.set_rep_rg(s["remd_rg"])
    if s.get("remd_dcf") is not None:
        o.set_rep_dcf(s["remd_dcf"], s["dcf_dir"])
    if s.get("ss") is not None:
        o.set_ss(s["ss"])
    if s.get("anneal") is not None:
        o.set_anneal(*s["anneal"])
    if s.get("resize") is not None:
        o.set_box_resize(*s["resize"])
    for r in range(1, nrep + 1):
        o.set_tape(r, ref0[f"tape{r}"])
    if nrep > 1:
        o.set_tape(-1, ref0["tape_rep"])
    for r in range(1, nrep + 1):
        o.neighbor_list(r)
    o.force(1, fmc)
    used_static = [o.tape_pos(r) for r in range(1, nrep + 1)]
    o2 = corc.Oracle(s)
    if s.get("remd_dims") is not None:
        o2.set_rep_dims(s["remd_dims"])
    if s.get("remd_ion") is not None:
        o2.set_ionic(s["remd_ion"])
    if ext:
        o2.set_external(ext)
    if s.get("rg_schedule") is not None:
        o2.set_rg_schedule(*s["rg_schedule"])
    if s.get("remd_rg") is not None:
        o2.set_rep_rg(s["remd_rg"])
    if s.get("remd_dcf") is not None:
        o2.set_rep_dcf(s["remd_dcf"], s["dcf_dir"])
    if s.get("ss") is not None:
        o2.set_ss(s["ss"])
    if s.get("anneal") is not None:
        o2.set_anneal(*s["anneal"])
    if s.get("resize") is not None:
        o2.set_box_resize(*s["resize"])
    for r in range(1, nrep + 1):
        o2.set_tape(r, ref0[f"tape{r}"])
    if nrep > 1:
        o2.set_tape(-1, ref0["tape_rep"])
    o2.init_velocities()
    o2.md(nstep, nex)
    used_run = [o2.tape_pos(r) for r in range(1, nrep + 1)]
    keep = max(max(used_static), max(used_run)) + 64
    out = dict(
        nchains=s["nchains"], nmp_chain=s["nmp_chain"], moltype=s["moltype"], seq=s["seq"],
        xyz=s["xyz"], temps=s["temps"], nrep=nrep, nstep=nstep, nstep_exchange=nex,
        tapes=np.stack([ref0[f"tape{r}"][:keep] for r in range(1, nrep + 1)]),
        tape_rep=ref0.get("tape_rep", np.zeros(1, np.uint32))[:256],
        wca=ref0["L"]["wca"], ele=ref0["L"]["ele"], bp=ref0["L"]["bp"], bpcoef=ref0["L"]["bpcoef"],
        force=ref0["F"], bp_status=ref0["B"][0], ene_bp=ref0["B"][1], energies=ref0["E"],
        ele_params=ref0["ele"], bp3_dH=ref0["T"][0], bp3_dS=ref0["T"][1], bp_cut=ref0["T"][2],
        md_xyz=ref1["R"]["xyz"], md_velos=ref1["R"]["velos"], md_accels=ref1["R"]["accels"],
        md_energies=ref1["R"]["energies"], md_rep2lab=ref1["R"]["rep2lab"],
    )
    # parameters as a flat float/int record
    p = s["params"]
    scal = {k: v for k, v in p.items() if isinstance(v, (int, float))}
    out["param_names"] = np.array(sorted(scal))
    out["param_values"] = np.array([float(scal[k]) for k in sorted(scal)])
    out["box"] = np.array(p["box"], float)
    for k in LIST_PARAMS:
        out["param_" + k] = np.asarray(p[k], np.float64)
    # per-type base-pair parameters (GC, AU, GU) in BP_FIELDS order
    from sis_amd.abi import BP_FIELDS
    out["bp_params"] = np.array([[float(p["bp"][t].get(f, 0.0)) for f in BP_FIELDS] for t in ("GC", "AU", "GU")])
    for k, v in (ext or {}).items():
        out["ext_" + k] = np.asarray(v)
    if s.get("ss") is not None:
        out["ss_pairs"] = np.asarray(s["ss"], np.int32)
    if twz is not None:
        out["twz_text"] = np.array(twz.read_text())
    if s.get("remd_ion") is not None:
        out["remd_ion"] = np.asarray(s["remd_ion"], np.float64)
    if s.get("remd_dims") is not None:
        out["remd_dims"] = np.asarray(s["remd_dims"], np.int32)
        out["grid_temps"] = np.asarray(s["remd_grid"]["temps"], np.float64)
        out["grid_ion"] = np.asarray(s["remd_grid"]["ion"], np.float64)
    if s.get("remd_rg") is not None:
        out["remd_rg"] = np.asarray(s["remd_rg"], np.float64)
    if s.get("remd_dcf") is not None:
        out["remd_dcf"] = np.asarray(s["remd_dcf"], np.float64)
        out["dcf_dir"] = np.asarray(s["dcf_dir"], np.float64)
    if s.get("rg_schedule") is not None:
        out["rgs_steps"], out["rgs_k"], out["rgs_0"] = [np.asarray(a) for a in s["rg_schedule"]]
    if s.get("resize") is not None:
        out["resize_step"] = np.int64(s["resize"][0])
        out["resize_change"] = np.asarray(s["resize"][1], np.float64)
        out["resize_target"] = np.asarray(s["resize"][2], np.float64)
    if s.get("anneal") is not None:
        out["anneal_steps"] = np.asarray(s["anneal"][0], np.int64)
        out["anneal_temps"] = np.asarray(s["anneal"][1], np.float64)
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, "tape words kept", keep, "lists", {k: len(ref0["L"][k]) for k in ("wca", "ele", "bp")})


def load(name):
    """Rebuild the system dict of a fixture (params from the defaults plus
    the stored scalar overrides)."""
    d = dict(np.load(HERE / f"{name}.npz", allow_pickle=False))
    pn = [str(x) for x in d["param_names"]]
    over = {k: float(v) for k, v in zip(pn, d["param_values"])}
    p = configs.default_params()
    for k, v in over.items():
        p[k] = int(v) if isinstance(p.get(k), int) else v
    seqs = None
    sysd = dict(name=name, nchains=int(d["nchains"]), nmp_chain=d["nmp_chain"], moltype=d["moltype"],
                seq=d["seq"], xyz=d["xyz"], params=p, nrep=int(d["nrep"]), temps=d["temps"],
                nstep=int(d["nstep"]), seqs=seqs)
    p["box"] = [float(b) for b in d["box"]]
    for k in LIST_PARAMS:
        if "param_" + k in d:
            p[k] = [float(v) for v in d["param_" + k]]
    if "bp_params" in d:
        from sis_amd.abi import BP_FIELDS
        p["bp"] = {t: {f: float(v) for f, v in zip(BP_FIELDS, row)} for t, row in zip(("GC", "AU", "GU"), d["bp_params"])}
    ext = {k[4:]: d[k] for k in d if k.startswith("ext_")}
    if ext:
        for k in ("stage", "rg"):
            ext[k] = int(ext[k])
        for k in ("stage_sigma", "stage_eps", "rg_k"):
            ext[k] = float(ext[k])
        sysd["external"] = ext
    if "ss_pairs" in d:
        sysd["ss"] = d["ss_pairs"]
    if "remd_ion" in d:
        sysd["remd_ion"] = d["remd_ion"]
    if "remd_dims" in d:
        sysd["remd_dims"] = d["remd_dims"]
        sysd["remd_grid"] = dict(temps=d["grid_temps"], ion=d["grid_ion"])
    if "remd_rg" in d:
        sysd["remd_rg"] = d["remd_rg"]
    if "remd_dcf" in d:
        sysd["remd_dcf"], sysd["dcf_dir"] = d["remd_dcf"], d["dcf_dir"]
    if "rgs_steps" in d:
        sysd["rg_schedule"] = (d["rgs_steps"], d["rgs_k"], d["rgs_0"])
    if "resize_step" in d:
        sysd["resize"] = (int(d["resize_step"]), d["resize_change"], d["resize_target"])
    if "anneal_steps" in d:
        sysd["anneal"] = (d["anneal_steps"], d["anneal_temps"])
    return sysd, d


def dcd_split(noise=0.5, nframes=6, seed=17):
    """A two-chain duplex with a circular chain beside it, for the DCD job's
    lists: T2S2HP_folded.pdb cut in its loop into chains of 13 and 14 nt
    (the stem pairs become pairs between the chains), and a flat circular
    chain of 16 nt 30 A away.  Force field check/sis_dih_ReB_NN.ff with the
    electrostatics and limit MC of check/input_T2S2HP.toml (cutoff type 2,
    10 Debye lengths, 0.15 M, nstep_MC = 1).  Frames: the folded
    coordinates, then the same with N(0, noise^2) A displacements, stored
    as float32 as a DCD file holds them."""
    from sis_amd.forcefield import read_ff
    seq, x = _t2s2hp("folded")
    rng = np.random.default_rng(seed)
    n = 16
    ring_seq = configs.random_seq(rng, n, 0.6)
    R = 5.9 / (2 * np.sin(np.pi / n))
    th = 2 * np.pi * np.arange(n) / n
    ring = np.stack([R * np.cos(th), R * np.sin(th), np.zeros(n)], axis=1) + x.mean(0) + np.array([30.0, 0.0, 0.0])
    p = read_ff(REF_DIR / "check/sis_dih_ReB_NN.ff", tempK=300.0, ionic_strength=0.15, ele_cutoff_type=2,
                ele_cutoff=10.0, nl_margin=10.0, viscosity_pas=1e-5, dt=0.05, bp_model=5, bp_min_loop=3,
                max_bp_per_nt=1, nstep_bp_mc=1)
    xyz = np.vstack([x, ring])
    sysd = configs._mk("dcd_split", [seq[:13], seq[13:], ring_seq], xyz, p)
    sysd["moltype"][2] = MOLT_CIRCRNA
    frames = [xyz] + [xyz + rng.normal(0.0, noise, xyz.shape) for _ in range(nframes - 1)]
    return sysd, np.array(frames).astype(np.float32)


def make_dcd(name="dcd_split", seed=4321):
    """job_dcd on the reference build: the lists of list_bp / list_exv /
    list_ele (src/main.F90:144-154) and, per frame, energy_sumup with the
    limit MC on replica 1's stream (src/job_dcd.F90:59-80) and the base-pair
    state it leaves."""
    s, frames = dcd_split()
    T = float(s["temps"][0])
    ops = [("tape", 1, 400_000, "t"), ("analysis_lists",), ("lists", 1, "L")]
    for k, f in enumerate(frames):
        ops += [("set_xyz", f.astype(np.float64)[None]), ("energy", 1, T, 1, 0, f"E{k}"), ("bp_state", 1, f"B{k}")]
    ref = refdrv.run(s, ops, seed=seed)
    o = corc.Oracle(s)
    o.set_tape(1, ref["t"])
    o.analysis_lists(1)
    for f in frames:
        o.set_xyz(f.astype(np.float64)[None])
        o.energy(1, T, 1, 0)
    keep = o.tape_pos(1) + 64
    st = [ref[f"B{k}"][0].astype(np.uint8) for k in range(len(frames))]
    out = dict(
        nchains=s["nchains"], nmp_chain=s["nmp_chain"], moltype=s["moltype"], seq=s["seq"],
        xyz=s["xyz"], temps=s["temps"], nrep=1, nstep=0, nstep_exchange=0,
        tapes=ref["t"][:keep][None], tape_rep=np.zeros(1, np.uint32), frames=frames,
        wca=ref["L"]["wca"], ele=ref["L"]["ele"], bp=ref["L"]["bp"], bpcoef=ref["L"]["bpcoef"],
        energies=np.array([ref[f"E{k}"] for k in range(len(frames))]),
        bp_status=np.array(st), ene_bp=np.array([ref[f"B{k}"][1] for k in range(len(frames))]),
    )
    p = s["params"]
    scal = {k: v for k, v in p.items() if isinstance(v, (int, float))}
    out["param_names"] = np.array(sorted(scal))
    out["param_values"] = np.array([float(scal[k]) for k in sorted(scal)])
    out["box"] = np.array(p["box"], float)
    for k in LIST_PARAMS:
        out["param_" + k] = np.asarray(p[k], np.float64)
    from sis_amd.abi import BP_FIELDS
    out["bp_params"] = np.array([[float(p["bp"][t].get(f, 0.0)) for f in BP_FIELDS] for t in ("GC", "AU", "GU")])
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, "tape words kept", keep, "lists", {k: len(ref["L"][k]) for k in ("wca", "ele", "bp")},
          "formed", [int(x.sum()) for x in st])


def make_cag47_dcd(name="cag47_dcd"):
    """The reference build's job_dcd energies over the 100 frames of
    check/test_DCD_CAG47/CAG47.dcd (NHT22, base-pair model 1, no limit MC):
    list_bp / list_exv / list_ele, then energy_sumup per frame."""
    from sis_amd.engine import read_dcd_file
    s = cag47_system()
    frames, _ = read_dcd_file(CAG47_DIR / "CAG47.dcd", 141)
    ops = [("analysis_lists",)]
    for k, f in enumerate(frames):
        ops += [("set_xyz", f[None]), ("energy", 1, 295.15, 0, 0, f"E{k}")]
    ref = refdrv.run(s, ops, seed=123)
    E = np.array([ref[f"E{k}"] for k in range(len(frames))])
    np.savez_compressed(HERE / f"{name}.npz", energies=E)
    print(name, "E_bp range", E[:, 4].min(), E[:, 4].max())


if __name__ == "__main__":
    names = sys.argv[1:] or list(SYSTEMS) + ["dcd_split", "cag47_dcd"]
    for nm in names:
        if nm == "dcd_split":
            make_dcd()
            continue
        if nm == "cag47_dcd":
            make_cag47_dcd()
            continue
        b, n, nex = SYSTEMS[nm]
        make(nm, b, n, nex)
