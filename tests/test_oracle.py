"""The C restatement (oracle/liboracle.so) pinned against the reference:
golden fixtures written by the reference Fortran (tests/golden/), the live
reference build (oracle/_ref) where present, and closed forms."""
import math
from pathlib import Path

import numpy as np
import pytest

from sis_amd import configs
from oracle import corc, refdrv
from tests.golden.make_golden import force_mc, load

FIXTURES = ["hairpin64", "small_pbc", "mixed", "remd4", "remd4mc3", "remd4bd", "remd2ext", "hairpin_ss", "hairpin_anneal", "pbc_resize", "cag47_m1", "remd2rgs", "remd4ion", "remd4rg", "remd4dcf",
            "t2s2hp_sisff", "t2s2hp_ion8", "t2s2hp_temp8", "t2s2hp_ion8c3", "t2s2hp_temp8c2", "remd2x3"]


def _oracle_from_fixture(name):
    sysd, d = load(name)
    o = corc.Oracle(sysd)
    if "remd_dims" in sysd:
        o.set_rep_dims(sysd["remd_dims"])
    if "remd_ion" in sysd:
        o.set_ionic(sysd["remd_ion"])
    if "external" in sysd:
        o.set_external(sysd["external"])
    if "rg_schedule" in sysd:
        o.set_rg_schedule(*sysd["rg_schedule"])
    if "remd_rg" in sysd:
        o.set_rep_rg(sysd["remd_rg"])
    if "remd_dcf" in sysd:
        o.set_rep_dcf(sysd["remd_dcf"], sysd["dcf_dir"])
    if "ss" in sysd:
        o.set_ss(sysd["ss"])
    if "anneal" in sysd:
        o.set_anneal(*sysd["anneal"])
    if "resize" in sysd:
        o.set_box_resize(*sysd["resize"])
    for r in range(1, sysd["nrep"] + 1):
        o.set_tape(r, d["tapes"][r - 1])
    if sysd["nrep"] > 1:
        o.set_tape(-1, d["tape_rep"])
    return sysd, d, o


@pytest.mark.parametrize("name", FIXTURES)
def test_lists_bit_exact(name):
    sysd, d, o = _oracle_from_fixture(name)
    for r in range(1, sysd["nrep"] + 1):
        o.neighbor_list(r)
    L = o.lists(1)
    assert np.array_equal(L["wca"], d["wca"])
    assert np.array_equal(L["ele"], d["ele"])
    assert np.array_equal(L["bp"], d["bp"])          # same candidates, same order
    assert np.array_equal(L["bpcoef"], d["bpcoef"])  # triplet dH/dS bit-exact


@pytest.mark.parametrize("name", FIXTURES)
def test_force_and_bp_state(name):
    sysd, d, o = _oracle_from_fixture(name)
    for r in range(1, sysd["nrep"] + 1):
        o.neighbor_list(r)
    f = o.force(1, force_mc(sysd))
    fr = d["force"]
    rel = np.abs(f - fr).max() / np.abs(fr).max()
    assert rel < 1e-12
    st, en = o.bp_state(1)
    assert np.array_equal(st, d["bp_status"].astype(np.uint8))   # limit MC on the same words
    assert np.allclose(en, d["ene_bp"], rtol=1e-12, atol=1e-300)
    e = o.energy(1, float(sysd["temps"][0]), 0, 1)
    assert np.allclose(e[1:], d["energies"][1:], rtol=1e-10, atol=1e-12)
    assert np.allclose(o.ele_params(1), d["ele_params"], rtol=1e-14)


@pytest.mark.parametrize("name", FIXTURES)
def test_tables(name):
    sysd, d, o = _oracle_from_fixture(name)
    dh, ds, cut = o.tables()
    assert np.array_equal(dh, d["bp3_dH"]) and np.array_equal(ds, d["bp3_dS"])
    assert np.array_equal(cut, d["bp_cut"])
    if sysd["params"]["bp_model"] == 5:
        assert np.count_nonzero(dh) > 400   # 468 triplets (src/init_bp.F90:170-171)


@pytest.mark.parametrize("name", FIXTURES)
def test_md_trajectory(name):
    """job_md for a few dozen steps with the reference's own MT streams:
    noise, limit MC, rebuilds and (remd4) exchanges consumed in order."""
    sysd, d, o = _oracle_from_fixture(name)
    o.init_velocities()
    o.md(int(d["nstep"]), int(d["nstep_exchange"]))
    x, v, a = o.state()
    assert np.abs(x - d["md_xyz"]).max() < 1e-10
    assert np.abs(v - d["md_velos"]).max() < 1e-12
    assert np.abs(a - d["md_accels"]).max() < 1e-14
    if sysd["nrep"] > 1:
        assert np.array_equal(o.labels(), d["md_rep2lab"][: sysd["nrep"]])


@pytest.mark.parametrize("name", ["t2s2hp_ion8c3", "t2s2hp_temp8c2"])
def test_cutoff_follows_rebuilds_not_exchanges(name):
    """DH cutoff type 2 under an ionic-strength or temperature ladder: the
    reference's set_ele after an exchange changes lambdaD and the
    coefficient, while ele_cutoff moves only at the replica's next
    neighbor_list (src/job_md.F90:556, src/neighbor_list.F90:80-82).  The
    fixture's run has exchanges between rebuilds, so an oracle with the
    cutoff moved at set_ele (planted) leaves the reference trajectory."""
    sysd, d, o = _oracle_from_fixture(name)
    o.init_velocities()
    nreb, _ = o.md(int(d["nstep"]), int(d["nstep_exchange"]))
    assert nreb >= 1
    assert np.abs(o.state()[0] - d["md_xyz"]).max() < 1e-10
    sysd, d, o = _oracle_from_fixture(name)
    o.plant(1)
    o.init_velocities()
    o.md(int(d["nstep"]), int(d["nstep_exchange"]))
    assert np.abs(o.state()[0] - d["md_xyz"]).max() > 1e-5


def test_fixtures_use_reference_default_physics():
    """The T2S2HP fixtures carry the reference's own example physics: the
    cosine dihedral and harmonic angle of sis.ff, DH cutoff type 2 at 10
    Debye lengths, NNend corrections (check/para.ff), base-pair model 3, and
    both REMD ladders of check/T2S2HP_REMD_C/T."""
    p = {n: load(n)[0]["params"] for n in ("t2s2hp_sisff", "t2s2hp_ion8", "t2s2hp_temp8")}
    assert p["t2s2hp_sisff"]["dih_type"] == 1 and p["t2s2hp_sisff"]["angl_reb"] == 0
    assert p["t2s2hp_sisff"]["bp_model"] == 3
    assert all(q["ele_cutoff_type"] == 2 and q["ele_cutoff"] == 10.0 for q in p.values())
    assert p["t2s2hp_ion8"]["nnend"] == 1 and any(v != 0.0 for v in p["t2s2hp_ion8"]["nnend_dH"])
    d = load("t2s2hp_ion8")[1]
    # the NNend corrections change the triplet table the reference built
    sysd = load("t2s2hp_ion8")[0]
    q = dict(sysd["params"], nnend=0)
    o = corc.Oracle(dict(sysd, params=q))
    dh, _, _ = o.tables()
    assert not np.array_equal(dh, d["bp3_dH"])


@pytest.mark.skipif(not __import__("os").path.exists("/root/reference/check/sis_dih_ReB_NN.ff"),
                    reason="reference force-field files not present")
def test_read_ff_matches_benchmark_parameters():
    """read_ff on check/sis_dih_ReB_NN.ff gives the benchmark force field
    (default_params) except the bond, which the benchmark systems set to
    k 15, r0 5.9; on sis.ff the harmonic angle and cosine dihedral."""
    from sis_amd.forcefield import default_params, read_ff
    d = default_params()
    p = read_ff("/root/reference/check/sis_dih_ReB_NN.ff")
    diff = {k for k in d if d[k] != p[k]}
    assert diff == {"bond_k", "bond_r0"}
    q = read_ff("/root/reference/sis.ff")
    assert (q["angl_reb"], q["angl_k"], q["angl_t0"]) == (0, 10.0, 2.643)
    assert (q["dih_type"], q["dih_k"], q["dih_p0"]) == (1, 5.0, 2.861)
    n = read_ff("/root/reference/check/para.ff")
    assert n["nnend"] == 1 and n["nnend_dH"][0] == 4.36


def test_philox_known_answers():
    """Philox4x32-10 restatement against the known-answer vectors published
    with the Random123 library of Salmon et al. (SC'11): zero counter and
    key, all-ones counter and key, and the pi-digits counter/key."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert [int(v) for v in corc.philox4x32_10(ctr, key)] == want


def test_counter_rng_normals_are_standard():
    """The counter-RNG normals the oracle draws for the engine's performance
    form: mean 0, variance 1 per component over many beads and steps, and
    the initial velocities at the equipartition temperature."""
    s = configs.remd_c3(nrep=2, n=400, seed=3)
    s["params"] = dict(s["params"], nstep_bp_mc=0)
    o = corc.Oracle(s)
    o.set_counter_rng(1234)
    o.init_velocities()
    _, v, _ = o.state()
    m = np.array([{0: 328.212, 1: 305.164, 2: 344.212, 3: 304.182}[int(q)] for q in s["seq"]])
    kT = 0.0019872042586 * np.asarray(s["temps"])
    z = v / np.sqrt(kT[:, None, None] / m[None, :, None])
    assert abs(z.mean()) < 0.05 and abs(z.var() - 1.0) < 0.05


@pytest.mark.parametrize("which", ["c2", "mixed_lengths"])
def test_cell_enumeration_equals_reference_loops(which):
    """The oracle's cell-list enumeration of large periodic systems gives
    the O(N^2) loops' lists exactly: same pairs, same order (chain i, chain
    j, i, j), same base-pair candidates and triplet coefficients, and the
    same forces.  C2 at full size (9,400 beads), and 110 chains of 20-69 nt
    in a 230 A box (4 cells per axis at the 55 A DH list cutoff)."""
    if which == "c2":
        s = configs.condensate_c2()
    else:
        rng = np.random.default_rng(4)
        seqs = [configs.random_seq(rng, int(n), 0.6) for n in rng.integers(20, 70, size=110)]
        box = 230.0
        x = configs._place_chains(rng, seqs, box, clash=4.0)
        s = configs._mk("mixed_pbc", seqs, x, configs.default_params(pbc=1, box=[box] * 3, nl_margin=5.0))
    s["params"] = dict(s["params"], nstep_bp_mc=0)
    assert int(np.sum(s["nmp_chain"])) >= 4096
    out = []
    for brute in (True, False):
        o = corc.Oracle(s)
        o.set_nl_brute(brute)
        o.neighbor_list(1)
        out.append((o.lists(1), o.force(1, 0)))
    (la, fa), (lb, fb) = out
    for k in ("wca", "ele", "bp", "bpcoef"):
        assert np.array_equal(la[k], lb[k]), k
    assert len(la["ele"]) > 1000 and len(la["bp"]) > 10
    assert np.array_equal(fa, fb)


def test_h4_index_rule_exercised():
    """mixed has chains of 24 and 40 nt: some bp_mp(6) entries follow the
    reference's comparison with the length of chain i
    (src/neighbor_list.F90:212) rather than j+1."""
    sysd, d = load("mixed")
    bp = d["bp"]
    assert np.any(bp[:, 5] != bp[:, 1] + 1)


# ---------------- closed forms and invariants ----------------

def test_mt19937_known_answers():
    # first outputs of init_genrand(5489), the MT19937 reference values
    w = corc.mt_words(5489, 5)
    assert list(w) == [3499211612, 581869302, 3890346734, 3586334585, 545404204]
    # 10000th output of the default-seeded generator (C++11 std::mt19937)
    assert corc.mt_words(5489, 10000)[-1] == 4123659995


def test_uniform_conversions_closed_form():
    s = configs.hairpin64(1)
    o = corc.Oracle(s)
    words = np.array([0xFFFFFFFF, 0xFFFFFFFF, 0, 0, 0x80000000, 0x40000000] * 4, np.uint32)
    o.set_tape(1, words)
    d1 = o.genrand(1, 1, 3)
    assert d1[0] == 1.0 and d1[1] == 0.0           # [0,1] closed interval
    ia, ib = 0x80000000 >> 5, 0x40000000 >> 6
    assert d1[2] == (ia * 67108864.0 + ib) / 9007199254740991.0
    o.set_tape(1, words)
    d3 = o.genrand(1, 3, 2)
    assert 0.0 < d3[1] < d3[0] < 1.0               # (0,1) open interval
    assert d3[1] == 0.5 / 4503599627370496.0


def test_set_ele_closed_form():
    """T = 300 K, I = 0.15 M: eps = 77.648, lB = 7.173 A, lambda_D = 7.836 A,
    coefficient 1.5957 kcal A/mol (src/set_ele.F90:45-99)."""
    s = configs.hairpin64(1)
    o = corc.Oracle(s)
    lam, coef, cut, eps = o.ele_params(1)
    Tc = 300.0 - 273.15
    assert eps == pytest.approx(87.740 - 0.40008 * Tc + 9.398e-4 * Tc**2 - 1.410e-6 * Tc**3, rel=1e-14)
    e0, qe, kb, na = 8.8541878128e-12, 1.602176634e-19, 1.380649e-23, 6.02214076e23
    lb = qe**2 / (4 * math.pi * e0 * eps * kb * 300.0) * 1e10
    assert lb == pytest.approx(7.173, abs=1e-3)
    lamD = math.sqrt(e0 * eps * kb * 300.0 / (2 * na * 1e3 * 0.15 * qe**2)) * 1e10
    assert lam == pytest.approx(lamD, rel=1e-12)
    assert coef == pytest.approx((4.38178046 / lb) ** 2 * lb * kb * 300.0 * na / 4184.0, rel=1e-12)
    assert cut == 50.0


def _energy_total(o, x):
    o.set_xyz(x[None])
    return o.energy(1, None, 0, 0)[0]


def test_finite_difference_forces():
    """job_check_force semantics (src/job_check_force.F90:46-82): central
    differences of energy_sumup with h = 1e-3 against force, BP MC off."""
    s = configs.small_pbc(3, 20, 90.0, seed=3)
    s["params"]["nstep_bp_mc"] = 0
    o = corc.Oracle(s)
    o.neighbor_list(1)
    x0 = s["xyz"][0].copy()
    f = o.force(1, 0)
    h = 1e-3
    rng = np.random.default_rng(0)
    worst = 0.0
    for i in rng.choice(x0.shape[0], 12, replace=False):
        for k in range(3):
            xp = x0.copy(); xp[i, k] += h
            xm = x0.copy(); xm[i, k] -= h
            fd = -(_energy_total(o, xp) - _energy_total(o, xm)) / (2 * h)
            worst = max(worst, abs(fd - f[i, k]) / max(1.0, abs(f[i, k])))
    assert worst < 1e-4


def test_bond_force_zero_at_r0_and_wca_zero_beyond_sigma():
    s = configs.hairpin64(1)
    s["params"].update(angl_reb=0, angl_k=0.0, dih_type=0, ele_on=0, bp_model=5)
    n = s["xyz"].shape[1]
    x = np.zeros((n, 3))
    x[:, 0] = 5.9 * np.arange(n)                     # straight chain, bonds at r0
    s["xyz"] = x[None]
    o = corc.Oracle(s)
    o.neighbor_list(1)
    f = o.force(1, 0)
    assert np.abs(f).max() < 1e-12                   # i,i+3 at 17.7 A > sigma


def test_free_particle_velocity_decay():
    """GJF-2GJ without force or noise: v(t+h) = a v(t) (src/job_md.F90:424-428)."""
    s = configs.hairpin64(1)
    o = corc.Oracle(s)
    c = o.md_coef(1)
    a = c[:, 0]
    gamma_over_m = 6 * math.pi * 1.4115e-5 * math.sqrt(1e3 / 4184.0) * 6.02214076e23 * 1e-20 * 10.0
    mass = np.array([{0: 328.212, 1: 305.164, 2: 344.212, 3: 304.182}[q] for q in s["seq"]])
    c1 = 0.5 * 0.05 * gamma_over_m / mass
    assert np.allclose(a, (1 - c1) / (1 + c1), rtol=1e-14)
    assert np.allclose(c[:, 2], math.sqrt(1.0) * 0.05 / np.sqrt(1 + c1), rtol=1e-14)


def test_brownian_coefficients_stokes_einstein():
    """Ermak-McCammon (src/job_md.F90:102-108, 725-731, 755-761): the drift
    per unit force h/gamma and the noise amplitude sqrt(2 D h) obey the
    fluctuation-dissipation relation cr^2 / (2 a) = kT, and the diffusion
    constant is the Stokes-Einstein one, kB T / (6 pi eta R) with R = 10 A,
    computed here in SI units and converted to A^2/ps."""
    s = configs.remd_c3(nrep=2, n=20, seed=1)
    s["params"] = dict(s["params"], integrator=2, dt=0.004)
    o = corc.Oracle(s)
    kb_si, na = 1.380649e-23, 6.02214076e23
    for r in (1, 2):
        tK = float(s["temps"][r - 1])
        c = o.md_coef(r)
        a, cr = c[:, 0], c[:, 3]
        kT = kb_si * tK * na / 4184.0                      # kcal/mol
        assert np.allclose(cr ** 2 / (2 * a), kT, rtol=1e-13)
        d_si = kb_si * tK / (6 * math.pi * 1.4115e-5 * 10e-10)   # m^2/s
        assert np.allclose(cr ** 2 / (2 * 0.004), d_si * 1e8, rtol=1e-12)
        assert np.all(c[:, 1] == 0) and np.all(c[:, 2] == 0)


@pytest.mark.ref
def test_live_reference_brownian_random_system():
    """Brownian dynamics on a fresh 4-replica ladder with a 0.5 A skin
    (a rebuild on nearly every step) and exchanges every 10 steps: oracle
    trajectory vs the reference build."""
    s = configs.remd_c3(nrep=4, n=30, seed=14, margin=0.5)
    s["params"] = dict(s["params"], integrator=2, dt=0.005)
    nstep, nex = 40, 10
    ops = [("tape", r, 400000, f"t{r}") for r in range(1, 5)] + [("tape", -1, 4096, "tr")]
    ref0 = refdrv.run(s, ops, seed=7, nstep_exchange=nex)
    ref1 = refdrv.run(s, [("job_md", nstep, "R")], seed=7, nstep_exchange=nex)["R"]
    o = corc.Oracle(s)
    for r in range(1, 5):
        o.set_tape(r, ref0[f"t{r}"])
    o.set_tape(-1, ref0["tr"])
    o.init_velocities()
    nreb, _ = o.md(nstep, nex)
    assert nreb >= 40
    x, v, a = o.state()
    assert np.abs(x - ref1["xyz"]).max() < 1e-10
    assert np.abs(v - ref1["velos"]).max() == 0.0   # velocities are not integrated
    assert np.array_equal(o.labels(), ref1["rep2lab"][:4])


def test_external_terms_are_energy_gradients():
    """Stage wall, constant-force tweezers, Rg bias and sigmoid restraints
    (src/force_stage.F90 ... src/force_restraint.F90): the force they add
    is minus the central-difference gradient of the energy they add
    (src/energy_stage.F90 ... src/energy_restraint.F90).  The force ramp
    has no energy term in the reference and is left out here."""
    from tests.golden.make_golden import remd_external
    s = remd_external()
    ext = dict(s["external"])
    for k in ("fr_pairs", "fr_k", "fr_init", "fr_velo"):
        ext.pop(k)
    o0, o1 = corc.Oracle(s), corc.Oracle(s)
    o1.set_external(ext)
    x = np.asarray(s["xyz"], float).reshape(2, 40, 3)
    tK = float(s["temps"][0])

    def eext(o, xx):
        o.set_xyz(xx)
        return o.energy(1, tK, 0, 0)[7:].sum()

    for o in (o0, o1):
        o.neighbor_list(1)
    f = o1.force(1, 0) - o0.force(1, 0)
    h = 1e-5
    rng = np.random.default_rng(3)
    for a in list(range(0, 8)) + list(rng.choice(40, 6, replace=False)):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[0, a, c] += h
            xm[0, a, c] -= h
            g = (eext(o1, xp) - eext(o1, xm)) / (2 * h)
            assert abs(f[a, c] + g) <= 1e-5 * max(1.0, abs(g)), (a, c, f[a, c], -g)


# OpenMM energies of check/test_DCD_CAG47 (tests/golden/cag47): bond, angle
# and WCA agree to float32-coordinate rounding; the electrostatic energy
# differs by < 0.5% (the run used the alternative dielectric of
# -D_HTN_CONSISTENT, src/set_ele.F90:29-44, or its own constants); the
# base-pair column comes from a different base-pair model than the stale
# input selects and is not compared.
CAG47_COLS = (("bond", 1, 1, 1e-4, 0.0), ("angl", 2, 2, 1e-4, 0.0), ("exv", 3, 5, 0.0, 1e-3),
              ("ele", 4, 6, 1e-2, 0.0))


def cag47_frames_and_reference():
    from sis_amd.engine import read_dcd_file
    from tests.golden.make_golden import CAG47_DIR, cag47_system
    s = cag47_system()
    frames, _ = read_dcd_file(CAG47_DIR / "CAG47.dcd", 141)
    return s, frames, np.loadtxt(CAG47_DIR / "CAG47.energy.out")


def test_cag47_openmm_energies():
    """job_dcd over the 100 OpenMM frames of CAG x 47 with the NHT22 force
    field (harmonic angle, model-1 base pairs): the oracle's energies
    against the OpenMM column of each term."""
    s, frames, ref = cag47_frames_and_reference()
    o = corc.Oracle(s)
    E = []
    o.analysis_lists(1)
    for x in frames:
        o.set_xyz(x[None])
        E.append(o.energy(1, 295.15, 0, 0))
    E = np.array(E)
    for name, ci, ei, rtol, atol in CAG47_COLS:
        assert np.allclose(E[:, ei], ref[:, ci], rtol=rtol, atol=atol), name
    assert np.all(E[:, 4] < 0)   # base pairs form in every frame
    # every column, the base-pair one included, against the reference
    # build's own job_dcd over the same frames
    want = np.load(Path(__file__).parent / "golden" / "cag47_dcd.npz")["energies"]
    assert np.allclose(E, want, rtol=1e-12, atol=1e-12)


def _grid_pairs(dims):
    """Exchange partners of every label on a replica grid, set by set,
    from the grid coordinates alone: set (parity, d) moves along dimension
    d, pairing values (1,2), (3,4), ... for the odd parity and (2,3),
    (4,5), ... for the even one; labels count with the first dimension
    slowest.  0 = no partner."""
    dims = list(dims)
    R = int(np.prod(dims))
    idx = np.array(np.unravel_index(np.arange(R), dims)).T   # 0-based values
    sets = []
    for par in (0, 1):
        for d, n in enumerate(dims):
            row = np.zeros(R + 2, np.int32)
            for lab in range(R):
                v = idx[lab, d] + 1
                if par == 0:
                    w = v + 1 if v % 2 == 1 else v - 1
                else:
                    w = v + 1 if v % 2 == 0 else v - 1
                if par == 1 and v == 1:
                    w = 0
                if w < 1 or w > n:
                    continue
                j = idx[lab].copy()
                j[d] = w - 1
                row[lab + 1] = np.ravel_multi_index(tuple(j), dims) + 1
            sets.append(row)
    return np.array(sets)


@pytest.mark.parametrize("dims", [[6], [5], [2, 3], [3, 2], [4, 4], [3, 5, 2], [2, 2, 2]])
def test_replica_grid_pair_table(dims):
    """make_replica_exchange_pair_tb (src/var_replica.F90:162-259) of the
    oracle against the partners the grid coordinates give."""
    s = configs.remd_c3(nrep=int(np.prod(dims)), n=12, seed=3)
    o = corc.Oracle(s)
    o.set_rep_dims(dims)
    assert np.array_equal(o.pair_table(), _grid_pairs(dims))


def test_dcd_split_fixture():
    """job_dcd's lists (list_exv, list_ele, list_bp; src/main.F90:144-154)
    and per-frame energies with the limit MC on a duplex of two chains and
    a circular chain: the oracle against the reference build's output for
    the same frames and MT words.  The lists of the MD job give other
    energies here (WCA between a circular chain's ends, list_ele's local
    index exclusions between chains), so the fixture tells the two rules
    apart."""
    sysd, d = load("dcd_split")
    T = float(sysd["temps"][0])
    o = corc.Oracle(sysd)
    o.set_tape(1, d["tapes"][0])
    o.analysis_lists(1)
    L = o.lists(1)
    for k in ("wca", "ele", "bp"):
        assert np.array_equal(L[k], d[k]), k
    assert np.array_equal(L["bpcoef"], d["bpcoef"])
    for k, f in enumerate(d["frames"]):
        o.set_xyz(f.astype(np.float64)[None])
        e = o.energy(1, T, 1, 0)
        assert np.allclose(e, d["energies"][k], rtol=1e-12, atol=1e-12), k
        assert np.array_equal(o.bp_state(1)[0], d["bp_status"][k]), k
    assert d["bp_status"].sum(axis=1).min() > 0
    md = corc.Oracle(sysd)
    md.set_tape(1, d["tapes"][0])
    md.set_xyz(d["frames"][0].astype(np.float64)[None])
    md.neighbor_list(1)
    e_md = md.energy(1, T, 1, 0)
    for col in (5, 6):   # exv, ele
        assert abs(e_md[col] - d["energies"][0][col]) > 1e-3, col


@pytest.mark.ref
def test_dropin_refuses_what_it_does_not_cover():
    """The drop-in layer under the reference's job_md stops with a message,
    before any device call, on a setting it does not cover (here the
    limit MC every 3 steps, whose bp_status_MC copies job_md keeps on the
    host)."""
    if not refdrv.REF_GPU_SO.exists():
        pytest.skip("oracle/_ref/libsisref_gpu.so not built")
    sysd, d = load("remd4mc3")
    with pytest.raises(RuntimeError, match="exited with code 1"):
        refdrv.run(sysd, [("job_md", 2, "R")], seed=4321, nstep_exchange=10, dropin=True, timeout=120)


@pytest.mark.ref
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_live_reference_analysis_lists(seed):
    """list_exv / list_ele / list_bp of the oracle against the reference
    build on systems of a circular chain and two linear chains of unequal
    lengths (the same builder as the mixed fixture), base-pair models 5 and
    1, and energies with the limit MC on the same MT words."""
    from tests.golden.make_golden import mixed_chains
    for model in (5, 1):
        s = mixed_chains(seed)
        s["params"] = dict(s["params"], bp_model=model)
        ops = [("tape", 1, 100000, "t"), ("analysis_lists",), ("lists", 1, "L"),
               ("energy", 1, 300.0, int(model == 5), 0, "E"), ("bp_state", 1, "B")]
        ref = refdrv.run(s, ops, seed=5)
        o = corc.Oracle(s)
        o.set_tape(1, ref["t"])
        o.analysis_lists(1)
        L = o.lists(1)
        for k in ("wca", "ele", "bp"):
            assert np.array_equal(L[k], ref["L"][k]), (model, k)
        assert np.array_equal(L["bpcoef"], ref["L"]["bpcoef"])
        e = o.energy(1, 300.0, int(model == 5), 0)
        assert np.allclose(e, ref["E"], rtol=1e-12, atol=1e-12), model
        assert np.array_equal(o.bp_state(1)[0], ref["B"][0].astype(np.uint8))


@pytest.mark.ref
@pytest.mark.parametrize("model", [1, 3])
def test_live_reference_base_pair_models_1_and_3(model):
    """Base-pair models 1 and 3 without the limit MC (src/force_bp.F90,
    src/energy_bp.F90; model 3 adds the isolated-pair rule of
    src/init_bp.F90): lists, forces, energies and a 40-step trajectory of
    the oracle against the reference build."""
    from tests.golden.make_golden import cag47_system
    s = cag47_system(frame=50, model=model)
    ops = [("tape", 1, 400000, "t"), ("nl", 1), ("lists", 1, "L"), ("force", 1, 0, "F"),
           ("energy", 1, 295.15, 0, 0, "E")]
    ref = refdrv.run(s, ops, seed=7)
    ref1 = refdrv.run(s, [("job_md", 40, "R")], seed=7)["R"]
    o = corc.Oracle(s)
    o.set_tape(1, ref["t"])
    o.neighbor_list(1)
    assert np.array_equal(o.lists(1)["bp"], ref["L"]["bp"])
    f = o.force(1, 0)
    assert np.abs(f - ref["F"]).max() / np.abs(ref["F"]).max() < 1e-12
    e = o.energy(1, 295.15, 0, 0)
    assert np.allclose(e, ref["E"], rtol=1e-12, atol=1e-12) and e[4] < 0
    o2 = corc.Oracle(s)
    o2.set_tape(1, ref["t"])
    o2.init_velocities()
    o2.md(40, 0)
    assert np.abs(o2.state()[0] - ref1["xyz"]).max() < 1e-10


@pytest.mark.ref
def test_live_reference_anneal_keeps_input_kT():
    """Annealing (src/job_md.F90:207-230,600-612): the same schedule from two
    input temperatures.  Only the force's limit MC sees the input
    temperature (its kT is never updated), so the two 400-step runs differ;
    the oracle follows the reference build in both."""
    xs = []
    for tin in (300.0, 900.0):
        s = configs.hairpin64(1)
        s["params"] = dict(s["params"], tempK=tin)
        s["temps"] = np.array([tin])
        steps, temps = [0, 150, 300], [300.0, 290.0, 280.0]
        ref0 = refdrv.run(s, [("tape", 1, 3000000, "t")], seed=99)
        ref1 = refdrv.run(s, [("anneal", steps, temps), ("job_md", 400, "R")], seed=99)["R"]
        o = corc.Oracle(s)
        o.set_tape(1, ref0["t"])
        o.set_anneal(steps, temps)
        o.init_velocities()
        o.md(400, 0)
        x = o.state()[0]
        assert np.abs(x - ref1["xyz"]).max() < 1e-10
        xs.append(x)
    assert np.abs(xs[0] - xs[1]).max() > 0.1


@pytest.mark.ref
def test_live_reference_ss_restriction():
    """A base-pair list on two chains (inter- and intra-chain pairs, plus
    pairs read_ss drops): the oracle's eligibility map equals the
    reference's bp3_map after read_ss + init_bp (src/read_ss.F90,
    src/init_bp.F90:177-313), and the BP candidate lists agree."""
    import ctypes as C
    from tests.golden.make_golden import mixed_chains, ss_pairs_for
    s = mixed_chains()
    pairs = ss_pairs_for(s, seed=5, keep=0.3)
    ref = refdrv.run(s, [("ss", pairs), ("bp_maps", "M"), ("nl", 1), ("lists", 1, "L")])
    o = corc.Oracle(s)
    o.set_ss(pairs)
    n = o.nmp
    elig = np.zeros((n, n), np.int32)
    for i in range(n):
        for j in range(i + 1, n):
            bt, b3 = C.c_int32(), C.c_int32()
            corc.lib().orc_bp_pair(o.h, i, j, C.byref(bt), C.byref(b3))
            elig[i, j] = b3.value
    assert np.array_equal(elig, np.triu(ref["M"][1], 1))
    assert 0 < np.count_nonzero(elig) < len(pairs)
    o.neighbor_list(1)
    assert np.array_equal(o.lists(1)["bp"], ref["L"]["bp"])


@pytest.mark.ref
def test_live_reference_external_terms_periodic():
    """Tweezers (constant force and a ramp at step 40), flat-bottomed Rg bias
    and restraints across the periodic boundary: oracle forces and energies
    vs the reference build."""
    from tests.golden.make_golden import remd_external
    sp = configs.small_pbc(3, 12, 80.0, seed=4)
    n = 36
    ext = dict(remd_external()["external"])
    ext.update(stage=0, rg=2, rg_0=[6.0], dcf_forces=np.asarray(ext["dcf_forces"])[:1],
               dcf_pairs=[[1, n], [3, n - 4]], fr_pairs=[[2, n - 1]],
               rest_id=[[5, n - 10, 1], [7, n - 15, 2]], rest_rcut=[60.0, 60.0],
               rest_para=[[2.0, 20.0, 3.0], [1.5, 25.0, 2.0]])
    ops = [("external", ext, 40), ("nl", 1), ("force", 1, 0, "F"),
           ("energy", 1, float(sp["temps"][0]), 0, 1, "E")]
    ref = refdrv.run(sp, ops)
    o = corc.Oracle(sp)
    o.set_external(ext)
    o.set_step(40)
    o.neighbor_list(1)
    f = o.force(1, 0)
    e = o.energy(1, float(sp["temps"][0]), 0, 1)
    assert np.abs(f - ref["F"]).max() / np.abs(ref["F"]).max() < 1e-12
    assert np.allclose(e, ref["E"], rtol=1e-12, atol=1e-12)
    assert np.all(e[8:] != 0.0)


@pytest.mark.ref
@pytest.mark.parametrize("seed", [21, 22])
def test_live_reference_random_systems(seed):
    """Fresh random periodic systems: oracle vs the reference build."""
    s = configs.small_pbc(5, 26, 110.0, seed=seed, tempK=290.0 + seed)
    ref = refdrv.run(s, [("tape", 1, 200000, "t"), ("nl", 1), ("lists", 1, "L"), ("force", 1, 1, "F"),
                         ("bp_state", 1, "B")])
    o = corc.Oracle(s)
    o.set_tape(1, ref["t"])
    o.neighbor_list(1)
    L = o.lists(1)
    for k in ("wca", "ele", "bp"):
        assert np.array_equal(L[k], ref["L"][k])
    f = o.force(1, 1)
    assert np.abs(f - ref["F"]).max() / np.abs(ref["F"]).max() < 1e-12
    assert np.array_equal(o.bp_state(1)[0], ref["B"][0].astype(np.uint8))


@pytest.mark.ref
def test_live_reference_shared_mc_flag():
    """Replica loop with limit MC every 3 steps and a 0.04 A skin, so that
    rebuilds fall on non-MC steps: the reference's single flg_bp_MC makes a
    rebuild in replica r force MC on the later replicas of that step
    (src/job_md.F90:364-401).  Oracle trajectory vs the reference build."""
    s = configs.remd_c3(nrep=4, n=40, seed=9, margin=0.04)
    s["params"] = dict(s["params"], nstep_bp_mc=3)
    nstep, nex = 40, 10
    ops = [("tape", r, 400000, f"t{r}") for r in range(1, 5)] + [("tape", -1, 4096, "tr")]
    ref0 = refdrv.run(s, ops, seed=99, nstep_exchange=nex)
    ref1 = refdrv.run(s, [("job_md", nstep, "R")], seed=99, nstep_exchange=nex)["R"]
    o = corc.Oracle(s)
    for r in range(1, 5):
        o.set_tape(r, ref0[f"t{r}"])
    o.set_tape(-1, ref0["tr"])
    o.init_velocities()
    nreb, _ = o.md(nstep, nex)
    assert nreb >= 4   # rebuilds happen inside the run
    x, v, a = o.state()
    assert np.abs(x - ref1["xyz"]).max() < 1e-10
    assert np.abs(v - ref1["velos"]).max() < 1e-12
    assert np.array_equal(o.labels(), ref1["rep2lab"][:4])


@pytest.mark.ref
def test_live_reference_kinetic_energy():
    """energy_kinetic: oracle against the reference build on random
    velocities of every replica (masses from the sequence)."""
    s = configs.remd_c3(nrep=3, n=50, seed=5)
    rng = np.random.default_rng(12)
    v = rng.normal(0.0, 2.0, size=(3, 50, 3))
    ref = refdrv.run(s, [("set_velos", v)] + [("kinetic", r, f"k{r}") for r in (1, 2, 3)])
    o = corc.Oracle(s)
    o.set_state(velos=v)
    for r in (1, 2, 3):
        assert abs(o.kinetic(r) - ref[f"k{r}"]) <= 1e-12 * ref[f"k{r}"]


def test_kinetic_energy_closed_form():
    """0.5 sum m v^2 with the masses of the sequence (A, U, G, C differ):
    oracle against numpy, and zero for zero velocities."""
    s = configs.remd_c3(nrep=2, n=30, seed=6)
    o = corc.Oracle(s)
    rng = np.random.default_rng(3)
    v = rng.normal(size=(2, 30, 3))
    o.set_state(velos=v)
    # masses per nucleotide type, src/job_md.F90:116-133 (A, U, G, C codes 0-3)
    m = np.array([328.212, 305.164, 344.212, 304.182])[np.asarray(s["seq"])]
    for r in (1, 2):
        want = 0.5 * float(np.sum(m[:, None] * v[r - 1] ** 2))
        assert abs(o.kinetic(r) - want) <= 1e-12 * want
    o.set_state(velos=np.zeros_like(v))
    assert o.kinetic(1) == 0.0
