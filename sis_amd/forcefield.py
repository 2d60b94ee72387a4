"""Force-field and run parameters for the benchmark systems.

Values are data taken from the reference's force-field files:
  * bond / angle_ReB / dihedral_exp / basepair blocks: check/sis_dih_ReB_NN.ff:3-85
    (bond k and r0 overridden to 15.0 and 5.9 for the benchmark systems)
  * nearest-neighbour dH / dS (Zuber et al. 2022): sis.ff:125-169
  * electrostatics of the benchmark systems: ionic strength and length per
    charge of check/input_T2S2HP.toml:33-38; cutoff type 1 at 50 A, the
    cutoff of the condensate benchmark workloads (that input file itself uses
    cutoff type 2, 10 Debye lengths; read_ff below plus a [Electrostatic]
    override reproduces it)
Units: Angstrom, kcal/mol, K, reduced time (tau = 48.888 fs).

read_ff() turns a reference force-field file (TOML, src/read_force_field.F90)
into the same flat parameter dict.
"""
from __future__ import annotations

import copy

NN_DH = {
    "GC_CG": -16.52, "CC_GG": -13.94, "GA_CU": -13.75, "CG_GC": -9.61, "AC_UG": -11.98,
    "CA_GU": -10.47, "AG_UC": -9.34, "UA_AU": -9.16, "AU_UA": -8.91, "AA_UU": -7.44,
    "GC_UG": -14.73, "CU_GG": -9.26, "GG_CU": -12.41, "CG_GU": -5.64, "AU_UG": -9.23,
    "GA_UU": -10.58, "UG_GU": -8.76, "UA_GU": -2.72, "GG_UU": -9.06, "GU_UG": -7.66,
    "AG_UU": -5.10,
}
NN_DS = {
    "GC_CG": -42.13, "CC_GG": -34.41, "GA_CU": -36.53, "CG_GC": -23.46, "AC_UG": -31.37,
    "CA_GU": -27.08, "AG_UC": -23.66, "UA_AU": -25.40, "AU_UA": -25.22, "AA_UU": -20.98,
    "GC_UG": -40.32, "CU_GG": -23.64, "GG_CU": -34.23, "CG_GU": -14.83, "AU_UG": -27.32,
    "GA_UU": -32.19, "UG_GU": -27.04, "UA_GU": -8.08, "GG_UU": -28.57, "GU_UG": -24.11,
    "AG_UU": -16.53,
}

_BP = {
    "GC": dict(U0=3.0, bond_k=1.86, bond_r=13.7, angl_k1=12.0, angl_k2=12.0, angl_k3=15.0,
               angl_k4=15.0, angl_theta1=1.83, angl_theta2=1.83, angl_theta3=0.95,
               angl_theta4=0.95, dihd_k1=5.0, dihd_k2=4.0, dihd_phi1=1.85, dihd_phi2=0.97),
    "AU": dict(U0=3.0, bond_k=1.31, bond_r=13.6, angl_k1=12.0, angl_k2=12.0, angl_k3=15.0,
               angl_k4=15.0, angl_theta1=1.81, angl_theta2=1.81, angl_theta3=0.98,
               angl_theta4=0.98, dihd_k1=5.0, dihd_k2=4.0, dihd_phi1=1.85, dihd_phi2=0.93),
    "GU": dict(U0=3.0, bond_k=1.20, bond_r=13.5, angl_k1=12.0, angl_k2=12.0, angl_k3=20.0,
               angl_k4=20.0, angl_theta1=1.80, angl_theta2=1.80, angl_theta3=0.98,
               angl_theta4=0.98, dihd_k1=5.0, dihd_k2=4.0, dihd_phi1=1.80, dihd_phi2=0.99),
}


def _nht22_bp(u0):
    # para/rna_NHT22.ff:14-50; the file sets no angl_k3/k4 terms
    return dict(U0=u0, bond_k=3.0, bond_r=13.8, angl_k1=1.5, angl_k2=1.5, angl_k3=0.0, angl_k4=0.0,
                angl_theta1=1.8326, angl_theta2=0.9425, angl_theta3=0.0, angl_theta4=0.0,
                dihd_k1=0.5, dihd_k2=0.5, dihd_phi1=1.8326, dihd_phi2=1.1345)


def nht22_params(**over) -> dict:
    """The NHT22 force field (para/rna_NHT22.ff): harmonic bond and angle,
    no dihedral, base-pair model 1 (U0 exp(-u), no limit MC), WCA; with the
    electrostatics of check/test_DCD_CAG47/input_DCD_CAG47.toml (295.15 K,
    0.15 M, 50 A cutoff)."""
    p = default_params(tempK=295.15, nl_margin=5.0, bond_k=15.0, bond_r0=5.9, angl_reb=0, angl_k=10.0,
                       angl_t0=2.618, dih_type=0, wca_sigma=10.0, wca_eps=2.0, ionic_strength=0.15,
                       ele_cutoff_type=1, ele_cutoff=50.0, bp_model=1, nstep_bp_mc=0, bp_cutoff_energy=-0.01,
                       bp={"GC": _nht22_bp(-5.0), "AU": _nht22_bp(-3.33333), "GU": _nht22_bp(-3.33333)})
    p.update(over)
    return p


def default_params(**over) -> dict:
    """Flat parameter dict matching struct sis_params."""
    from .abi import NNT_ORDER
    p = dict(
        bond_k=15.0, bond_r0=5.9,
        angl_reb=1, angl_k=5.0, angl_t0=2.643,
        dih_type=2, dih_k=1.4, dih_w=3.0, dih_p0=0.28,
        wca_sigma=10.0, wca_eps=2.0,
        ele_on=1, ele_cutoff_type=1, ele_cutoff=50.0, length_per_charge=4.38178046,
        ionic_strength=0.15, ele_excl_bond=1, ele_excl_angl=1, dummy_has_charge=0,
        bp_model=5, bp_min_loop=3, max_bp_per_nt=1, nstep_bp_mc=1,
        bp_cutoff_energy=0.0, bp_dH0=0.0, bp_dS0=-8.0, bp_dHend0=0.0, bp_dSend0=0.0,
        bp_coef_dG=1.0, nnend=0,
        nn_dH=[NN_DH[k] for k in NNT_ORDER], nn_dS=[NN_DS[k] for k in NNT_ORDER],
        nnend_dH=[0.0] * 6, nnend_dS=[0.0] * 6,
        bp=copy.deepcopy(_BP),
        tempK=300.0, dt=0.05, viscosity_pas=1.4115e-5, nl_margin=10.0,
        pbc=0, box=[0.0, 0.0, 0.0],
        integrator=1,            # 1 GJF-2GJ Langevin, 2 Ermak-McCammon Brownian
    )
    p.update(over)
    return p


def _lower_keys(t):
    if isinstance(t, dict):
        return {str(k).lower(): _lower_keys(v) for k, v in t.items()}
    return t


def read_ff(path, **over) -> dict:
    """Parameters of a reference force-field file, as src/read_force_field.F90
    reads it: [potential.bond] k/r0; [potential.angle] (harmonic) or
    [potential.angle_ReB] k/a0; [potential.dihedral] (cosine: k, phi0) or
    [potential.dihedral_exp] (k, w, phi0), neither meaning no dihedral term;
    [potential.basepair] cutoff_energy, coef_dG, dH0, dS0, NNend, dHend0,
    dSend0 and the GC/AU/GU blocks; [potential.wca] sigma/epsilon; [NN.dH] and
    [NN.dS] by stack name; [NNend.dH]/[NNend.dS] when NNend is true.  Keys are
    matched case-insensitively, as the reference's TOML reader does.  Block
    fields a file leaves out are 0.  Everything the file does not set (run
    conditions, electrostatics, base-pair model) comes from default_params()
    and `over`."""
    import tomli
    from .abi import NNT_ORDER, NNEND_ORDER, BP_FIELDS
    with open(path, "rb") as fh:
        t = _lower_keys(tomli.load(fh))
    pot = t["potential"]
    p = default_params()
    b = pot["bond"]
    p.update(bond_k=float(b["k"]), bond_r0=float(b["r0"]))
    if "angle" in pot and "angle_reb" in pot:   # src/read_force_field.F90:85-105
        raise ValueError("[potential.angle] and [potential.angle_ReB] cannot be specified together")
    if "angle" in pot:
        p.update(angl_reb=0, angl_k=float(pot["angle"]["k"]), angl_t0=float(pot["angle"]["a0"]))
    elif "angle_reb" in pot:
        p.update(angl_reb=1, angl_k=float(pot["angle_reb"]["k"]), angl_t0=float(pot["angle_reb"]["a0"]))
    if "dihedral" in pot and "dihedral_exp" in pot:   # :133-137
        raise ValueError("[dihedral] and [dihedral_exp] cannot be specified together")
    if "dihedral" in pot:
        d = pot["dihedral"]
        p.update(dih_type=1, dih_k=float(d["k"]), dih_w=0.0, dih_p0=float(d["phi0"]))
    elif "dihedral_exp" in pot:
        d = pot["dihedral_exp"]
        p.update(dih_type=2, dih_k=float(d["k"]), dih_w=float(d["w"]), dih_p0=float(d["phi0"]))
    else:
        p.update(dih_type=0)
    bp = pot["basepair"]
    p.update(bp_cutoff_energy=float(bp.get("cutoff_energy", 0.0)), bp_coef_dG=float(bp.get("coef_dg", 1.0)),
             bp_dH0=float(bp.get("dh0", 0.0)), bp_dS0=float(bp.get("ds0", 0.0)),
             nnend=int(bool(bp.get("nnend", False))),
             bp_dHend0=float(bp.get("dhend0", 0.0)), bp_dSend0=float(bp.get("dsend0", 0.0)))
    p["bp"] = {}
    for ty in ("GC", "AU", "GU"):
        blk = bp.get(ty.lower(), {})
        p["bp"][ty] = {f: float(blk.get(f.lower(), 0.0)) for f in BP_FIELDS if f != "cutoff_ddist"}
    w = pot["wca"]
    p.update(wca_sigma=float(w["sigma"]), wca_eps=float(w["epsilon"]))
    nn = t.get("nn", {})
    if "dh" in nn and "ds" in nn:
        p["nn_dH"] = [float(nn["dh"][k.lower()]) for k in NNT_ORDER]
        p["nn_dS"] = [float(nn["ds"][k.lower()]) for k in NNT_ORDER]
    if p["nnend"]:
        ne = t["nnend"]
        p["nnend_dH"] = [float(ne["dh"][k.lower()]) for k in NNEND_ORDER]
        p["nnend_dS"] = [float(ne["ds"][k.lower()]) for k in NNEND_ORDER]
    p.update(over)
    return p
