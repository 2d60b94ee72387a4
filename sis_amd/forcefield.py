"""Force-field and run parameters for the benchmark systems.

Values are data taken from the reference's force-field files:
  * bond / angle_ReB / dihedral_exp / basepair blocks: check/sis_dih_ReB_NN.ff:3-85
    (bond k and r0 overridden to 15.0 and 5.9 for the benchmark systems)
  * nearest-neighbour dH / dS (Zuber et al. 2022): sis.ff:125-169
  * electrostatics defaults: check/input_T2S2HP.toml:34-39 (cutoff type 1, 50 A)
Units: Angstrom, kcal/mol, K, reduced time (tau = 48.888 fs).
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
