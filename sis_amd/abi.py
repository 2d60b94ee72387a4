"""ctypes mirror of the C ABI declared in include/sis_gpu.h.

The structs here have exactly the layout of ``struct sis_bp_para`` and
``struct sis_params``; the same layout is bound from Fortran by
``sis_amd/fortran/sis_gpu_iface.F90``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libsis_gpu.so"

# Enumerations (values of the reference's const_idx module,
# src/const_idx.F90:7-69).
SEQT = {"A": 0, "U": 1, "G": 2, "C": 3, "D": 4}
SEQT_CHAR = {v: k for k, v in SEQT.items()}
MOLT_RNA, MOLT_CIRCRNA = 1, 2
BPT = {"GC": 1, "AU": 2, "GU": 3, "CG": 4, "UA": 5, "UG": 6}
ENE = {"TOTAL": 0, "BOND": 1, "ANGL": 2, "DIHE": 3, "BP": 4, "EXV": 5, "ELE": 6,
       "STAGE": 7, "TWZ": 8, "RG": 9, "REST": 10}
NNT_ORDER = ["GC_CG", "CC_GG", "GA_CU", "CG_GC", "AC_UG", "CA_GU", "AG_UC", "UA_AU",
             "AU_UA", "AA_UU", "GC_UG", "CU_GG", "GG_CU", "CG_GU", "AU_UG", "GA_UU",
             "UG_GU", "UA_GU", "GG_UU", "GU_UG", "AG_UU"]
NNEND_ORDER = ["AUonAU", "AUonCG", "AUonGU", "GUonCG", "GUonAU", "GUonGU"]

BP_FIELDS = ["cutoff_ddist", "U0", "bond_k", "bond_r",
             "angl_k1", "angl_k2", "angl_k3", "angl_k4",
             "angl_theta1", "angl_theta2", "angl_theta3", "angl_theta4",
             "dihd_k1", "dihd_k2", "dihd_phi1", "dihd_phi2"]


class SisBpPara(C.Structure):
    _fields_ = [(n, C.c_double) for n in BP_FIELDS]


class SisParams(C.Structure):
    _fields_ = [
        ("bond_k", C.c_double), ("bond_r0", C.c_double),
        ("angl_k", C.c_double), ("angl_t0", C.c_double),
        ("dih_k", C.c_double), ("dih_w", C.c_double), ("dih_p0", C.c_double),
        ("wca_sigma", C.c_double), ("wca_eps", C.c_double),
        ("ele_cutoff", C.c_double), ("length_per_charge", C.c_double),
        ("ionic_strength", C.c_double),
        ("bp_cutoff_energy", C.c_double),
        ("bp_dH0", C.c_double), ("bp_dS0", C.c_double),
        ("bp_dHend0", C.c_double), ("bp_dSend0", C.c_double), ("bp_coef_dG", C.c_double),
        ("nn_dH", C.c_double * 21), ("nn_dS", C.c_double * 21),
        ("nnend_dH", C.c_double * 6), ("nnend_dS", C.c_double * 6),
        ("bp", SisBpPara * 3),
        ("tempK", C.c_double),
        ("dt", C.c_double), ("viscosity_pas", C.c_double), ("nl_margin", C.c_double),
        ("box", C.c_double * 3),
        ("angl_reb", C.c_int32), ("dih_type", C.c_int32),
        ("ele_on", C.c_int32), ("ele_cutoff_type", C.c_int32),
        ("ele_excl_bond", C.c_int32), ("ele_excl_angl", C.c_int32),
        ("dummy_has_charge", C.c_int32),
        ("bp_model", C.c_int32), ("bp_min_loop", C.c_int32),
        ("max_bp_per_nt", C.c_int32), ("nstep_bp_mc", C.c_int32), ("nnend", C.c_int32),
        ("pbc", C.c_int32), ("integrator", C.c_int32),
    ]


class SisExternal(C.Structure):
    """struct sis_external: stage wall, tweezers, Rg bias, restraints."""
    _fields_ = [
        ("stage", C.c_int32), ("n_dcf", C.c_int32), ("n_fr", C.c_int32), ("rg", C.c_int32),
        ("n_rest", C.c_int32), ("pad", C.c_int32),
        ("stage_sigma", C.c_double), ("stage_eps", C.c_double), ("rg_k", C.c_double),
        ("rg_0", C.POINTER(C.c_double)),
        ("dcf_pairs", C.POINTER(C.c_int32)), ("dcf_forces", C.POINTER(C.c_double)),
        ("fr_pairs", C.POINTER(C.c_int32)), ("fr_k", C.POINTER(C.c_double)),
        ("fr_init", C.POINTER(C.c_double)), ("fr_velo", C.POINTER(C.c_double)),
        ("rest_id", C.POINTER(C.c_int32)), ("rest_rcut", C.POINTER(C.c_double)),
        ("rest_para", C.POINTER(C.c_double)),
    ]


def external_to_struct(ext: dict, nrep: int):
    """Fill a SisExternal from a dict: stage, stage_sigma, stage_eps;
    dcf_pairs (n, 2) 1-based, dcf_forces (nrep, n, 3); fr_pairs (n, 2),
    fr_k (n, 2), fr_init / fr_velo (n, 2, 3); rg (0 off, 1 harmonic,
    2 flat-bottomed), rg_k, rg_0 (nrep); rest_id (n, 3), rest_rcut (n),
    rest_para (n, 3).  Returns (struct, arrays to keep alive)."""
    s = SisExternal()
    keep = []

    def arr(name, dtype, shape_tail):
        a = ext.get(name)
        if a is None:
            return None, 0
        a = np.ascontiguousarray(a, dtype)
        keep.append(a)
        return a, (a.shape[0] if a.ndim else 0)

    s.stage = int(bool(ext.get("stage", 0)))
    s.stage_sigma = float(ext.get("stage_sigma", 10.0))
    s.stage_eps = float(ext.get("stage_eps", 1.2))
    a, n = arr("dcf_pairs", np.int32, 2)
    if a is not None:
        s.n_dcf = n
        s.dcf_pairs = _ptr(a, C.c_int32)
        f = np.ascontiguousarray(ext["dcf_forces"], np.float64).reshape(nrep, n, 3)
        keep.append(f)
        s.dcf_forces = _ptr(f, C.c_double)
    a, n = arr("fr_pairs", np.int32, 2)
    if a is not None:
        s.n_fr = n
        s.fr_pairs = _ptr(a, C.c_int32)
        for nm, shp in (("fr_k", (n, 2)), ("fr_init", (n, 2, 3)), ("fr_velo", (n, 2, 3))):
            b = np.ascontiguousarray(ext[nm], np.float64).reshape(shp)
            keep.append(b)
            setattr(s, nm, _ptr(b, C.c_double))
    s.rg = int(ext.get("rg", 0))
    if s.rg:
        s.rg_k = float(ext["rg_k"])
        r0 = np.ascontiguousarray(np.broadcast_to(np.asarray(ext["rg_0"], np.float64), (nrep,)))
        keep.append(r0)
        s.rg_0 = _ptr(r0, C.c_double)
    a, n = arr("rest_id", np.int32, 3)
    if a is not None:
        s.n_rest = n
        s.rest_id = _ptr(a, C.c_int32)
        rc = np.ascontiguousarray(ext["rest_rcut"], np.float64).reshape(n)
        rp = np.ascontiguousarray(ext["rest_para"], np.float64).reshape(n, 3)
        keep += [rc, rp]
        s.rest_rcut = _ptr(rc, C.c_double)
        s.rest_para = _ptr(rp, C.c_double)
    return s, keep


def params_to_struct(p: dict) -> SisParams:
    """Fill a SisParams from a flat dict (see forcefield.default_params)."""
    s = SisParams()
    for name, ctype in SisParams._fields_:
        if name == "bp":
            for k, key in enumerate(("GC", "AU", "GU")):
                blk = p["bp"][key]
                for f in BP_FIELDS:
                    setattr(s.bp[k], f, float(blk.get(f, 0.0)))
            continue
        if name not in p:
            continue
        val = p[name]
        if hasattr(ctype, "_length_"):
            arr = getattr(s, name)
            for i, v in enumerate(val):
                arr[i] = v
        else:
            setattr(s, name, val)
    return s


def load_library(path: os.PathLike | None = None) -> C.CDLL:
    """Load libsis_gpu.so.  Raises if it has not been built: there is no
    fallback implementation of the hot path."""
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    _declare(lib)
    return lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


dptr = lambda a: _ptr(a, C.c_double)  # noqa: E731
iptr = lambda a: _ptr(a, C.c_int32)  # noqa: E731
u32ptr = lambda a: _ptr(a, C.c_uint32)  # noqa: E731
i64ptr = lambda a: _ptr(a, C.c_int64)  # noqa: E731
u8ptr = lambda a: _ptr(a, C.c_uint8)  # noqa: E731


EXPORTED = [
    "sis_gpu_create", "sis_gpu_destroy", "sis_gpu_last_error",
    "sis_gpu_set_state", "sis_gpu_get_state", "sis_gpu_set_positions", "sis_gpu_get_rng_tape_pos", "sis_gpu_set_rep_conditions",
    "sis_gpu_rebuild", "sis_gpu_analysis_lists", "sis_gpu_force", "sis_gpu_energy", "sis_gpu_energy_noalloc", "sis_gpu_autotune",
    "sis_gpu_get_pair_count", "sis_gpu_get_pairs", "sis_gpu_get_bp",
    "sis_gpu_set_rng", "sis_gpu_set_rng_tape", "sis_gpu_init_velocities",
    "sis_gpu_step", "sis_gpu_exchange", "sis_gpu_set_rep_dims", "sis_gpu_get_labels",
    "sis_gpu_get_info", "sis_gpu_last_kernel_ms",
    "sis_pool_create", "sis_pool_submit", "sis_pool_poll", "sis_pool_forget",
    "sis_pool_stats", "sis_pool_destroy",
    "sis_gpu_dcd_open", "sis_gpu_dcd_write", "sis_gpu_dcd_close", "sis_dcd_write_file",
    "sis_gpu_set_step", "sis_gpu_set_labels",
    "sis_gpu_bp_open", "sis_gpu_bp_write", "sis_gpu_bp_close", "sis_bp_write_file",
    "sis_gpu_kinetic", "sis_gpu_out_open", "sis_gpu_out_close", "sis_out_write_file",
    "sis_gpu_dcd_energies", "sis_dcd_read_file", "sis_gpu_check_force",
    "sis_gpu_rst_open", "sis_gpu_rst_write", "sis_gpu_rst_close", "sis_gpu_rst_read",
    "sis_rst_write_file", "sis_rst_read_file", "sis_gpu_set_external", "sis_gpu_set_ss",
    "sis_gpu_set_anneal", "sis_gpu_set_box_resize", "sis_gpu_out_write",
    "sis_gpu_set_rg_schedule", "sis_gpu_set_rep_rg", "sis_gpu_set_rep_dcf",
    "sis_gpu_twz_open", "sis_gpu_twz_close",
]


def _declare(lib: C.CDLL) -> None:
    vp = C.c_void_p
    i32, i64, d = C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    sig = {
        "sis_gpu_create": (i32, [P(SisTopo), P(SisParams), P(SisOptions), P(vp)]),
        "sis_gpu_destroy": (i32, [vp]),
        "sis_gpu_last_error": (C.c_char_p, []),
        "sis_gpu_set_state": (i32, [vp, P(d), P(d), P(d)]),
        "sis_gpu_get_state": (i32, [vp, P(d), P(d), P(d)]),
        "sis_gpu_set_positions": (i32, [vp, i32, P(d)]),
        "sis_gpu_get_rng_tape_pos": (i32, [vp, i32, P(i64)]),
        "sis_gpu_set_rep_conditions": (i32, [vp, P(d), P(d)]),
        "sis_gpu_rebuild": (i32, [vp, i32]),
        "sis_gpu_analysis_lists": (i32, [vp]),
        "sis_gpu_force": (i32, [vp, i32, i32, P(d)]),
        "sis_gpu_energy": (i32, [vp, i32, d, i32, P(d)]),
        "sis_gpu_energy_noalloc": (i32, [vp, i32, d, i32, P(d)]),
        "sis_gpu_autotune": (i32, [vp, P(d), i32, i64, P(d)]),
        "sis_gpu_get_pair_count": (i32, [vp, i32, i32, P(i64)]),
        "sis_gpu_get_pairs": (i32, [vp, i32, i32, P(i32), i64, P(i64)]),
        "sis_gpu_get_bp": (i32, [vp, i32, P(i32), P(d), P(C.c_uint8), P(d), i64, P(i64)]),
        "sis_gpu_set_rng": (i32, [vp, C.c_uint64]),
        "sis_gpu_set_rng_tape": (i32, [vp, i32, P(C.c_uint32), i64]),
        "sis_gpu_init_velocities": (i32, [vp]),
        "sis_gpu_step": (i32, [vp, i64, P(i64)]),
        "sis_gpu_exchange": (i32, [vp, P(d), P(i32)]),
        "sis_gpu_set_rep_dims": (i32, [vp, i32, P(i32)]),
        "sis_gpu_get_labels": (i32, [vp, P(i32)]),
        "sis_gpu_get_info": (i32, [vp, P(i64), i32]),
        "sis_gpu_last_kernel_ms": (d, [vp, i32]),
        "sis_pool_create": (i32, [C.c_uint64, i64, i32, P(vp)]),
        "sis_pool_submit": (i32, [vp, P(SisTopo), P(SisParams), P(d), i32, i64, C.c_uint64, P(i64)]),
        "sis_pool_poll": (i32, [vp, i64, P(i32), P(d), P(d)]),
        "sis_pool_forget": (i32, [vp, i64]),
        "sis_pool_stats": (i32, [vp, P(d), i32]),
        "sis_pool_destroy": (i32, [vp]),
        "sis_gpu_dcd_open": (i32, [vp, P(C.c_char_p), i32, i64]),
        "sis_gpu_dcd_write": (i32, [vp]),
        "sis_gpu_set_step": (i32, [vp, i64]),
        "sis_gpu_kinetic": (i32, [vp, i32, P(d)]),
        "sis_gpu_out_open": (i32, [vp, P(C.c_char_p), i64]),
        "sis_gpu_out_close": (i32, [vp]),
        "sis_gpu_out_write": (i32, [vp]),
        "sis_gpu_dcd_energies": (i32, [vp, C.c_char_p, C.c_char_p, i64, P(d), P(i64)]),
        "sis_gpu_check_force": (i32, [vp, i32, d, C.c_char_p, P(d), P(d)]),
        "sis_dcd_read_file": (i32, [C.c_char_p, i32, i64, P(d), P(d), P(i64)]),
        "sis_out_write_file": (i32, [C.c_char_p, i32, P(i64), P(i32), P(d), P(d), P(d), i32, i32]),
        "sis_gpu_bp_open": (i32, [vp, P(C.c_char_p), i32, i64]),
        "sis_gpu_bp_write": (i32, [vp]),
        "sis_gpu_bp_close": (i32, [vp]),
        "sis_bp_write_file": (i32, [C.c_char_p, i64, P(i32), P(C.c_uint8), P(d), i32, d, i32]),
        "sis_gpu_set_labels": (i32, [vp, P(i32)]),
        "sis_gpu_set_external": (i32, [vp, P(SisExternal)]),
        "sis_gpu_set_ss": (i32, [vp, i32, P(i32)]),
        "sis_gpu_set_anneal": (i32, [vp, i32, P(i64), P(d)]),
        "sis_gpu_set_box_resize": (i32, [vp, i64, P(d), P(d)]),
        "sis_gpu_set_rg_schedule": (i32, [vp, i32, P(i64), P(d), P(d)]),
        "sis_gpu_set_rep_rg": (i32, [vp, P(d)]),
        "sis_gpu_set_rep_dcf": (i32, [vp, P(d), P(d)]),
        "sis_gpu_twz_open": (i32, [vp, C.c_char_p, i64]),
        "sis_gpu_twz_close": (i32, [vp]),
        "sis_gpu_rst_open": (i32, [vp, P(C.c_char_p), i64]),
        "sis_gpu_rst_write": (i32, [vp]),
        "sis_gpu_rst_close": (i32, [vp]),
        "sis_gpu_rst_read": (i32, [vp, P(C.c_char_p)]),
        "sis_rst_write_file": (i32, [C.c_char_p, i32, i32, i32, i64, P(i32), P(d), P(d), P(d),
                                     P(C.c_uint32), P(C.c_uint32), P(d), P(C.c_uint64)]),
        "sis_rst_read_file": (i32, [C.c_char_p, i32, i32, P(i32), P(i32), P(i64), P(i32), P(d), P(d), P(d),
                                    P(d), P(C.c_uint64)]),
        "sis_gpu_dcd_close": (i32, [vp]),
        "sis_dcd_write_file": (i32, [C.c_char_p, i32, i32, P(d), i32, P(d), i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


class SisTopo(C.Structure):
    """struct sis_topo: chains, per-chain molecule type, per-bead sequence."""
    _fields_ = [
        ("nchains", C.c_int32), ("nmp", C.c_int32),
        ("nmp_chain", C.POINTER(C.c_int32)),
        ("moltype", C.POINTER(C.c_int32)),
        ("seq", C.POINTER(C.c_int32)),
        ("no_charge", C.POINTER(C.c_int32)), ("n_no_charge", C.c_int32),
        ("pad", C.c_int32),
    ]


class SisOptions(C.Structure):
    """struct sis_options: replicas, precision, parity and launch knobs."""
    _fields_ = [
        ("nrep", C.c_int32),
        ("precision", C.c_int32),      # 0 = FP64 pair math, 1 = FP32 pair math
        ("parity", C.c_int32),         # 1 = consume an injected MT word tape
        ("mode", C.c_int32),           # 0 = auto, 1 = per-kernel, 2 = persistent, 3 = batched
        ("block", C.c_int32),          # batched form: threads per replica (0 = 256; 128/256/512)
        ("max_nbr", C.c_int32),        # neighbour slots per bead (0 = auto)
        ("rebuild_every", C.c_int32),  # 0 = displacement trigger as the reference
        ("exchange_every", C.c_int32),  # replica exchange cadence (0 = none)
        ("seed", C.c_uint64),
        ("grid", C.c_int32),           # persistent form: workgroups (0 = as many as fit)
        ("flags", C.c_int32),          # bit0: batched exchanges on the host
    ]
