"""ctypes wrapper of the C restatement (oracle/liboracle.so).  Test oracle
only: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg."""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle.so"
sys.path.insert(0, str(HERE.parent))
from sis_amd import abi  # noqa: E402
from sis_amd.abi import SisParams, SisTopo, params_to_struct  # noqa: E402

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not SO.exists():
            raise RuntimeError(f"{SO} not built (make -C oracle liboracle.so)")
        L = C.CDLL(str(SO))
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        P = C.POINTER
        L.orc_create.restype = vp
        L.orc_create.argtypes = [P(SisTopo), P(SisParams), i32, P(d)]
        L.orc_destroy.argtypes = [vp]
        for n in ("orc_set_xyz",):
            getattr(L, n).argtypes = [vp, P(d)]
        L.orc_get_state.argtypes = [vp, P(d), P(d), P(d)]
        L.orc_set_state.argtypes = [vp, P(d), P(d), P(d)]
        L.orc_neighbor_list.argtypes = [vp, i32]
        L.orc_count.restype = i64
        L.orc_count.argtypes = [vp, i32, i32]
        L.orc_lists.argtypes = [vp, i32, i32, P(i32), P(d)]
        L.orc_force.argtypes = [vp, i32, i32, P(d)]
        L.orc_energy.argtypes = [vp, i32, d, i32, i32, P(d)]
        L.orc_bp_state.argtypes = [vp, i32, P(C.c_uint8), P(d)]
        L.orc_tables.argtypes = [vp, P(d), P(d), P(d)]
        L.orc_bp_pair.argtypes = [vp, i32, i32, P(i32), P(i32)]
        L.orc_ele_params.argtypes = [vp, i32, P(d)]
        L.orc_md_coef.argtypes = [vp, i32, P(d)]
        L.orc_set_external.argtypes = [vp, P(abi.SisExternal)]
        L.orc_set_step.argtypes = [vp, i64]
        L.orc_set_anneal.argtypes = [vp, i32, P(i64), P(d)]
        L.orc_set_ionic.argtypes = [vp, P(d)]
        L.orc_plant.argtypes = [vp, i32, i32]
        L.orc_set_counter_rng.argtypes = [vp, C.c_uint64, i32]
        L.orc_set_nl_brute.argtypes = [vp, i32]
        L.orc_analysis_lists.argtypes = [vp, i32]
        L.orc_set_rep_dims.argtypes = [vp, i32, P(i32)]
        L.orc_pair_table.argtypes = [vp, P(i32)]
        L.orc_philox.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.orc_set_rep_rg.argtypes = [vp, P(d)]
        L.orc_set_rep_dcf.argtypes = [vp, P(d), P(d)]
        L.orc_set_rg_schedule.argtypes = [vp, i32, P(i64), P(d), P(d)]
        L.orc_set_box_resize.argtypes = [vp, i64, P(d), P(d)]
        L.orc_get_box.argtypes = [vp, P(d)]
        L.orc_set_ss.restype = i32
        L.orc_set_ss.argtypes = [vp, i32, P(i32)]
        L.orc_set_tape.argtypes = [vp, i32, P(C.c_uint32), i64]
        L.orc_tape_pos.restype = i64
        L.orc_tape_pos.argtypes = [vp, i32]
        L.orc_boxmuller.argtypes = [vp, i32, i64, P(d)]
        L.orc_mt_words.argtypes = [C.c_uint32, i64, P(C.c_uint32)]
        L.orc_genrand.argtypes = [vp, i32, i32, i64, P(d)]
        L.orc_init_velocities.argtypes = [vp]
        L.orc_kinetic.restype = d
        L.orc_kinetic.argtypes = [vp, i32]
        L.orc_md.restype = i64
        L.orc_md.argtypes = [vp, i64, i32, P(d)]
        L.orc_labels.argtypes = [vp, P(i32)]
        L.orc_set_temps.argtypes = [vp, P(d)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_topo(sysd):
    keep = dict(
        nmp_chain=np.ascontiguousarray(sysd["nmp_chain"], np.int32),
        moltype=np.ascontiguousarray(sysd["moltype"], np.int32),
        seq=np.ascontiguousarray(sysd["seq"], np.int32),
    )
    t = SisTopo()
    t.nchains = int(sysd["nchains"])
    t.nmp = int(keep["seq"].size)
    t.nmp_chain = _i(keep["nmp_chain"])
    t.moltype = _i(keep["moltype"])
    t.seq = _i(keep["seq"])
    t.no_charge = None
    t.n_no_charge = 0
    return t, keep


def mt_words(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_mt_words(C.c_uint32(seed), n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


class Oracle:
    """One system in the C restatement; methods mirror the reference's
    subroutines (neighbor_list, force, energy_sumup, job_md)."""

    def __init__(self, sysd: dict):
        self.sysd = sysd
        self.nmp = int(np.sum(sysd["nmp_chain"]))
        self.nrep = int(sysd["nrep"])
        self.topo, self._keep = make_topo(sysd)
        self.prm = params_to_struct(sysd["params"])
        self.temps = np.ascontiguousarray(sysd["temps"], np.float64)
        self.h = lib().orc_create(C.byref(self.topo), C.byref(self.prm), self.nrep, _d(self.temps))
        self._tapes = {}
        self.set_xyz(sysd["xyz"])

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def set_xyz(self, xyz):
        x = np.ascontiguousarray(xyz, np.float64).reshape(self.nrep, self.nmp, 3)
        lib().orc_set_xyz(self.h, _d(x))

    def set_state(self, xyz=None, velos=None, accels=None):
        f = lambda a: None if a is None else _d(np.ascontiguousarray(a, np.float64))  # noqa: E731
        keep = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (xyz, velos, accels)]
        lib().orc_set_state(self.h, *[(_d(a) if a is not None else None) for a in keep])
        del f

    def state(self):
        x = np.zeros((self.nrep, self.nmp, 3))
        v = np.zeros_like(x)
        a = np.zeros_like(x)
        lib().orc_get_state(self.h, _d(x), _d(v), _d(a))
        return x, v, a

    def neighbor_list(self, irep=1):
        lib().orc_neighbor_list(self.h, irep)

    def lists(self, irep=1):
        out = {}
        for kind, name, w in ((0, "wca", 2), (1, "ele", 2), (2, "bp", 7)):
            n = lib().orc_count(self.h, irep, kind)
            a = np.zeros((max(n, 1), w), np.int32)
            c = np.zeros((max(n, 1), 2), np.float64)
            lib().orc_lists(self.h, irep, kind, _i(a), _d(c))
            out[name] = a[:n].copy()
            if kind == 2:
                out["bpcoef"] = c[:n].copy()
        return out

    def force(self, irep=1, bp_mc=1):
        f = np.zeros((self.nmp, 3))
        lib().orc_force(self.h, irep, bp_mc, _d(f))
        return f

    def energy(self, irep=1, tK=None, bp_mc=1, bp_energy_valid=0):
        e = np.zeros(11)
        tK = self.temps[irep - 1] if tK is None else tK
        lib().orc_energy(self.h, irep, tK, bp_mc, bp_energy_valid, _d(e))
        return e

    def bp_state(self, irep=1):
        n = lib().orc_count(self.h, irep, 2)
        st = np.zeros(max(n, 1), np.uint8)
        en = np.zeros(max(n, 1))
        lib().orc_bp_state(self.h, irep, st.ctypes.data_as(C.POINTER(C.c_uint8)), _d(en))
        return st[:n].copy(), en[:n].copy()

    def tables(self):
        dh, ds, cut = np.zeros(468), np.zeros(468), np.zeros(7)
        lib().orc_tables(self.h, _d(dh), _d(ds), _d(cut))
        return dh, ds, cut

    def bp_pair(self, imp, jmp):
        t = C.c_int32()
        b = C.c_int32()
        lib().orc_bp_pair(self.h, imp, jmp, C.byref(t), C.byref(b))
        return t.value, b.value

    def ele_params(self, irep=1):
        e = np.zeros(4)
        lib().orc_ele_params(self.h, irep, _d(e))
        return e

    def set_external(self, ext):
        """External terms (abi.external_to_struct dict), None clears."""
        if ext is None:
            lib().orc_set_external(self.h, None)
            return
        s, keep = abi.external_to_struct(ext, self.nrep)
        lib().orc_set_external(self.h, C.byref(s))
        del keep

    def set_box_resize(self, resize_step, change, target=None):
        ch = np.ascontiguousarray(change, np.float64)
        tg = np.ascontiguousarray(target, np.float64) if target is not None else None
        lib().orc_set_box_resize(self.h, int(resize_step), _d(ch), _d(tg) if tg is not None else None)

    def box(self):
        b = np.zeros(3)
        lib().orc_get_box(self.h, _d(b))
        return b

    def set_rg_schedule(self, steps, k, r0):
        st = np.ascontiguousarray(steps, np.int64)
        kk = np.ascontiguousarray(k, np.float64)
        rr = np.ascontiguousarray(r0, np.float64)
        lib().orc_set_rg_schedule(self.h, len(st), st.ctypes.data_as(C.POINTER(C.c_int64)), _d(kk), _d(rr))

    def plant(self, which: int, on: bool = True):
        """Switch on a planted mistake (1: set_ele moves the DH cutoff at
        once), for tests that show a fixture catches it."""
        lib().orc_plant(self.h, which, int(on))

    def set_counter_rng(self, seed: int, on: bool = True):
        """Draw the noise, initial velocities and exchange uniforms from the
        engine's performance-form counter RNG (Philox4x32-10 keyed by seed,
        replica, step and bead) instead of the MT tapes."""
        lib().orc_set_counter_rng(self.h, int(seed), int(on))

    def set_nl_brute(self, on: bool = True):
        """Neighbour lists by the reference's O(N^2) loops even where the
        cell-list enumeration would be used (large periodic systems)."""
        lib().orc_set_nl_brute(self.h, int(on))

    def analysis_lists(self, irep: int = 1):
        """The cutoff-free lists of the DCD / CHECK_FORCE jobs (list_bp,
        list_exv, list_ele) in place of neighbor_list."""
        lib().orc_analysis_lists(self.h, irep)

    def set_rep_dims(self, nvalues):
        """Replica grid of len(nvalues) dimensions (the reference's REPT
        order of the variables that have ladders)."""
        self._dims = np.ascontiguousarray(nvalues, np.int32)
        lib().orc_set_rep_dims(self.h, len(self._dims), self._dims.ctypes.data_as(C.POINTER(C.c_int32)))

    def pair_table(self):
        nset = 2 * len(getattr(self, "_dims", [0]))
        out = np.zeros((nset, self.nrep + 2), np.int32)
        lib().orc_pair_table(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return out

    def set_ionic(self, ionic):
        """Per-label ionic strengths (REMD over REPT%ION)."""
        self._ion = np.ascontiguousarray(ionic, np.float64)
        lib().orc_set_ionic(self.h, _d(self._ion))

    def set_rep_dcf(self, k, direction):
        """f-REMD: force magnitude per label along the normalised direction."""
        self._dk = np.ascontiguousarray(k, np.float64)
        self._dd = np.ascontiguousarray(direction, np.float64)
        lib().orc_set_rep_dcf(self.h, _d(self._dk), _d(self._dd))

    def set_rep_rg(self, rg0):
        """Per-label Rg-bias targets (REMD over REPT%RG)."""
        self._rg = np.ascontiguousarray(rg0, np.float64)
        lib().orc_set_rep_rg(self.h, _d(self._rg))

    def set_anneal(self, steps, temps):
        st = np.ascontiguousarray(steps, np.int64)
        tk = np.ascontiguousarray(temps, np.float64)
        lib().orc_set_anneal(self.h, len(st), st.ctypes.data_as(C.POINTER(C.c_int64)), _d(tk))

    def set_ss(self, pairs):
        """Restrict base pairs to those of a CT/BPSEQ/BPL file, (n, 2)
        1-based bead ids; None clears.  Returns the number kept."""
        if pairs is None:
            return lib().orc_set_ss(self.h, -1, None)
        p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        return int(lib().orc_set_ss(self.h, len(p), p.ctypes.data_as(C.POINTER(C.c_int32))))

    def set_step(self, istep):
        lib().orc_set_step(self.h, int(istep))

    def md_coef(self, irep=1):
        c = np.zeros((self.nmp, 4))
        lib().orc_md_coef(self.h, irep, _d(c))
        return c

    def set_tape(self, irep, words):
        w = np.ascontiguousarray(words, np.uint32)
        self._tapes[irep] = w
        lib().orc_set_tape(self.h, irep, w.ctypes.data_as(C.POINTER(C.c_uint32)), w.size)

    def tape_pos(self, irep):
        return lib().orc_tape_pos(self.h, irep)

    def boxmuller(self, irep, n):
        o = np.zeros(n)
        lib().orc_boxmuller(self.h, irep, n, _d(o))
        return o

    def genrand(self, irep, kind, n):
        o = np.zeros(n)
        lib().orc_genrand(self.h, irep, kind, n, _d(o))
        return o

    def kinetic(self, irep=1):
        return float(lib().orc_kinetic(self.h, irep))

    def init_velocities(self):
        lib().orc_init_velocities(self.h)

    def md(self, nsteps, nstep_exchange=0):
        e = np.zeros((self.nrep, 11))
        nreb = lib().orc_md(self.h, nsteps, nstep_exchange, _d(e))
        return nreb, e

    def labels(self):
        lab = np.zeros(self.nrep, np.int32)
        lib().orc_labels(self.h, _i(lab))
        return lab


def philox4x32_10(ctr, key):
    """Philox4x32-10 block (the oracle's restatement) of a 4-word counter
    and a 2-word key."""
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    P = C.POINTER(C.c_uint32)
    lib().orc_philox(c.ctypes.data_as(P), k.ctypes.data_as(P), out.ctypes.data_as(P))
    return out
