"""Host-side mirror of the reference's hot-path interface over the C ABI.

Method names follow the reference subroutines they replace:
  neighbor_list(irep)             src/neighbor_list.F90:1
  force(irep, bp_mc)              src/force.F90:1
  energy_sumup(irep, tempK, ...)  src/energy_sumup.F90:1
  init_velocities()               src/job_md.F90:157-176
  step(nsteps)                    src/job_md.F90:360-460 (loop body)
  exchange()                      src/energy_replica.F90 + src/replica_exchange.F90
Every call goes through libsis_gpu.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import SisOptions, SisTopo, params_to_struct

MODES = {"auto": 0, "phases": 1, "persistent": 2, "batched": 3}


class SisGpuError(RuntimeError):
    pass


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = abi.load_library()
    return _LIB


def _check(rc: int):
    if rc != 0:
        msg = lib().sis_gpu_last_error().decode(errors="replace")
        raise SisGpuError(f"sis_gpu error {rc}: {msg}")


class Engine:
    def __init__(self, sysd: dict, precision: int = 0, parity: bool = False, mode: str | int = "phases",
                 seed: int = 20241, exchange_every: int = 0, rebuild_every: int = 0, block: int = 0,
                 max_nbr: int = 0, grid: int = 0, host_exchange: bool = False, no_graph: bool = False):
        self.sysd = sysd
        self.nmp = int(np.sum(sysd["nmp_chain"]))
        self.nrep = int(sysd["nrep"])
        self._keep = dict(
            nmp_chain=np.ascontiguousarray(sysd["nmp_chain"], np.int32),
            moltype=np.ascontiguousarray(sysd["moltype"], np.int32),
            seq=np.ascontiguousarray(sysd["seq"], np.int32),
        )
        t = SisTopo()
        t.nchains = int(sysd["nchains"])
        t.nmp = self.nmp
        t.nmp_chain = abi.iptr(self._keep["nmp_chain"])
        t.moltype = abi.iptr(self._keep["moltype"])
        t.seq = abi.iptr(self._keep["seq"])
        t.no_charge = None
        t.n_no_charge = 0
        self.topo = t
        self.prm = params_to_struct(sysd["params"])
        o = SisOptions()
        o.nrep = self.nrep
        o.precision = precision
        o.parity = 1 if parity else 0
        o.mode = MODES.get(mode, mode) if isinstance(mode, str) else int(mode)
        o.block = block
        o.max_nbr = max_nbr
        o.rebuild_every = rebuild_every
        o.exchange_every = exchange_every
        o.seed = seed
        o.grid = grid
        o.flags = (1 if host_exchange else 0) | (2 if no_graph else 0)
        self.opt = o
        h = C.c_void_p()
        _check(lib().sis_gpu_create(C.byref(t), C.byref(self.prm), C.byref(o), C.byref(h)))
        self.h = h
        if self.nrep > 1 or np.any(np.asarray(sysd["temps"]) != sysd["params"]["tempK"]):
            self.set_rep_conditions(np.asarray(sysd["temps"], float), None)
        self.set_state(sysd["xyz"])

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().sis_gpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    def set_state(self, xyz=None, velos=None, accels=None):
        arrs = [None if a is None else np.ascontiguousarray(np.asarray(a, np.float64).reshape(self.nrep, self.nmp, 3))
                for a in (xyz, velos, accels)]
        _check(lib().sis_gpu_set_state(self.h, *[abi.dptr(a) if a is not None else None for a in arrs]))

    def get_state(self):
        x = np.zeros((self.nrep, self.nmp, 3))
        v = np.zeros_like(x)
        a = np.zeros_like(x)
        _check(lib().sis_gpu_get_state(self.h, abi.dptr(x), abi.dptr(v), abi.dptr(a)))
        return x, v, a

    def set_rep_conditions(self, temps=None, ionic=None):
        t = None if temps is None else np.ascontiguousarray(temps, np.float64)
        i = None if ionic is None else np.ascontiguousarray(ionic, np.float64)
        _check(lib().sis_gpu_set_rep_conditions(self.h, abi.dptr(t) if t is not None else None,
                                                abi.dptr(i) if i is not None else None))

    def set_rep_dims(self, nvalues):
        """Replicas on a grid (REPT order of the variables with ladders,
        first dimension slowest): the pair table of each dimension in turn."""
        v = np.ascontiguousarray(nvalues, np.int32)
        _check(lib().sis_gpu_set_rep_dims(self.h, len(v), abi.iptr(v)))

    # ---- hot path ----
    def neighbor_list(self, irep: int = 0):
        _check(lib().sis_gpu_rebuild(self.h, irep))

    def analysis_lists(self):
        """The lists of the DCD / CHECK_FORCE jobs (list_bp, list_exv,
        list_ele) for every replica; the next step rebuilds the MD lists."""
        _check(lib().sis_gpu_analysis_lists(self.h))

    def force(self, irep: int = 1, bp_mc: int = 1) -> np.ndarray:
        f = np.zeros((self.nmp, 3))
        _check(lib().sis_gpu_force(self.h, irep, bp_mc, abi.dptr(f)))
        return f

    def energy_sumup(self, irep: int = 1, tempK: float = -1.0, bp_mc: int = 1,
                     bp_energy_valid: int = 0) -> np.ndarray:
        e = np.zeros(11)
        flags = (1 if bp_mc else 0) | (2 if bp_energy_valid else 0)
        _check(lib().sis_gpu_energy(self.h, irep, float(tempK), flags, abi.dptr(e)))
        return e

    def energy_noalloc(self, irep: int = 1, tempK: float = -1.0, bp_energy_valid: int = 1) -> np.ndarray:
        """energy_sumup without device allocation (sis_gpu_energy_noalloc);
        the limit MC is not available there."""
        e = np.zeros(11)
        flags = 2 if bp_energy_valid else 0
        _check(lib().sis_gpu_energy_noalloc(self.h, irep, float(tempK), flags, abi.dptr(e)))
        return e

    def autotune(self, skins=(3.0, 5.0, 8.0), nsteps: int = 200) -> dict:
        """sis_gpu_autotune: times every launch shape x skin from the
        current state and keeps the fastest (state restored)."""
        sk = np.ascontiguousarray(skins, np.float64)
        rep = np.zeros(2 + 9 * len(sk))
        _check(lib().sis_gpu_autotune(self.h, abi.dptr(sk), len(sk), int(nsteps), abi.dptr(rep)))
        trials = [dict(shape=int(rep[2 + 3 * k]), skin=float(rep[3 + 3 * k]), ms_per_step=round(float(rep[4 + 3 * k]), 5))
                  for k in range(3 * len(sk)) if rep[3 + 3 * k] > 0]   # forms with one shape fill nskin trials
        return dict(shape=int(rep[0]), skin=float(rep[1]), trials=trials)

    def kinetic(self, irep: int = 0):
        """energy_kinetic of replica irep (1-based), or of all replicas (0)."""
        out = np.zeros(max(self.nrep, 1))
        _check(lib().sis_gpu_kinetic(self.h, int(irep), abi.dptr(out)))
        return out if irep == 0 else float(out[0])

    def pairs(self, irep: int = 1, kind: int = 0) -> np.ndarray:
        n = C.c_int64()
        _check(lib().sis_gpu_get_pair_count(self.h, irep, kind, C.byref(n)))
        w = 7 if kind == 2 else 2
        out = np.zeros((max(n.value, 1), w), np.int32)
        m = C.c_int64()
        _check(lib().sis_gpu_get_pairs(self.h, irep, kind, abi.iptr(out), n.value, C.byref(m)))
        return out[: m.value].copy()

    def bp(self, irep: int = 1) -> dict:
        n = C.c_int64()
        _check(lib().sis_gpu_get_pair_count(self.h, irep, 2, C.byref(n)))
        k = max(n.value, 1)
        mp = np.zeros((k, 7), np.int32)
        cf = np.zeros((k, 2))
        st = np.zeros(k, np.uint8)
        en = np.zeros(k)
        m = C.c_int64()
        _check(lib().sis_gpu_get_bp(self.h, irep, abi.iptr(mp), abi.dptr(cf), abi.u8ptr(st), abi.dptr(en),
                                    k, C.byref(m)))
        n = m.value
        return dict(bp=mp[:n].copy(), bpcoef=cf[:n].copy(), status=st[:n].copy(), ene=en[:n].copy())

    def lists(self, irep: int = 1) -> dict:
        b = self.bp(irep)
        return dict(wca=self.pairs(irep, 0), ele=self.pairs(irep, 1), bp=b["bp"], bpcoef=b["bpcoef"])

    # ---- random numbers ----
    def set_rng(self, seed: int):
        _check(lib().sis_gpu_set_rng(self.h, C.c_uint64(seed)))

    def set_rng_tape(self, irep: int, words):
        w = np.ascontiguousarray(words, np.uint32)
        _check(lib().sis_gpu_set_rng_tape(self.h, irep, abi.u32ptr(w), w.size))

    # ---- dynamics ----
    def init_velocities(self):
        _check(lib().sis_gpu_init_velocities(self.h))

    def step(self, nsteps: int) -> int:
        n = C.c_int64()
        _check(lib().sis_gpu_step(self.h, int(nsteps), C.byref(n)))
        return n.value

    def exchange(self):
        e = np.zeros((self.nrep, 2))
        acc = np.zeros(max(self.nrep // 2 + 1, 1), np.int32)
        _check(lib().sis_gpu_exchange(self.h, abi.dptr(e), abi.iptr(acc)))
        return e, acc

    # ---- trajectory output (DCD, one file per replica) ----
    def dcd_open(self, paths, fix_com_origin: int = 0, nstep_save: int = 0):
        arr = (C.c_char_p * self.nrep)(*[str(x).encode() for x in paths])
        _check(lib().sis_gpu_dcd_open(self.h, arr, int(fix_com_origin), int(nstep_save)))

    def dcd_write(self):
        _check(lib().sis_gpu_dcd_write(self.h))

    def dcd_close(self):
        _check(lib().sis_gpu_dcd_close(self.h))

    def bp_open(self, paths, kind: int = 0, nstep_save: int = 0):
        arr = (C.c_char_p * self.nrep)(*[str(x).encode() for x in paths])
        _check(lib().sis_gpu_bp_open(self.h, arr, int(kind), int(nstep_save)))

    def bp_write(self):
        _check(lib().sis_gpu_bp_write(self.h))

    def bp_close(self):
        _check(lib().sis_gpu_bp_close(self.h))

    def out_open(self, paths, nstep_save: int):
        arr = (C.c_char_p * self.nrep)(*[str(x).encode() for x in paths])
        _check(lib().sis_gpu_out_open(self.h, arr, int(nstep_save)))

    def out_write(self):
        """Energy lines of the current state now (the initial line)."""
        _check(lib().sis_gpu_out_write(self.h))

    def out_close(self):
        _check(lib().sis_gpu_out_close(self.h))

    def set_external(self, ext: dict | None):
        """Stage wall, tweezers, Rg bias and restraints (see
        abi.external_to_struct); None clears them."""
        if ext is None:
            _check(lib().sis_gpu_set_external(self.h, None))
            return
        s, keep = abi.external_to_struct(ext, self.nrep)
        _check(lib().sis_gpu_set_external(self.h, C.byref(s)))
        del keep

    def set_anneal(self, steps, temps):
        """Annealing schedule (read_anneal): temps[k] from step steps[k];
        empty turns it off."""
        st = np.ascontiguousarray(steps, np.int64)
        tk = np.ascontiguousarray(temps, np.float64)
        _check(lib().sis_gpu_set_anneal(self.h, len(st), st.ctypes.data_as(C.POINTER(C.c_int64)), abi.dptr(tk)))

    def twz_open(self, path, nstep_save: int):
        """Tweezers' force-ramp lines (.twz) every nstep_save steps."""
        _check(lib().sis_gpu_twz_open(self.h, str(path).encode(), int(nstep_save)))

    def twz_close(self):
        _check(lib().sis_gpu_twz_close(self.h))

    def set_rep_dcf(self, k, direction):
        """f-REMD: constant-force magnitude k per replica's label along the
        normalised direction of the one constant-force pair."""
        kk = np.ascontiguousarray(k, np.float64)
        dd = np.ascontiguousarray(direction, np.float64)
        _check(lib().sis_gpu_set_rep_dcf(self.h, abi.dptr(kk), abi.dptr(dd)))

    def set_rep_rg(self, rg0):
        """Replica exchange over the Rg-bias target: rg0 per replica's label."""
        if rg0 is None:
            _check(lib().sis_gpu_set_rep_rg(self.h, None))
            return
        g = np.ascontiguousarray(rg0, np.float64)
        _check(lib().sis_gpu_set_rep_rg(self.h, abi.dptr(g)))

    def set_rg_schedule(self, steps, k, r0):
        """Timed Rg bias: (k[i], r0[i]) from step steps[i] on."""
        st = np.ascontiguousarray(steps, np.int64)
        kk = np.ascontiguousarray(k, np.float64)
        rr = np.ascontiguousarray(r0, np.float64)
        _check(lib().sis_gpu_set_rg_schedule(self.h, len(st), st.ctypes.data_as(C.POINTER(C.c_int64)),
                                             abi.dptr(kk), abi.dptr(rr)))

    def set_box_resize(self, resize_step: int, change, target=None):
        """[PBC_box] resizing: every resize_step steps the box moves by
        change per axis, up to target (None: no target)."""
        ch = np.ascontiguousarray(change, np.float64)
        tg = np.ascontiguousarray(target, np.float64) if target is not None else None
        _check(lib().sis_gpu_set_box_resize(self.h, int(resize_step), abi.dptr(ch),
                                            abi.dptr(tg) if tg is not None else None))

    def set_ss(self, pairs):
        """Restrict base pairs to a CT/BPSEQ/BPL file's pairs, (n, 2)
        1-based bead ids; None lifts the restriction."""
        if pairs is None:
            _check(lib().sis_gpu_set_ss(self.h, -1, None))
            return
        p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        _check(lib().sis_gpu_set_ss(self.h, len(p), abi.iptr(p)))

    def rst_open(self, paths, nstep_save_rst: int = 0):
        """Restart files (src/write_rst.F90 blocks), one per replica."""
        arr = (C.c_char_p * self.nrep)(*[str(x).encode() for x in paths])
        _check(lib().sis_gpu_rst_open(self.h, arr, int(nstep_save_rst)))

    def rst_write(self):
        _check(lib().sis_gpu_rst_write(self.h))

    def rst_close(self):
        _check(lib().sis_gpu_rst_close(self.h))

    def rst_read(self, paths):
        """read_rst for a restarted run: labels, state and step count."""
        arr = (C.c_char_p * self.nrep)(*[str(x).encode() for x in paths])
        _check(lib().sis_gpu_rst_read(self.h, arr))

    def dcd_energies(self, path, out_path=None, max_frames: int = 100000):
        """job_dcd: energies (nframes, 11) of the frames of a DCD file."""
        e = np.zeros((max_frames, 11))
        n = C.c_int64()
        _check(lib().sis_gpu_dcd_energies(self.h, str(path).encode(),
                                          None if out_path is None else str(out_path).encode(), int(max_frames),
                                          abi.dptr(e), C.byref(n)))
        return e[: min(n.value, max_frames)].copy()

    def check_force(self, irep: int = 1, h: float = 1e-3, out_path=None):
        """job_check_force: (f_energy, f_force), each (nmp, 3)."""
        fe = np.zeros((self.nmp, 3))
        ff = np.zeros((self.nmp, 3))
        _check(lib().sis_gpu_check_force(self.h, int(irep), float(h),
                                         None if out_path is None else str(out_path).encode(), abi.dptr(fe),
                                         abi.dptr(ff)))
        return fe, ff

    # ---- restart ----
    def set_step(self, istep: int):
        _check(lib().sis_gpu_set_step(self.h, int(istep)))

    def set_labels(self, rep2lab):
        lab = np.ascontiguousarray(rep2lab, np.int32)
        _check(lib().sis_gpu_set_labels(self.h, abi.iptr(lab)))

    def labels(self) -> np.ndarray:
        lab = np.zeros(self.nrep, np.int32)
        _check(lib().sis_gpu_get_labels(self.h, abi.iptr(lab)))
        return lab

    def info(self) -> dict:
        v = np.zeros(11, np.int64)
        _check(lib().sis_gpu_get_info(self.h, abi.i64ptr(v), 11))
        keys = ["nmp", "nrep", "istep", "rebuilds", "nwca", "nele", "nbp", "max_nbr", "mode", "block",
                "device_exchanges"]
        return dict(zip(keys, [int(x) for x in v]))

    def last_ms(self, which: int = 0) -> float:
        return float(lib().sis_gpu_last_kernel_ms(self.h, which))


def write_dcd_file(path, xyz_frames, box=None, fix_com_origin: int = 0):
    """Host-side DCD writer of the library (the format of src/dcd.F90):
    xyz_frames (nframes, nmp, 3); box (3,) adds the unit-cell blocks."""
    x = np.ascontiguousarray(np.asarray(xyz_frames, np.float64))
    nfr, nmp = x.shape[0], x.shape[1]
    b = np.ascontiguousarray(np.asarray(box if box is not None else [0.0, 0.0, 0.0], np.float64))
    _check(lib().sis_dcd_write_file(str(path).encode(), nmp, nfr, abi.dptr(x), 1 if box is not None else 0,
                                    abi.dptr(b), int(fix_com_origin)))


def write_bp_file(path, sets, tempK: float, kind: int = 0):
    """Host-side base-pair writer of the library (src/write_bp.F90 layout):
    sets is a list of (ij (n, 2) 1-based, status (n,), ene (n,)) of equal n."""
    n = len(sets[0][1]) if sets else 0
    ij = np.ascontiguousarray(np.concatenate([np.asarray(s[0], np.int32).reshape(n, 2) for s in sets]) if sets
                              else np.zeros((0, 2), np.int32))
    st = np.ascontiguousarray(np.concatenate([np.asarray(s[1], np.uint8) for s in sets]) if sets
                              else np.zeros(0, np.uint8))
    en = np.ascontiguousarray(np.concatenate([np.asarray(s[2], np.float64) for s in sets]) if sets
                              else np.zeros(0))
    _check(lib().sis_bp_write_file(str(path).encode(), n, abi.iptr(ij), abi.u8ptr(st), abi.dptr(en), len(sets),
                                   float(tempK), int(kind)))


def write_out_file(path, isteps, labels, temps, ekin, energies, replica: bool, ele: bool):
    """Host-side .out line writer of the library (src/write_out.F90 format);
    energies (nline, 11)."""
    ist = np.ascontiguousarray(isteps, np.int64)
    lab = np.ascontiguousarray(labels, np.int32)
    tk = np.ascontiguousarray(temps, np.float64)
    ek = np.ascontiguousarray(ekin, np.float64)
    en = np.ascontiguousarray(energies, np.float64)
    _check(lib().sis_out_write_file(str(path).encode(), len(ist), ist.ctypes.data_as(C.POINTER(C.c_int64)),
                                    abi.iptr(lab), abi.dptr(tk), abi.dptr(ek), abi.dptr(en), int(bool(replica)),
                                    int(bool(ele))))


def write_rst_file(path, grep, nrep, istep, xyz, velos, accels, rep2lab=None, mts_rep=None, mts=None,
                   box=None, seed=None):
    """Host-side restart writer of the library for one replica; xyz,
    velos, accels (nmp, 3); mts_rep / mts 643 uint32 words or None."""
    x = np.ascontiguousarray(xyz, np.float64)
    v = np.ascontiguousarray(velos, np.float64)
    a = np.ascontiguousarray(accels, np.float64)
    lab = np.ascontiguousarray(rep2lab if rep2lab is not None else np.zeros(1), np.int32)
    u32 = C.POINTER(C.c_uint32)
    mr = np.ascontiguousarray(mts_rep, np.uint32) if mts_rep is not None else None
    ms = np.ascontiguousarray(mts, np.uint32) if mts is not None else None
    bx = np.ascontiguousarray(box, np.float64) if box is not None else None
    sd = C.c_uint64(int(seed)) if seed is not None else None
    _check(lib().sis_rst_write_file(
        str(path).encode(), int(grep), int(nrep), x.shape[0], int(istep), abi.iptr(lab), abi.dptr(x), abi.dptr(v),
        abi.dptr(a), mr.ctypes.data_as(u32) if mr is not None else None,
        ms.ctypes.data_as(u32) if ms is not None else None, abi.dptr(bx) if bx is not None else None,
        C.byref(sd) if sd is not None else None))


def read_rst_file(path, nmp: int, nrep: int):
    """Host-side restart reader of the library: dict of the blocks found."""
    blocks, grep = C.c_int32(), C.c_int32()
    istep = C.c_int64()
    lab = np.zeros(nrep, np.int32)
    x, v, a = np.zeros((nmp, 3)), np.zeros((nmp, 3)), np.zeros((nmp, 3))
    box = np.zeros(3)
    seed = C.c_uint64()
    _check(lib().sis_rst_read_file(str(path).encode(), int(nmp), int(nrep), C.byref(blocks), C.byref(grep),
                                   C.byref(istep), abi.iptr(lab), abi.dptr(x), abi.dptr(v), abi.dptr(a),
                                   abi.dptr(box), C.byref(seed)))
    ids = [k + 1 for k in range(31) if blocks.value >> k & 1]
    return dict(blocks=ids, grep=grep.value, istep=istep.value, rep2lab=lab, xyz=x, velos=v, accels=a, box=box,
                seed=seed.value)


def read_dcd_file(path, nmp: int, max_frames: int = 100000):
    """Host-side DCD reader of the library: (frames (n, nmp, 3), box (3,) or
    -1 when the file has no unit cell)."""
    x = np.zeros((max_frames, nmp, 3))
    b = np.zeros(3)
    n = C.c_int64()
    _check(lib().sis_dcd_read_file(str(path).encode(), int(nmp), int(max_frames), abi.dptr(x), abi.dptr(b),
                                   C.byref(n)))
    return x[: min(n.value, max_frames)].copy(), b
