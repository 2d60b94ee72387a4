"""Job pool over the C ABI: the serving form (sis_pool_* in sis_gpu.h).

Jobs are whole SIS MD runs (one replica each): their own chains, sequence,
temperature, salt, box, step count and seed, with the force field shared by
the pool.  `submit` returns at once; a native scheduler thread admits jobs
by HBM footprint, keeps them resident and advances all of them in shared
launches.  `poll` reports state, timing and, when done, final coordinates.
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import SisTopo, params_to_struct
from .engine import SisGpuError, lib

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


class JobPool:
    def __init__(self, hbm_budget_bytes: int = 0, quantum_bead_steps: int = 0, max_active: int = 1024):
        h = C.c_void_p()
        self._check(lib().sis_pool_create(int(hbm_budget_bytes), int(quantum_bead_steps), int(max_active),
                                          C.byref(h)))
        self.h = h
        self._nmp = {}

    @staticmethod
    def _check(rc: int):
        if rc != 0:
            msg = lib().sis_gpu_last_error().decode(errors="replace")
            raise SisGpuError(f"sis_pool error {rc}: {msg}")

    def submit(self, sysd: dict, nsteps: int, seed: int, precision: int = 0) -> int:
        """sysd as built by sis_amd.configs (one replica); returns the job id."""
        keep = dict(nmp_chain=np.ascontiguousarray(sysd["nmp_chain"], np.int32),
                    moltype=np.ascontiguousarray(sysd["moltype"], np.int32),
                    seq=np.ascontiguousarray(sysd["seq"], np.int32))
        nmp = int(keep["seq"].size)
        t = SisTopo()
        t.nchains = int(sysd["nchains"])
        t.nmp = nmp
        t.nmp_chain = abi.iptr(keep["nmp_chain"])
        t.moltype = abi.iptr(keep["moltype"])
        t.seq = abi.iptr(keep["seq"])
        t.no_charge = None
        t.n_no_charge = 0
        prm = params_to_struct(sysd["params"])
        x = np.ascontiguousarray(np.asarray(sysd["xyz"], np.float64).reshape(-1, nmp, 3)[0])
        jid = C.c_int64()
        self._check(lib().sis_pool_submit(self.h, C.byref(t), C.byref(prm), abi.dptr(x), int(precision),
                                          int(nsteps), int(seed), C.byref(jid)))
        self._nmp[jid.value] = nmp
        return jid.value

    def poll(self, job_id: int, want_xyz: bool = False):
        """(state, times[5] = submit, start, end, steps done, rebuilds, xyz or None).
        A failed job raises SisGpuError with its message."""
        st = C.c_int32()
        tm = np.zeros(5)
        x = np.zeros((self._nmp[job_id], 3)) if want_xyz else None
        self._check(lib().sis_pool_poll(self.h, int(job_id), C.byref(st), abi.dptr(tm),
                                        abi.dptr(x) if x is not None else None))
        return st.value, tm, (x if st.value == DONE else None)

    def forget(self, job_id: int):
        self._check(lib().sis_pool_forget(self.h, int(job_id)))
        self._nmp.pop(job_id, None)

    def stats(self) -> dict:
        o = np.zeros(9)
        self._check(lib().sis_pool_stats(self.h, abi.dptr(o), 9))
        keys = ("queued", "running", "done", "failed", "bead_steps", "gpu_s", "launches", "bytes", "elapsed")
        return dict(zip(keys, o.tolist()))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().sis_pool_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
