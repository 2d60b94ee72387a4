"""Python access to the reference SIS Fortran built in oracle/_ref/ (test
oracle only: imported by tests/ and by bench.py's CPU-baseline leg).

The reference keeps its state in module globals and job_md is single-use
per process, so every call here runs in a fresh spawned process.
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp
import os
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_SO = HERE / "_ref" / "libsisref.so"
REF_OMP_SO = HERE / "_ref" / "libsisref_omp.so"
# the serial build with force / neighbor_list / energy_sumup replaced by the
# drop-in layer on sis_amd/libsis_gpu.so (oracle/Makefile target dropin)
REF_GPU_SO = HERE / "_ref" / "libsisref_gpu.so"


def available(omp: bool = False) -> bool:
    return (REF_OMP_SO if omp else REF_SO).exists()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _setup(lib, sysd, prefix, seed, nstep_exchange=0):
    import sys
    sys.path.insert(0, str(HERE.parent))
    from sis_amd.abi import params_to_struct
    p = params_to_struct(sysd["params"])
    nmp_chain = np.ascontiguousarray(sysd["nmp_chain"], np.int32)
    molt = np.ascontiguousarray(sysd["moltype"], np.int32)
    seq = np.ascontiguousarray(sysd["seq"], np.int32)
    temps = np.ascontiguousarray(sysd["temps"], np.float64)
    xyz = np.ascontiguousarray(sysd["xyz"], np.float64)
    grid = sysd.get("remd_grid")   # {"temps": [nT], "ion": [nI]}: a temperature x ionic-strength grid
    if grid is not None:
        tv = np.ascontiguousarray(grid["temps"], np.float64)
        iv = np.ascontiguousarray(grid["ion"], np.float64)
        lib.sisref_remd_temps(C.c_int(len(tv)), _dp(tv))
        lib.sisref_remd_var(C.c_int(3), C.c_int(len(iv)), _dp(iv))
    for key, kind in (("remd_ion", 3), ("remd_rg", 4), ("remd_dcf", 2)):   # ladders over REPT%ION / RG / TWZDCF
        if grid is not None:
            break
        if sysd.get(key) is not None:
            vals = np.ascontiguousarray(sysd[key], np.float64)
            lib.sisref_remd_var(C.c_int(kind), C.c_int(len(vals)), _dp(vals))
    lib.sisref_setup(C.byref(p), C.c_int(int(sysd["nchains"])), _ip(nmp_chain), _ip(molt),
                     _ip(seq), C.c_int(int(sysd["nrep"])), _dp(temps), C.c_int(int(seed)),
                     _dp(xyz), prefix.encode() + b"\0", C.c_int64(int(nstep_exchange)))


def _worker(q, so, sysd, ops, seed, nstep_exchange):
    try:
        # quiet the reference's stdout banner
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        lib = C.CDLL(str(so))
        tmp = tempfile.mkdtemp(prefix="sisref_")
        _setup(lib, sysd, os.path.join(tmp, "ref"), seed, nstep_exchange)
        nmp = int(np.sum(sysd["nmp_chain"]))
        nrep = int(sysd["nrep"])
        out = {}
        for op in ops:
            kind = op[0]
            if kind == "nl":
                lib.sisref_neighbor_list(C.c_int(op[1]))
            elif kind == "analysis_lists":
                lib.sisref_analysis_lists()
            elif kind == "lists":
                irep = op[1]
                cnt = np.zeros(3, np.int32)
                lib.sisref_counts(C.c_int(irep), _ip(cnt))
                wca = np.zeros((max(cnt[0], 1), 2), np.int32)
                ele = np.zeros((max(cnt[1], 1), 2), np.int32)
                bp = np.zeros((max(cnt[2], 1), 7), np.int32)
                bpc = np.zeros((max(cnt[2], 1), 2), np.float64)
                lib.sisref_lists(C.c_int(irep), _ip(wca), _ip(ele), _ip(bp), _dp(bpc))
                out[op[-1]] = dict(wca=wca[: cnt[0]].copy(), ele=ele[: cnt[1]].copy(),
                                   bp=bp[: cnt[2]].copy(), bpcoef=bpc[: cnt[2]].copy())
            elif kind == "force":
                f = np.zeros((nmp, 3), np.float64)
                lib.sisref_force(C.c_int(op[1]), C.c_int(op[2]), _dp(f))
                out[op[-1]] = f
            elif kind == "bp_state":
                cnt = np.zeros(3, np.int32)
                lib.sisref_counts(C.c_int(op[1]), _ip(cnt))
                st = np.zeros(max(cnt[2], 1), np.int8)
                en = np.zeros(max(cnt[2], 1), np.float64)
                lib.sisref_bp_state(C.c_int(op[1]), st.ctypes.data_as(C.POINTER(C.c_int8)), _dp(en))
                out[op[-1]] = (st[: cnt[2]].copy(), en[: cnt[2]].copy())
            elif kind == "energy":
                e = np.zeros(11, np.float64)
                lib.sisref_energy(C.c_int(op[1]), C.c_double(op[2]), C.c_int(op[3]),
                                  C.c_int(op[4]), _dp(e))
                out[op[-1]] = e
            elif kind == "set_velos":
                v = np.ascontiguousarray(op[1], np.float64)
                lib.sisref_set_velos(_dp(v))
            elif kind == "get_velos":
                vo = np.zeros((nrep, nmp, 3))
                ao = np.zeros((nrep, nmp, 3))
                lib.sisref_get_state(_dp(vo), _dp(ao))
                out[op[-1]] = vo
            elif kind == "kinetic":
                ek = C.c_double(0.0)
                lib.sisref_kinetic(C.c_int(op[1]), C.byref(ek))
                out[op[-1]] = ek.value
            elif kind == "ele":
                e = np.zeros(4, np.float64)
                lib.sisref_ele_params(C.c_int(op[1]), _dp(e))
                out[op[-1]] = e
            elif kind == "bp_tables":
                dh = np.zeros(468)
                ds = np.zeros(468)
                cut = np.zeros(7)
                lib.sisref_bp_tables(_dp(dh), _dp(ds), _dp(cut))
                out[op[-1]] = (dh, ds, cut)
            elif kind == "bp_maps":
                a = np.zeros((nmp, nmp), np.int32)
                b = np.zeros((nmp, nmp), np.int32)
                lib.sisref_bp_maps(_ip(a), _ip(b))
                out[op[-1]] = (a.T.copy(), b.T.copy())   # -> [imp-1, jmp-1]
            elif kind == "tape":
                n = int(op[2])
                w = np.zeros(n, np.uint32)
                lib.sisref_mt_tape(C.c_int(op[1]), C.c_int64(n),
                                   w.ctypes.data_as(C.POINTER(C.c_int32)))
                out[op[-1]] = w
            elif kind == "genrand":
                n = int(op[3])
                r = np.zeros(n)
                lib.sisref_genrand(C.c_int(op[1]), C.c_int(op[2]), C.c_int64(n), _dp(r))
                out[op[-1]] = r
            elif kind == "boxmuller":
                n = int(op[2])
                r = np.zeros(n)
                lib.sisref_boxmuller(C.c_int(op[1]), C.c_int64(n), _dp(r))
                out[op[-1]] = r
            elif kind == "set_xyz":
                x = np.ascontiguousarray(op[1], np.float64)
                lib.sisref_set_xyz(_dp(x))
            elif kind == "get_xyz":
                x = np.zeros((nrep, nmp, 3))
                lib.sisref_get_xyz(_dp(x))
                out[op[-1]] = x
            elif kind == "job_md_timed":
                import time
                e = np.zeros((nrep, 11))
                t0 = time.perf_counter()
                lib.sisref_job_md(C.c_int64(int(op[1])), _dp(e))
                out[op[-1]] = time.perf_counter() - t0
            elif kind == "external":
                # op[1]: external-term dict (sis_amd.abi.external_to_struct keys), op[2]: istep
                ext = op[1]
                i32a = lambda a, n: np.ascontiguousarray(a if a is not None else np.zeros(n), np.int32)  # noqa: E731
                f64a = lambda a, n: np.ascontiguousarray(a if a is not None else np.zeros(n), np.float64)  # noqa: E731
                dp_ = i32a(ext.get("dcf_pairs"), 2)
                nd = len(dp_) if ext.get("dcf_pairs") is not None else 0
                df = f64a(ext.get("dcf_forces"), 3 * max(nrep, 1))
                fp = i32a(ext.get("fr_pairs"), 2)
                nf = len(fp) if ext.get("fr_pairs") is not None else 0
                fk, fi, fv = f64a(ext.get("fr_k"), 2), f64a(ext.get("fr_init"), 6), f64a(ext.get("fr_velo"), 6)
                rg0 = f64a(np.broadcast_to(np.asarray(ext.get("rg_0", 0.0), float), (nrep,)), nrep)
                ri = i32a(ext.get("rest_id"), 3)
                nr = len(ri) if ext.get("rest_id") is not None else 0
                rc, rp = f64a(ext.get("rest_rcut"), 1), f64a(ext.get("rest_para"), 3)
                lib.sisref_set_external(
                    C.c_int(int(bool(ext.get("stage", 0)))), C.c_double(float(ext.get("stage_sigma", 10.0))),
                    C.c_double(float(ext.get("stage_eps", 1.2))), C.c_int(nd), _ip(dp_), _dp(df), C.c_int(nf),
                    _ip(fp), _dp(fk), _dp(fi), _dp(fv), C.c_int(int(ext.get("rg", 0))),
                    C.c_double(float(ext.get("rg_k", 0.0))), _dp(rg0), C.c_int(nr), _ip(ri), _dp(rc), _dp(rp),
                    C.c_int64(int(op[2])))
            elif kind == "twz":
                fn = str(op[1]).encode()
                lib.sisref_open_twz(C.c_char_p(fn), C.c_int(len(fn)), C.c_int64(int(op[2])))
            elif kind == "dcf_remd":
                dd = np.ascontiguousarray(op[1], np.float64)
                lib.sisref_set_dcf_remd(_dp(dd))
            elif kind == "rg_schedule":
                fn = os.path.join(tmp, "rg.sched")
                with open(fn, "w") as fh:
                    for st_, k_, r_ in zip(op[1], op[2], op[3]):
                        fh.write(f"{int(st_)} {float(k_)!r} {float(r_)!r}\n")
                lib.sisref_set_rg_schedule(C.c_char_p(fn.encode()), C.c_int(len(fn)))
            elif kind == "resize":
                ch = np.ascontiguousarray(op[2], np.float64)
                tg = np.ascontiguousarray(op[3] if op[3] is not None else np.zeros(3), np.float64)
                lib.sisref_set_box_resize(C.c_int64(int(op[1])), _dp(ch), C.c_int(int(op[3] is not None)), _dp(tg))
            elif kind == "box":
                b = np.zeros(3)
                lib.sisref_get_box(_dp(b))
                out[op[-1]] = b
            elif kind == "anneal":
                st = np.ascontiguousarray(op[1], np.int64)
                tk = np.ascontiguousarray(op[2], np.float64)
                lib.sisref_set_anneal(C.c_int(len(st)), st.ctypes.data_as(C.POINTER(C.c_int64)), _dp(tk))
            elif kind == "ss":
                # op[1]: (n, 2) 1-based bead pairs, written as a BPL file
                bpl = os.path.join(tmp, "ss.bpl")
                with open(bpl, "w") as fh:
                    for a, b in np.asarray(op[1], int).reshape(-1, 2):
                        fh.write(f"{a} {b}\n")
                lib.sisref_set_ss(C.c_char_p(bpl.encode()), C.c_int(len(bpl)))
            elif kind == "write_rst":
                pre = str(op[1]).encode()
                lib.sisref_write_rst(C.c_char_p(pre), C.c_int(len(pre)))
            elif kind == "job_md":
                e = np.zeros((nrep, 11))
                lib.sisref_job_md(C.c_int64(int(op[1])), _dp(e))
                x = np.zeros((nrep, nmp, 3))
                v = np.zeros((nrep, nmp, 3))
                a = np.zeros((nrep, nmp, 3))
                lib.sisref_get_xyz(_dp(x))
                lib.sisref_get_state(_dp(v), _dp(a))
                lab = np.zeros(max(nrep, 1), np.int32)
                lib.sisref_replica_labels(_ip(lab))
                out[op[-1]] = dict(xyz=x, velos=v, accels=a, energies=e, rep2lab=lab)
            else:
                raise ValueError(kind)
        q.put(("ok", out))
    except BaseException as exc:  # report to the parent
        import traceback
        q.put(("err", traceback.format_exc() + repr(exc)))


def run(sysd: dict, ops: list, seed: int = 12345, omp: bool = False, nstep_exchange: int = 0,
        timeout: float = 600.0, threads: int = 1, dropin: bool = False) -> dict:
    """Run `ops` against a fresh reference process set up with `sysd`.

    ops entries (last element is the result key where one is produced):
      ("nl", irep) ; ("analysis_lists",) ; ("lists", irep, key) ; ("force", irep, bp_mc, key)
      ("bp_state", irep, key) ; ("energy", irep, tempK, bp_mc, bp_energy_valid, key)
      ("ele", irep, key) ; ("bp_tables", key) ; ("bp_maps", key)
      ("tape", irep, n, key) ; ("genrand", irep, kind, n, key)
      ("boxmuller", irep, n, key) ; ("set_xyz", xyz) ; ("get_xyz", key)
      ("job_md", nstep, key)
    """
    so = REF_GPU_SO if dropin else (REF_OMP_SO if omp else REF_SO)
    if not so.exists():
        raise RuntimeError(f"{so} not built (make -C oracle ref dropin)")
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    env_old = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_STACKSIZE")}
    os.environ["OMP_NUM_THREADS"] = str(threads)
    # the reference keeps per-replica work arrays as stack automatics
    # (src/job_md.F90:40-44, src/force.F90:20): the child gets the largest
    # stack the hard limit allows, and large OpenMP thread stacks
    os.environ.setdefault("OMP_STACKSIZE", "1G")
    import resource
    stack_old = resource.getrlimit(resource.RLIMIT_STACK)
    try:
        resource.setrlimit(resource.RLIMIT_STACK, (stack_old[1], stack_old[1]))
    except (ValueError, OSError):
        pass
    try:
        pr = ctx.Process(target=_worker, args=(q, so, sysd, ops, seed, nstep_exchange))
        pr.start()
        resource.setrlimit(resource.RLIMIT_STACK, stack_old)
        import queue
        import time
        t0 = time.time()
        while True:
            try:
                status, payload = q.get(timeout=1.0)
                break
            except queue.Empty:
                if not pr.is_alive():
                    status, payload = "err", f"reference process exited with code {pr.exitcode}"
                    break
                if time.time() - t0 > timeout:
                    pr.kill()
                    status, payload = "err", "reference process timed out"
                    break
        pr.join(timeout=30)
    finally:
        try:
            resource.setrlimit(resource.RLIMIT_STACK, stack_old)
        except (ValueError, OSError):
            pass
        for k, v in env_old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    if status != "ok":
        raise RuntimeError("reference run failed:\n" + payload)
    return payload


def _dcd_worker(q, so, path, xyz, box, fix_com):
    try:
        lib = C.CDLL(str(so))
        x = np.ascontiguousarray(xyz, np.float64)
        b = np.ascontiguousarray(box if box is not None else np.zeros(3), np.float64)
        p = str(path).encode()
        lib.sisref_write_dcd(C.c_char_p(p), C.c_int(len(p)), C.c_int(x.shape[1]), C.c_int(x.shape[0]), _dp(x),
                             C.c_int(1 if box is not None else 0), _dp(b), C.c_int(int(fix_com)))
        q.put(("ok", None))
    except Exception as e:  # pragma: no cover - reported to the caller
        q.put(("err", repr(e)))


def write_dcd(path, xyz_frames, box=None, fix_com=0, timeout=60):
    """The reference's DCD writer (file_dcd, src/dcd.F90) on xyz_frames
    (nframes, nmp, 3), run in a child process."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    pr = ctx.Process(target=_dcd_worker, args=(q, str(REF_SO), str(path), np.asarray(xyz_frames), box, fix_com))
    pr.start()
    try:
        st, msg = q.get(timeout=timeout)
    finally:
        pr.join(timeout=10)
    if st != "ok":
        raise RuntimeError(f"reference DCD writer failed: {msg}")


def _out_worker(q, so, path, isteps, labels, temps, ekin, ene, replica, ele):
    try:
        lib = C.CDLL(str(so))
        p = str(path).encode()
        n = len(isteps)
        ist = np.ascontiguousarray(isteps, np.int64)
        lab = np.ascontiguousarray(labels, np.int32)
        tk = np.ascontiguousarray(temps, np.float64)
        ek = np.ascontiguousarray(ekin, np.float64)
        en = np.ascontiguousarray(ene, np.float64)   # (n, 11) = Fortran (0:10, n)
        lib.sisref_write_out(C.c_char_p(p), C.c_int(len(p)), C.c_int(n), ist.ctypes.data_as(C.POINTER(C.c_int64)),
                             _ip(lab), _dp(tk), _dp(ek), _dp(en), C.c_int(int(replica)), C.c_int(int(ele)))
        q.put(("ok", None))
    except Exception as e:  # pragma: no cover - reported to the caller
        q.put(("err", repr(e)))


def write_out(path, isteps, labels, temps, ekin, ene, replica=False, ele=True, timeout=60):
    """The reference's .out writer (write_out_header + write_out,
    src/write_out.F90) on the given lines, run in a child process."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    pr = ctx.Process(target=_out_worker, args=(q, str(REF_SO), str(path), np.asarray(isteps), np.asarray(labels),
                                               np.asarray(temps), np.asarray(ekin), np.asarray(ene), replica, ele))
    pr.start()
    try:
        st, msg = q.get(timeout=timeout)
    finally:
        pr.join(timeout=10)
    if st != "ok":
        raise RuntimeError(f"reference .out writer failed: {msg}")
