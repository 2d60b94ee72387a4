"""The reference's own job_md driving the device.

oracle/_ref/libsisref_gpu.so is the reference SIS Fortran built from its
own sources with force.F90, neighbor_list.F90 and energy_sumup.F90 left out
and the project's drop-in layer (sis_amd/fortran/sis_gpu_dropin.F90) linked
in their place, on sis_amd/libsis_gpu.so (oracle/Makefile, target dropin).
Its unmodified job_md (src/job_md.F90: skin test, GJF-2GJ integrator,
temperature replica exchange, output) then runs on the device's lists,
forces, energies and limit MC.  Bars: against the reference's CPU run of
the same system and seed (the golden fixtures), the same labels, positions
within 1e-7 A, velocities within 1e-6 A/tau and final energies within 1e-7
relative, the bars of the engine's own trajectory parity tests."""
import numpy as np
import pytest

from oracle import refdrv
from tests.golden.make_golden import load

pytestmark = pytest.mark.gpu

CASES = ["hairpin64", "small_pbc", "mixed", "remd4", "t2s2hp_temp8", "t2s2hp_sisff", "cag47_m1"]


@pytest.mark.skipif(not refdrv.REF_GPU_SO.exists(), reason="oracle/_ref/libsisref_gpu.so not built")
@pytest.mark.parametrize("name", CASES)
def test_reference_job_md_drives_the_device(name):
    sysd, d = load(name)
    print(f"[case] reference job_md on the device: {name}", flush=True)
    out = refdrv.run(sysd, [("job_md", int(d["nstep"]), "R")], seed=4321,
                     nstep_exchange=int(d["nstep_exchange"]), dropin=True, timeout=300)["R"]
    assert np.array_equal(out["rep2lab"], d["md_rep2lab"])
    assert np.abs(out["xyz"] - d["md_xyz"]).max() < 1e-7
    assert np.abs(out["velos"] - d["md_velos"]).max() < 1e-6
    assert np.allclose(out["energies"], d["md_energies"], rtol=1e-7, atol=1e-7)
