"""sis_gpu_energy_noalloc: energy_sumup with its scratch in the library's
static device array, checked through the C ABI against sis_gpu_energy on
the same contexts (same kernels and summation order, so the bar is
equality), on two calls in a row where the second has the smaller input
(the array keeps the first call's contents), with external terms, and the
refusal of a context beyond the array's size."""
import numpy as np
import pytest

from sis_amd import configs
from sis_amd.engine import Engine, SisGpuError
from tests.golden.make_golden import load

pytestmark = pytest.mark.gpu


def _both(eng, r, T, valid):
    a = eng.energy_sumup(r, T, 0, valid)
    b = eng.energy_noalloc(r, T, valid)
    return a, b


def test_noalloc_energy_equals_allocating_entry_point():
    big = configs.cag_system(40, 47, 1.0, 310.0)   # 1,880 beads: 8 partial-sum blocks
    print("[case] noalloc energy: 1,880-bead condensate, then a 4 x 64 ladder", flush=True)
    e1 = Engine(big, mode="phases", seed=1)
    e1.neighbor_list(0)
    e1.force(1, 1)
    for valid in (1, 0):
        for T in (310.0, 290.0):
            a, b = _both(e1, 1, T, valid)
            assert np.array_equal(a, b), (valid, T, a, b)
    assert abs(a[0]) > 0.0
    small = configs.remd_c3(nrep=4, n=64, seed=5)   # second call: fewer beads, more replicas
    e2 = Engine(small, mode="phases", seed=2)
    e2.neighbor_list(0)
    for r in (1, 3, 4):
        e2.force(r, 1)
        for valid in (1, 0):
            a, b = _both(e2, r, -1.0, valid)
            assert np.array_equal(a, b), (r, valid)
            a, b = _both(e2, r, 350.0, valid)
            assert np.array_equal(a, b), (r, valid)
    # and the first context again after the smaller one
    a, b = _both(e1, 1, 305.0, 1)
    assert np.array_equal(a, b)


def test_noalloc_energy_with_external_terms():
    sysd, d = load("remd2ext")
    print("[case] noalloc energy: remd2ext external terms", flush=True)
    e = Engine(sysd, mode="phases", seed=3)
    e.set_external(sysd["external"])
    e.neighbor_list(0)
    for r in (1, 2):
        e.force(r, 1)
        a, b = _both(e, r, -1.0, 1)
        assert np.array_equal(a, b)
        assert np.any(a[7:11] != 0.0)


def test_noalloc_energy_refusals():
    s = configs.remd_c3(nrep=2, n=20, seed=1)
    print("[case] noalloc energy: refusals", flush=True)
    e = Engine(s, mode="phases", seed=1)
    with pytest.raises(SisGpuError):      # before any rebuild
        e.energy_noalloc(1, -1.0, 1)
    e.neighbor_list(0)
    from sis_amd.engine import lib
    import ctypes as C
    out = np.zeros(11)
    rc = lib().sis_gpu_energy_noalloc(e.h, 1, -1.0, 1, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc != 0   # the limit MC is refused
    many = configs.remd_c3(nrep=17000, n=4, seed=1)   # (8 + 16 + R) doubles + 2R ints > 256 KiB
    big = Engine(many, mode="phases", seed=1)
    with pytest.raises(SisGpuError, match="scratch"):
        big.energy_noalloc(1, -1.0, 1)
