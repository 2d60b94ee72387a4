"""Host-side checks that need no GPU: the C ABI library loads and exports
every entry point include/sis_gpu.h declares; the Fortran binding module
compiles; the synthetic configurations have the sizes the benchmark
configurations state."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from sis_amd import abi, configs

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    txt = (ROOT / "include" / "sis_gpu.h").read_text()
    return sorted(set(re.findall(r"\b(sis_(?:gpu|pool|dcd|bp|out|rst|twz)_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(abi.LIB_PATH))
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(abi.EXPORTED) == names


def test_struct_layouts():
    assert C.sizeof(abi.SisBpPara) == 16 * 8
    assert C.sizeof(abi.SisParams) == (18 + 42 + 12 + 48 + 4 + 3) * 8 + 14 * 4
    assert C.sizeof(abi.SisOptions) == 8 * 4 + 8 + 2 * 4


@pytest.mark.skipif(not Path("/opt/rocm/llvm/bin/flang").exists(), reason="flang absent")
def test_fortran_binding_compiles(tmp_path):
    src = ROOT / "sis_amd" / "fortran" / "sis_gpu_iface.F90"
    r = subprocess.run(["/opt/rocm/llvm/bin/flang", "-c", "-module-dir", str(tmp_path), str(src),
                        "-o", str(tmp_path / "x.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_config_sizes():
    c1 = configs.hairpin64(1)
    assert c1["xyz"].shape == (1, 64, 3)
    s = c1["seqs"][0]
    stem = s[5:29]
    assert configs.revcomp(stem) == s[37:61]
    c3 = configs.remd_c3(nrep=256, n=128)
    assert c3["nrep"] == 256 and len(c3["seq"]) == 128
    assert c3["temps"][0] == pytest.approx(290.0) and c3["temps"][-1] == pytest.approx(380.0)
    assert configs.box_for_concentration(9400, 1.0) == pytest.approx(2499.1, abs=0.1)
    assert configs.box_for_concentration(94000, 2.0) == pytest.approx(4273.5, abs=0.1)
    bonds = np.linalg.norm(np.diff(c1["xyz"][0], axis=0), axis=1)
    assert np.allclose(bonds, 5.9)
