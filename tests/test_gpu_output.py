"""Device-fed trajectory output: frames written during sis_gpu_step (pinned
staging buffers, background writer) are byte-identical to the host writer
applied to state snapshots of the same trajectory taken between shorter
sis_gpu_step calls."""
import numpy as np
import pytest

from sis_amd import configs
from sis_amd.engine import Engine, write_dcd_file

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["batched", "pbc_phases"])
def test_dcd_frames_during_steps(tmp_path, case):
    if case == "batched":
        s = configs.remd_c3(nrep=4, n=128, seed=3)
        kw = dict(mode="batched", precision=1, exchange_every=50)
        fix_com, box = 0, None
    else:
        s = configs.small_pbc()
        kw = dict(mode="phases", precision=0)
        fix_com, box = 1, s["params"]["box"]
    R = int(s["nrep"])
    print(f"[case] dcd {case}", flush=True)
    eng = Engine(s, seed=41, **kw)
    eng.init_velocities()
    eng.neighbor_list(0)
    paths = [tmp_path / f"dev_{r}.dcd" for r in range(R)]
    eng.dcd_open(paths, fix_com_origin=fix_com, nstep_save=25)
    eng.dcd_write()
    eng.step(60)
    eng.step(40)
    eng.dcd_close()
    ref = Engine(s, seed=41, **kw)
    ref.init_velocities()
    ref.neighbor_list(0)
    frames = [ref.get_state()[0]]
    for _ in range(4):
        ref.step(25)
        frames.append(ref.get_state()[0])
    frames = np.array(frames)   # (5, R, N, 3)
    for r in range(R):
        want = tmp_path / f"host_{r}.dcd"
        write_dcd_file(want, frames[:, r], box=box, fix_com_origin=fix_com)
        assert paths[r].read_bytes() == want.read_bytes(), r
