"""Trajectory output (DCD) of the library against the reference's own
writer (file_dcd of src/dcd.F90, compiled from the reference sources into
oracle/_ref): byte-identical files on the same frames.  CPU only: the host
writer needs no GPU."""
import numpy as np
import pytest

from oracle import refdrv
from sis_amd.engine import write_dcd_file

needs_ref = pytest.mark.skipif(not refdrv.available(), reason="reference build (oracle/_ref) absent")


def _frames(seed, nframes, nmp, scale=300.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, size=(nframes, nmp, 3))


@needs_ref
@pytest.mark.parametrize("box,fix_com", [(None, 0), (None, 1), ((250.0, 260.5, 1.0e3 / 3.0), 0),
                                         ((2499.1, 2499.1, 2499.1), 1)])
def test_dcd_bytes_match_reference_writer(tmp_path, box, fix_com):
    x = _frames(7, 5, 129)
    ours, ref = tmp_path / "ours.dcd", tmp_path / "ref.dcd"
    write_dcd_file(ours, x, box=box, fix_com_origin=fix_com)
    refdrv.write_dcd(ref, x, box=None if box is None else np.array(box), fix_com=fix_com)
    a, b = ours.read_bytes(), ref.read_bytes()
    assert len(a) == len(b)
    assert a == b


def test_dcd_layout_known_answer(tmp_path):
    """Header and one frame decoded field by field (the layout of
    src/dcd.F90:287-355,538-596), without the reference build."""
    x = np.array([[[1.0, 2.0, 3.0], [4.0, 5.0, 6.5]]])
    p = tmp_path / "a.dcd"
    write_dcd_file(p, x, box=(10.0, 20.0, 30.0), fix_com_origin=1)
    raw = p.read_bytes()
    i32 = lambda o: int(np.frombuffer(raw, np.int32, 1, o)[0])  # noqa: E731
    assert i32(0) == 84 and raw[4:8] == b"CORD"
    ctrl = np.frombuffer(raw, np.int32, 20, 8)
    assert ctrl[10] == 1 and ctrl[19] == 24 and ctrl[:10].sum() == 0
    assert i32(88) == 84 and i32(92) == 84 and i32(96) == 1 and raw[100:180] == b" " * 80
    assert i32(180) == 84 and i32(184) == 4 and i32(188) == 2 and i32(192) == 4
    o = 196
    assert i32(o) == 48
    assert np.array_equal(np.frombuffer(raw, np.float64, 6, o + 4), [10.0, 0.0, 20.0, 0.0, 0.0, 30.0])
    o += 56
    com = x[0].mean(axis=0)
    for k in range(3):
        assert i32(o) == 8
        assert np.array_equal(np.frombuffer(raw, np.float32, 2, o + 4), (x[0, :, k] - com[k]).astype(np.float32))
        assert i32(o + 12) == 8
        o += 16
    assert o == len(raw)


def _read_bp(path):
    raw = open(path, "rb").read()
    assert np.frombuffer(raw, np.int32, 2, 0).tolist() == [2, 4]
    rec = np.frombuffer(raw[8:], dtype=[("i", "<i2"), ("j", "<i2"), ("e", "<f4")])
    sets, cur = [], []
    for r in rec:
        if r["i"] == 0 and r["j"] == 0 and r["e"] == 0.0:
            sets.append(cur)
            cur = []
        else:
            cur.append((int(r["i"]), int(r["j"]), float(r["e"])))
    assert not cur
    return sets


def test_bp_layout_known_answer(tmp_path):
    """Records of src/write_bp.F90: formed pairs in list order, E <= -kT for
    .bp, every formed pair for .bpall, float32 energies, a zero terminator
    per save step."""
    from sis_amd.engine import write_bp_file
    kT = 300.0 * 1.380649e-23 * ((1.0 / 4184.0) * 6.02214076e23)
    ij = np.array([[3, 40], [5, 38], [7, 36], [9, 34]])
    st = np.array([1, 1, 0, 1], np.uint8)
    en = np.array([-2.0 * kT, -0.5 * kT, -3.0, -kT])
    p0, p1 = tmp_path / "a.bp", tmp_path / "a.bpall"
    write_bp_file(p0, [(ij, st, en), (ij, 1 - st, en)], 300.0, kind=0)
    write_bp_file(p1, [(ij, st, en)], 300.0, kind=1)
    s0 = _read_bp(p0)
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    assert s0 == [[(3, 40, f32(en[0])), (9, 34, f32(en[3]))], [(7, 36, f32(-3.0))]]
    assert _read_bp(p1) == [[(3, 40, f32(en[0])), (5, 38, f32(en[1])), (9, 34, f32(en[3]))]]


def _out_cases():
    edge = [0.0, -0.0, 1e-8, -1e-8, 0.0999999, 0.09999996, 0.1, 0.123456789, 1.0, 0.9999996, 9.9999996,
            99999.94, 999999.4, 999999.6, 1e6, -1e6, 123456.789, -0.001234, 3.14159265, -42.0, 1e-101,
            1.23e120, -7.5e-5, 0.05, 0.5, 5.0, 50.0, 500.0, 5000.0, 50000.0, 500000.0, 5e6, 12.3456749999]
    rng = np.random.default_rng(5)
    rand = list(rng.normal(size=40) * 10.0 ** rng.integers(-9, 9, size=40))
    return np.array(edge + rand)


@needs_ref
@pytest.mark.parametrize("replica,ele", [(True, True), (False, True), (False, False)])
def test_out_lines_match_reference_writer(tmp_path, replica, ele):
    """Energy lines (i12, i4, f6.2, g13.6 fields) byte-identical to the
    reference's write_out_header + write_out on edge values (zero, the
    F/E switch points of G editing, three-digit exponents) and random
    magnitudes from 1e-9 to 1e9."""
    from sis_amd.engine import write_out_file
    v = _out_cases()
    n = len(v)
    rng = np.random.default_rng(9)
    ene = v[:, None] * rng.uniform(-2.0, 2.0, size=(n, 11))
    ene[:, 0] = v
    ekin = np.abs(v[::-1]) * 1.5
    temps = rng.uniform(280.0, 380.0, size=n)
    isteps = np.arange(n, dtype=np.int64) * 1000
    labels = np.arange(1, n + 1) % 256 + 1
    ours, ref = tmp_path / "ours.out", tmp_path / "ref.out"
    write_out_file(ours, isteps, labels, temps, ekin, ene, replica, ele)
    refdrv.write_out(ref, isteps, labels, temps, ekin, ene, replica=replica, ele=ele)
    a, b = ours.read_text().splitlines(), ref.read_text().splitlines()
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x == y


@needs_ref
@pytest.mark.parametrize("box", [None, (300.0, 310.0, 320.0)])
def test_dcd_reader_round_trip_of_reference_file(tmp_path, box):
    """Frames written by the reference's writer read back by the library's
    reader (read_onestep of src/dcd.F90): float32 values and the unit cell."""
    from sis_amd.engine import read_dcd_file
    x = _frames(3, 4, 77)
    p = tmp_path / "ref.dcd"
    refdrv.write_dcd(p, x, box=None if box is None else np.array(box), fix_com=0)
    got, b = read_dcd_file(p, 77)
    assert got.shape == x.shape
    assert np.array_equal(got, x.astype(np.float32).astype(np.float64))
    assert np.array_equal(b, [-1.0, -1.0, -1.0] if box is None else box)


def _rst_blocks(raw):
    """(id, size, offset of the fields) of each block, walking by size as
    read_rst's skip does (src/read_rst.F90:48-53)."""
    out, p = [], 0
    while p < len(raw):
        i, n = np.frombuffer(raw, np.int32, 2, p)
        out.append((int(i), int(n), p + 8))
        p += 8 + int(n)
    assert p == len(raw)
    return out


@needs_ref
@pytest.mark.parametrize("case", ["remd3", "pbc"])
def test_rst_bytes_match_reference_writer(tmp_path, case):
    """Restart files of write_rst (src/write_rst.F90:25-145) after 30 steps
    of the reference's job_md: the library's writer, given the same state,
    labels, step and the reference's MT states (opaque 643-word blocks),
    writes the same bytes; the library's reader returns that state."""
    from sis_amd import configs
    from sis_amd.engine import read_rst_file, write_rst_file
    s = configs.remd_c3(nrep=3, n=20, seed=2) if case == "remd3" else configs.small_pbc(3, 10, 80.0, seed=3)
    pre = str(tmp_path / "ref")
    r = refdrv.run(s, [("job_md", 30, "R"), ("write_rst", pre)], seed=5, nstep_exchange=10)["R"]
    nrep, nmp = int(s["nrep"]), int(np.sum(s["nmp_chain"]))
    for g in range(1, nrep + 1):
        raw = (tmp_path / f"ref{g:04d}.rst").read_bytes()
        blocks = {i: (n, o) for i, n, o in _rst_blocks(raw)}
        mts_rep = np.frombuffer(raw, np.uint32, 643, blocks[8][1]) if 8 in blocks else None
        mts = np.frombuffer(raw, np.uint32, 643, blocks[7][1] + 4)
        box = np.array(s["params"]["box"]) if s["params"]["pbc"] else None
        ours = tmp_path / f"ours{g}.rst"
        write_rst_file(ours, g, nrep, 30, r["xyz"][g - 1], r["velos"][g - 1], r["accels"][g - 1],
                       rep2lab=r["rep2lab"][:nrep], mts_rep=mts_rep, mts=mts, box=box)
        assert ours.read_bytes() == raw
        back = read_rst_file(tmp_path / f"ref{g:04d}.rst", nmp, nrep)
        assert back["istep"] == 30 and back["grep"] == g
        assert np.array_equal(back["xyz"], r["xyz"][g - 1])
        assert np.array_equal(back["velos"], r["velos"][g - 1])
        assert np.array_equal(back["accels"], r["accels"][g - 1])
        if g == 1 and nrep > 1:
            assert np.array_equal(back["rep2lab"], r["rep2lab"][:nrep])
        if box is not None and g == 1:
            assert np.array_equal(back["box"], box)


def test_rst_layout_known_answer(tmp_path):
    """Block ids, sizes and fields of a restart file decoded by hand (the
    layout of src/write_rst.F90:40-145): REPLICA, STEP, XYZ, VELO, ACCEL,
    PBC and the counter-RNG seed block 11; the reader returns them."""
    from sis_amd.engine import read_rst_file, write_rst_file
    x = np.arange(6, dtype=float).reshape(2, 3) + 0.5
    v, a = -x, 2 * x
    p = tmp_path / "k.rst"
    write_rst_file(p, 1, 2, 123456789012, x, v, a, rep2lab=[2, 1], box=[10.0, 20.0, 30.0], seed=2**63 + 5)
    raw = p.read_bytes()
    bl = _rst_blocks(raw)
    assert [(i, n) for i, n, _ in bl] == [(6, 20), (1, 12), (3, 56), (4, 56), (5, 56), (10, 24), (11, 8)]
    i32 = lambda o, k=1: np.frombuffer(raw, np.int32, k, o)  # noqa: E731
    assert list(i32(bl[0][2], 5)) == [2, 1, 2, 2, 1]                       # nrep, (i, rep2lab(i))
    assert i32(bl[1][2])[0] == 1 and np.frombuffer(raw, np.int64, 1, bl[1][2] + 4)[0] == 123456789012
    for (i, n, o), arr in zip(bl[2:5], (x, v, a)):
        assert list(i32(o, 2)) == [1, 2]                                   # grep, nmp
        assert np.array_equal(np.frombuffer(raw, np.float64, 6, o + 8), arr.ravel())
    assert list(np.frombuffer(raw, np.float64, 3, bl[5][2])) == [10.0, 20.0, 30.0]
    assert np.frombuffer(raw, np.uint64, 1, bl[6][2])[0] == 2**63 + 5
    back = read_rst_file(p, 2, 2)
    assert back["blocks"] == [1, 3, 4, 5, 6, 10, 11] and back["istep"] == 123456789012
    assert list(back["rep2lab"]) == [2, 1] and back["seed"] == 2**63 + 5
    assert np.array_equal(back["velos"], v)
    # replica 2's file carries no REPLICA and no PBC block
    q = tmp_path / "k2.rst"
    write_rst_file(q, 2, 2, 7, x, v, a, box=[10.0, 20.0, 30.0])
    assert [i for i, _, _ in _rst_blocks(q.read_bytes())] == [1, 3, 4, 5]
