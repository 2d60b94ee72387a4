"""Seeded synthetic systems for the five benchmark configurations.

A system is a dict:
  nchains, nmp_chain (int32[nchains]), moltype (int32[nchains]),
  seq (int32[nmp], SEQT codes chain after chain), xyz (float64[nrep, nmp, 3]),
  params (flat dict for struct sis_params), nrep, temps (float64[nrep]),
  name, nstep.
Chains are numbered contiguously (bead ids of chain k follow those of
chain k-1), which is the numbering the reference's FASTA reader produces
(src/read_fasta.F90:99-140).
"""
from __future__ import annotations

import math

import numpy as np

from .abi import SEQT, MOLT_RNA
from .forcefield import default_params

N_AVOGADRO = 6.02214076e23
COMP = {"A": "U", "U": "A", "G": "C", "C": "G"}


def box_for_concentration(nmp: int, conc_mM: float) -> float:
    """Cubic box edge (Angstrom) holding nmp nucleotides at conc_mM."""
    per_a3 = conc_mM * 1e-3 * N_AVOGADRO / 1e27
    return (nmp / per_a3) ** (1.0 / 3.0)


def random_seq(rng, n: int, gc: float) -> str:
    out = []
    for _ in range(n):
        if rng.random() < gc:
            out.append("G" if rng.random() < 0.5 else "C")
        else:
            out.append("A" if rng.random() < 0.5 else "U")
    return "".join(out)


def revcomp(s: str) -> str:
    return "".join(COMP[c] for c in reversed(s))


def _place(p0, p1, p2, bond, theta, phi):
    """NeRF: new point at distance `bond` from p2, angle theta (p1,p2,new),
    dihedral phi (p0,p1,p2,new)."""
    bc = p2 - p1
    bc /= np.linalg.norm(bc)
    n = np.cross(p1 - p0, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-8:
        n = np.cross(bc, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(n) < 1e-8:
            n = np.cross(bc, np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(theta),
                   bond * math.sin(theta) * math.cos(phi),
                   bond * math.sin(theta) * math.sin(phi)])
    return p2 + d2[0] * bc + d2[1] * m + d2[2] * n


def random_coil(rng, n: int, start=None, bond=5.9, angle_deg=120.0, angle_sd=5.0,
                clash=9.0, tries=200):
    """Self-avoiding seeded random coil with fixed bond length and angles
    drawn near angle_deg; beads |i-j| >= 3 are kept >= clash apart."""
    start = np.zeros(3) if start is None else np.asarray(start, float)
    for _attempt in range(100):
        x = np.zeros((n, 3))
        x[0] = start
        if n > 1:
            d = rng.normal(size=3)
            x[1] = x[0] + bond * d / np.linalg.norm(d)
        if n > 2:
            th = math.radians(rng.normal(angle_deg, angle_sd))
            ax = x[1] - x[0]
            ax /= np.linalg.norm(ax)
            perp = np.cross(ax, rng.normal(size=3))
            perp /= np.linalg.norm(perp)
            x[2] = x[1] + bond * (-math.cos(th) * ax + math.sin(th) * perp)
        ok = True
        for i in range(3, n):
            placed = False
            for _t in range(tries):
                th = math.radians(rng.normal(angle_deg, angle_sd))
                phi = rng.uniform(-math.pi, math.pi)
                p = _place(x[i - 3], x[i - 2], x[i - 1], bond, th, phi)
                if i >= 3:
                    dd = np.linalg.norm(x[: i - 2] - p, axis=1)
                    if dd.min() < clash:
                        continue
                x[i] = p
                placed = True
                break
            if not placed:
                ok = False
                break
        if ok:
            return x
    raise RuntimeError("random_coil: could not grow a self-avoiding chain")


def _mk(name, seqs, xyz, params, nrep=1, temps=None, nstep=1000):
    nmp_chain = np.array([len(s) for s in seqs], np.int32)
    seq = np.array([SEQT[c] for s in seqs for c in s], np.int32)
    xyz = np.asarray(xyz, np.float64)
    if xyz.ndim == 2:
        xyz = np.repeat(xyz[None], nrep, axis=0)
    temps = np.full(nrep, params["tempK"]) if temps is None else np.asarray(temps, float)
    return dict(name=name, nchains=len(seqs), nmp_chain=nmp_chain,
                moltype=np.full(len(seqs), MOLT_RNA, np.int32), seq=seq,
                seqs=list(seqs), xyz=np.ascontiguousarray(xyz), params=params,
                nrep=nrep, temps=temps, nstep=nstep)


def read_xyz_coords(path) -> np.ndarray:
    """Coordinates of an XYZ file as the reference reads xyz_ini: a count
    line, a comment line, then one `name x y z` line per bead."""
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    return np.array([[float(v) for v in l.split()[1:4]] for l in lines[2:2 + n]], np.float64)


def read_pdb_coords(path) -> np.ndarray:
    """Coordinates of the ATOM/HETATM records of a one-bead-per-nucleotide
    PDB file (columns 31-54), in file order."""
    out = []
    for l in open(path):
        if l.startswith(("ATOM", "HETATM")):
            out.append([float(l[30:38]), float(l[38:46]), float(l[46:54])])
    return np.array(out, np.float64)


def hairpin64(seed: int = 1) -> dict:
    """C1: one 64-nt hairpin chain: 5-nt A/U flank, 24-bp stem (60% GC),
    8-nt loop, reverse-complement stem, 3-nt flank (5+24+8+24+3 = 64),
    non-periodic, T = 300 K, margin 10 A."""
    rng = np.random.default_rng(seed)
    flank5 = "".join(rng.choice(list("AU"), 5))
    stem = random_seq(rng, 24, 0.6)
    loop = random_seq(rng, 8, 0.5)
    flank3 = "".join(rng.choice(list("AU"), 3))
    s = flank5 + stem + loop + revcomp(stem) + flank3
    x = random_coil(rng, len(s))
    p = default_params(tempK=300.0, nl_margin=10.0)
    return _mk("C1_hairpin64", [s], x, p, nstep=100_000)


def _place_chains(rng, seqs, box, clash=10.0):
    """Random coils with start points uniform in the box; whole chains are
    rejected while any bead sits within `clash` of an earlier chain
    (minimum image)."""
    from scipy.spatial import cKDTree
    coords = []
    tree = None
    n_tree = 0          # beads of the first chains, held in the tree
    recent = []         # wrapped beads of the chains placed since the tree was built
    for s in seqs:
        for _t in range(1000):
            start = rng.uniform(0.0, box, size=3)
            x = random_coil(rng, len(s), start=start)
            xw = np.mod(x, box)
            # nearest earlier bead under the minimum image: the tree's beads,
            # then the recent ones directly
            if tree is not None and tree.query(xw, k=1)[0].min() < clash:
                continue
            if recent:
                r = np.vstack(recent)
                d = xw[:, None, :] - r[None, :, :]
                d -= box * np.round(d / box)
                if np.sqrt((d ** 2).sum(axis=2)).min() < clash:
                    continue
            coords.append(x)
            recent.append(xw)
            if sum(len(q) for q in recent) > 2048:
                n_tree += sum(len(q) for q in recent)
                recent = []
                tree = cKDTree(np.mod(np.vstack(coords), box), boxsize=box)
            break
        else:
            raise RuntimeError("could not place chain")
    return np.vstack(coords)


def cag_system(nchains: int, length: int = 47, conc_mM: float = 1.0, tempK: float = 310.0,
               margin: float = 5.0, seed: int = 2, name: str = "CAG") -> dict:
    """nchains copies of a (CAG)n chain of `length` nt, periodic cube at
    conc_mM nucleotides (C2: 200 chains at 1 mM; C4: 2,000 at 2 mM)."""
    rng = np.random.default_rng(seed)
    s = ("CAG" * (length // 3 + 1))[:length]
    seqs = [s] * nchains
    nmp = nchains * length
    box = box_for_concentration(nmp, conc_mM)
    x = _place_chains(rng, seqs, box)
    p = default_params(tempK=tempK, nl_margin=margin, pbc=1, box=[box] * 3)
    return _mk(name, seqs, x, p)


def condensate_c2(seed: int = 2) -> dict:
    sysd = cag_system(200, 47, 1.0, 310.0, 5.0, seed, "C2_CAG47x200")
    sysd["nstep"] = 1_000_000
    return sysd


def condensate_c4(seed: int = 4) -> dict:
    sysd = cag_system(2000, 47, 2.0, 310.0, 5.0, seed, "C4_CAG47x2000")
    sysd["nstep"] = 200_000
    return sysd


def remd_c3(nrep: int = 256, n: int = 128, seed: int = 3, margin: float = 5.0) -> dict:
    """C3: nrep temperature replicas of one seeded random 128-nt chain
    (50% GC), geometric ladder 290-380 K, non-periodic."""
    rng = np.random.default_rng(seed)
    s = random_seq(rng, n, 0.5)
    x = random_coil(rng, n)
    if nrep > 1:
        temps = 290.0 * (380.0 / 290.0) ** (np.arange(nrep) / (nrep - 1))
    else:
        temps = np.array([290.0])
    p = default_params(tempK=float(temps[0]), nl_margin=margin)
    sysd = _mk("C3_remd%dx%d" % (nrep, n), [s], x, p, nrep=nrep, temps=temps, nstep=1_000_000)
    sysd["exchange_every"] = 1000
    return sysd


def serving_job(rng, max_chains: int = 64) -> dict:
    """One C5 job: 1-64 chains of 20-400 nt (40-70% GC), T in [280, 360] K,
    salt in [0.05, 0.5] M, periodic box at 0.5-5 mM, 1e4-1e6 steps
    (log-uniform)."""
    nch = int(rng.integers(1, max_chains + 1))
    gc = rng.uniform(0.4, 0.7)
    seqs = [random_seq(rng, int(rng.integers(20, 401)), gc) for _ in range(nch)]
    nmp = sum(len(s) for s in seqs)
    conc = rng.uniform(0.5, 5.0)
    box = box_for_concentration(nmp, conc)
    x = _place_chains(rng, seqs, box)
    tK = rng.uniform(280.0, 360.0)
    ion = rng.uniform(0.05, 0.5)
    nstep = int(round(10 ** rng.uniform(4.0, 6.0)))
    p = default_params(tempK=tK, ionic_strength=ion, nl_margin=5.0, pbc=1, box=[box] * 3)
    return _mk("C5_job", seqs, x, p, nstep=nstep)


def small_pbc(nchains: int = 3, length: int = 24, box: float = 120.0, seed: int = 7,
              tempK: float = 300.0, gc: float = 0.6) -> dict:
    """Small periodic multi-chain system for parity tests (chains of random
    sequence, box small enough that minimum-image pairs occur)."""
    rng = np.random.default_rng(seed)
    seqs = [random_seq(rng, length, gc) for _ in range(nchains)]
    x = _place_chains(rng, seqs, box, clash=6.0)
    p = default_params(tempK=tempK, nl_margin=5.0, pbc=1, box=[box] * 3)
    return _mk("small_pbc", seqs, x, p)
