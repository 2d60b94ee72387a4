"""Per-replica observables read back from an Engine: radius of gyration,
total potential energy and the fraction of nucleotides in a formed base
pair (after the limit rule).  Used to compare ensembles (FP32 against FP64
pair arithmetic) by two-sample Kolmogorov-Smirnov tests."""
from __future__ import annotations

import numpy as np

from .abi import ENE


def radius_of_gyration(xyz: np.ndarray) -> np.ndarray:
    """Unweighted Rg of each replica, xyz (R, N, 3) -> (R,)."""
    c = xyz - xyz.mean(axis=1, keepdims=True)
    return np.sqrt((c ** 2).sum(axis=2).mean(axis=1))


def sample(eng) -> dict:
    """Rg, E_total (sis_gpu_energy, limit MC as in a step) and the paired
    nucleotide fraction of every replica of `eng`."""
    x = eng.get_state()[0]
    R, N = x.shape[0], x.shape[1]
    e = np.zeros(R)
    f = np.zeros(R)
    for r in range(1, R + 1):
        e[r - 1] = eng.energy_sumup(r)[ENE["TOTAL"]]
        b = eng.bp(r)
        f[r - 1] = 2.0 * float(np.sum(b["status"] > 0)) / N
    return dict(rg=radius_of_gyration(x), e_total=e, bp_fraction=f)


def ks_compare(a: dict, b: dict) -> dict:
    """Two-sample KS statistic and p-value per observable."""
    from scipy.stats import ks_2samp
    out = {}
    for k in a:
        res = ks_2samp(a[k], b[k])
        out[k] = dict(D=round(float(res.statistic), 4), p=round(float(res.pvalue), 4))
    return out


def unwrap_chains(xyz: np.ndarray, nmp_chain, box) -> np.ndarray:
    """Coordinates (N, 3) of one replica with every chain made whole: each
    bead follows its predecessor by the minimum-image bond vector (periodic
    boxes), so per-chain shapes do not depend on where the wrap cut them."""
    x = np.asarray(xyz, float).copy()
    L = np.asarray(box, float)
    if not np.all(L > 0):
        return x
    off = np.concatenate([[0], np.cumsum(nmp_chain)])
    d = np.diff(x, axis=0)
    d -= L * np.round(d / L)
    for c in range(len(nmp_chain)):
        a, b = off[c], off[c + 1]
        if b - a > 1:
            x[a + 1:b] = x[a] + np.cumsum(d[a:b - 1], axis=0)
    return x


def chain_sample(eng, irep: int = 1) -> dict:
    """Per-chain radius of gyration and paired-nucleotide fraction (after
    the limit rule) of one replica, and its total potential energy: the
    observables of a many-chain system (one value per chain)."""
    sysd = eng.sysd
    nmp_chain = np.asarray(sysd["nmp_chain"])
    x = eng.get_state()[0][irep - 1]
    box = sysd["params"]["box"] if sysd["params"].get("pbc") else [0, 0, 0]
    xu = unwrap_chains(x, nmp_chain, box)
    off = np.concatenate([[0], np.cumsum(nmp_chain)])
    rg = np.array([np.sqrt(((xu[a:b] - xu[a:b].mean(0)) ** 2).sum(1).mean()) for a, b in zip(off[:-1], off[1:])])
    b = eng.bp(irep)
    paired = np.zeros(x.shape[0])
    st = b["status"] > 0
    for i, j in b["bp"][st][:, :2]:
        paired[i - 1] = 1.0
        paired[j - 1] = 1.0
    frac = np.array([paired[a:b].mean() for a, b in zip(off[:-1], off[1:])])
    e = float(eng.energy_sumup(irep, -1.0, 0, 1)[ENE["TOTAL"]])   # BP state of the last step
    return dict(rg=rg, bp_fraction=frac, e_total=e)
