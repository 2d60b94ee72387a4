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
