"""DP compress oracle (TEST INFRASTRUCTURE): evaluation of the tabulated embedding.

  horner / horner_d   compress_calc / compress_calc_grad (src/op/kernel/calculateCompress.cu,
                      calculateCompress_grad.cu): G = sum_q c[q] f2^(K-1-q), highest power first,
                      and dG/df2; pinned against numpy.polyval / numpy.polyder.
  compress_forward    CalculateCompressFuncs::forward/backward (src/op/src/CalcOps.cpp:331-357):
                      G [n, L] and the f2 gradient sum_i gG_i dG_i/df2.
"""
from __future__ import annotations

import numpy as np


def horner(f2, coef):
    """f2 [n], coef [n, L, K] -> G [n, L]."""
    f2 = np.asarray(f2, np.float64).reshape(-1, 1)
    coef = np.asarray(coef, np.float64)
    r = coef[..., 0].copy()
    for q in range(1, coef.shape[-1]):
        r = r * f2 + coef[..., q]
    return r


def horner_d(f2, coef):
    """dG/df2 [n, L]."""
    f2 = np.asarray(f2, np.float64).reshape(-1, 1)
    coef = np.asarray(coef, np.float64)
    K = coef.shape[-1]
    r = (K - 1) * coef[..., 0]
    for q in range(1, K - 1):
        r = r * f2 + (K - 1 - q) * coef[..., q]
    return r


def compress_forward(f2, coef, gG=None):
    G = horner(f2, coef)
    if gG is None:
        return G
    return G, (np.asarray(gG, np.float64) * horner_d(f2, coef)).sum(-1)
