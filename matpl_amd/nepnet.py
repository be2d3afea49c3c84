"""NEP model configuration and weights.

Descriptor (src/model/nep_net.py:103-142, src/user/nep_param.py:60-83):
  radial   q_n   = sum_j g_n(r_ij),  g_n(r) = sum_k c2[ti,tj,n,k] f_k(r),  n < n_max_r + 1
  angular  s_nlm = sum_j g'_n(r_ij) b_lm(r_ij / |r_ij|) over j within rcut_angular,
           q_nl = sum_m C3B_lm s_nlm^2 (L = 1..l_max_3), optional 4-body
           (l_max_4 = 2) and 5-body (l_max_5 = 1) invariants
  f_k(r) = (T_k(x) + 1)/2 * fc(r), x = 2 (r/rc - 1)^2 - 1, fc = (1 + cos(pi r/rc))/2
Feature layout per atom: [q_0..q_{nr-1}, q_{l=1,n=0..}, ..., q_{l=4,n}, 4-body_n, 5-body_n].
Energy: Ei = fit_{t_i}(q_scaler * q); the fitting net is the DP-style
FittingNet (src/model/nep_fitting.py:40-163): first hidden layer plain tanh,
further hidden layers residual, linear output.

Weights are seeded (PCG64): c ~ U(-0.5, 0.5) + 0.1 N(0,1) (the reference's
init, nep_net.py:163-168), NN W ~ N(0, 1/fan_in), b ~ N(0, 1), q_scaler = 1/(max-min)
of the features over the initial frame(s) when given, else 1.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class NEPConfig:
    ntypes: int
    rcut_radial: float = 6.0
    rcut_angular: float = 4.0
    n_max_radial: int = 4        # features n = 0..n_max
    n_max_angular: int = 4
    basis_radial: int = 12       # Chebyshev k = 0..basis
    basis_angular: int = 12
    l_max: List[int] = field(default_factory=lambda: [4, 2, 0])
    neuron: List[int] = field(default_factory=lambda: [30])

    @property
    def nr(self) -> int:
        return self.n_max_radial + 1

    @property
    def na(self) -> int:
        return self.n_max_angular + 1

    @property
    def nbr(self) -> int:
        return self.basis_radial + 1

    @property
    def nba(self) -> int:
        return self.basis_angular + 1

    @property
    def n_angular_feats(self) -> int:
        l3, l4, l5 = self.l_max
        return self.na * l3 + (self.na if l4 > 0 else 0) + (self.na if l5 > 0 else 0)

    @property
    def nfeat(self) -> int:
        return self.nr + self.n_angular_feats


@dataclass
class NEPWeights:
    cfg: NEPConfig
    c2: np.ndarray               # [T, T, nr, nbr]
    c3: np.ndarray               # [T, T, na, nba]
    q_scaler: np.ndarray         # [F]
    fit: list                    # [T] nets, each a list of (W, b, rdt)

    @staticmethod
    def random(cfg: NEPConfig, seed: int = 2024, q_scaler=None) -> "NEPWeights":
        rng = np.random.Generator(np.random.PCG64(seed))
        T = cfg.ntypes
        n2 = T * T * cfg.nr * cfg.nbr
        n3 = T * T * cfg.na * cfg.nba
        c = rng.uniform(-0.5, 0.5, size=n2 + n3) + 0.1 * rng.normal(size=n2 + n3)
        c2 = c[:n2].reshape(T, T, cfg.nr, cfg.nbr)
        c3 = c[n2:].reshape(T, T, cfg.na, cfg.nba)
        sizes = [cfg.nfeat] + list(cfg.neuron) + [1]
        fit = []
        for _ in range(T):
            layers = []
            for i in range(len(sizes) - 1):
                W = rng.normal(0.0, 1.0 / np.sqrt(sizes[i]), size=(sizes[i], sizes[i + 1]))
                b = rng.normal(0.0, 1.0, size=sizes[i + 1])
                r = rng.normal(0.1, 0.001, size=sizes[i + 1]) if (0 < i < len(sizes) - 2) else None
                layers.append((W, b, r))
            fit.append(layers)
        qs = np.ones(cfg.nfeat) if q_scaler is None else np.asarray(q_scaler, np.float64)
        return NEPWeights(cfg, c2, c3, qs, fit)
