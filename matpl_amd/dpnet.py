"""DP se_e2_a model configuration and weights.

Net shapes and layer rules follow the reference's EmbeddingNet / FittingNet
(src/model/dp_embedding.py:40-160): the embedding net maps the scalar s_ij
through network_size = [1] + emb_sizes with tanh; a layer whose output width
equals its input width is residual (x = h + x, or h*resnet_dt + x), a layer
that doubles the width concatenates (x = [x, x] + h). The fitting net maps
the emb_sizes[-1]*M2 descriptor through fit_sizes; its first layer is plain,
later hidden layers are residual with resnet_dt, the last layer is linear.
Defaults are the reference's (src/user/nn_feature_type.py:10-29,
src/user/model_param.py:66-75): embedding [25,25,25] bias, no resnet_dt;
fitting [50,50,50,1] bias, resnet_dt; M2 = 16.

Weights are drawn from a seeded PCG64 stream: W ~ N(0, 1/fan_in), biases ~
N(0, 1) (the reference's bias init, dp_embedding.py:63,121,136, with zero
energy shift), resnet_dt ~ N(0.1, 0.001) (dp_embedding.py:126).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

Layer = Tuple[np.ndarray, Optional[np.ndarray], Optional[np.ndarray]]   # (W[in,out], b[out], rdt[out])


@dataclass
class DPConfig:
    ntypes: int
    rcut: float = 6.0            # Rmax (Rc_M)
    rcut_smooth: float = 5.5     # Rmin
    m: int = 100                 # maxNeighborNum, per neighbor type
    M2: int = 16
    emb_sizes: List[int] = field(default_factory=lambda: [25, 25, 25])
    fit_sizes: List[int] = field(default_factory=lambda: [50, 50, 50, 1])
    emb_bias: bool = True
    emb_resnet_dt: bool = False
    fit_bias: bool = True
    fit_resnet_dt: bool = True

    @property
    def S(self) -> int:
        return self.ntypes * self.m

    @property
    def emb_out(self) -> int:
        return self.emb_sizes[-1]


def _emb_layer_rdt(cfg: DPConfig, i: int) -> bool:
    sizes = [1] + cfg.emb_sizes
    return cfg.emb_resnet_dt and (sizes[i + 1] == sizes[i] or sizes[i + 1] == 2 * sizes[i])


def _fit_layer_rdt(cfg: DPConfig, i: int) -> bool:
    nl = len(cfg.fit_sizes)
    return cfg.fit_resnet_dt and 0 < i < nl - 1


@dataclass
class DPWeights:
    cfg: DPConfig
    emb: List[List[Layer]]       # [ntypes*ntypes] nets, index ti*ntypes + tj
    fit: List[List[Layer]]       # [ntypes] nets
    davg: np.ndarray             # [ntypes, 4]
    dstd: np.ndarray             # [ntypes, 4]

    @staticmethod
    def random(cfg: DPConfig, seed: int = 2023, davg=None, dstd=None) -> "DPWeights":
        rng = np.random.Generator(np.random.PCG64(seed))

        def layer(fan_in, fan_out, bias, rdt):
            W = rng.normal(0.0, 1.0 / np.sqrt(fan_in), size=(fan_in, fan_out))
            b = rng.normal(0.0, 1.0, size=fan_out) if bias else None
            r = rng.normal(0.1, 0.001, size=fan_out) if rdt else None
            return (W, b, r)

        T = cfg.ntypes
        esz = [1] + cfg.emb_sizes
        emb = [[layer(esz[i], esz[i + 1], cfg.emb_bias, _emb_layer_rdt(cfg, i))
                for i in range(len(esz) - 1)] for _ in range(T * T)]
        fsz = [cfg.emb_out * cfg.M2] + cfg.fit_sizes
        fit = [[layer(fsz[i], fsz[i + 1], cfg.fit_bias, _fit_layer_rdt(cfg, i))
                for i in range(len(fsz) - 1)] for _ in range(T)]
        if davg is None:
            davg = np.zeros((T, 4))
        if dstd is None:
            dstd = np.ones((T, 4))
        return DPWeights(cfg, emb, fit, np.asarray(davg, np.float64), np.asarray(dstd, np.float64))

    def davg_full(self) -> np.ndarray:
        """davg tiled to the reference layout [ntypes, S*4] (dp_dp.py:293-296)."""
        return np.tile(self.davg[:, None, :], (1, self.cfg.S, 1)).reshape(self.cfg.ntypes, -1)

    def dstd_full(self) -> np.ndarray:
        return np.tile(self.dstd[:, None, :], (1, self.cfg.S, 1)).reshape(self.cfg.ntypes, -1)
