"""Type-embedding DP (the reference's TypeDP, src/model/dp_dp_typ_emb.py:44-460).

The model differs from se_e2_a DP in the embedding only: one embedding net shared by every type
pair takes [s_ij, v(t_j)], the normalized physical-property vector of the neighbor's element
(dp_dp_typ_emb.py:361-403; vectors from utils/atom_type_emb_dict.py:143-161), instead of one net per
pair taking s_ij. Layer rules, the contraction A = Ri^T G / (m * type_nums), DR and the per-type
fitting nets are those of DP.

Because v(t_j) is constant within a neighbor-type bucket, the first layer's pre-activation splits:
    [s, v] W1 + b1 = s W1[0] + (v W1[1:] + b1),
so the model is exactly a per-pair DP whose pairs (t_i, t_j) share every weight except the first
layer's bias, which becomes b1 + v(t_j) W1[1:]. ``fold`` builds those DPWeights; every DP path here
(plain ops, fused BF16 kernels, MD, compress, the oracle) then runs type-embedding models unchanged.
``TypeDPTorch`` keeps the reference's concatenated-input form on the HIP ops.

Type vectors: ``type_vectors`` min-max normalizes per-element property values,
(value - min) / (max - min) over the property's whole table (atom_type_emb_dict.py:130-153). The
five default properties (nn_feature_type.py:35) come from the element table in
matpl_amd.elements; other properties take the caller's values and ranges.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .dp import DPTorch, _Net
from .dpnet import DPConfig, DPWeights, Layer

from . import elements

ATOMIC_NUMBER_RANGE = (1.0, 109.0)
DEFAULT_PROPERTIES = ("atomic_number", "atom_radius", "atom_mass", "electron_affin", "pauling")


def property_range(name: str) -> Tuple[float, float]:
    """(min, max) of a built-in property over the elements the table gives a number for."""
    if name == "atomic_number":
        return ATOMIC_NUMBER_RANGE
    vals = [v for v in elements.PROPERTIES[name] if v is not None and not np.isnan(v)]
    return float(min(vals)), float(max(vals))


def type_vectors(atomic_numbers: Sequence[int], properties: Sequence[str] = DEFAULT_PROPERTIES,
                 values: Optional[Dict[str, Dict[int, float]]] = None,
                 ranges: Optional[Dict[str, Tuple[float, float]]] = None) -> np.ndarray:
    """[T, P] normalized property vectors, one row per element (by atomic number). Built-in
    properties (DEFAULT_PROPERTIES) are read from the element table unless ``values`` gives them."""
    out = np.zeros((len(atomic_numbers), len(properties)))
    for p, name in enumerate(properties):
        if values and name in values:
            if not ranges or name not in ranges:
                raise ValueError(f"type_vectors: give the (min, max) range of property '{name}'")
            lo, hi = ranges[name]
            vals = [float(values[name][int(z)]) for z in atomic_numbers]
        elif name == "atomic_number":
            lo, hi = ATOMIC_NUMBER_RANGE
            vals = [float(z) for z in atomic_numbers]
        elif name in elements.PROPERTIES:
            lo, hi = property_range(name)
            vals = []
            for z in atomic_numbers:
                if not 1 <= int(z) <= len(elements.SYMBOLS):
                    raise ValueError(f"type_vectors: no element with atomic number {z}")
                v = elements.PROPERTIES[name][int(z) - 1]
                if v is None:
                    raise ValueError(f"type_vectors: the table has no '{name}' for {elements.SYMBOLS[int(z) - 1]}")
                vals.append(float(v))
        else:
            raise ValueError(f"type_vectors: give the values and (min, max) range of property '{name}'")
        for t, v in enumerate(vals):
            if np.isnan(v):
                raise ValueError(f"type_vectors: property '{name}' of element {atomic_numbers[t]} is NaN")
            out[t, p] = (v - lo) / (hi - lo)
    return out


@dataclass
class TypeDPWeights:
    cfg: DPConfig
    type_vec: np.ndarray          # [T, P]
    emb: List[Layer]              # shared net, first layer W [1 + P, E0]
    fit: List[List[Layer]]        # [T]
    davg: np.ndarray              # [T, 4]
    dstd: np.ndarray              # [T, 4]

    @staticmethod
    def random(cfg: DPConfig, type_vec, seed: int = 2025, davg=None, dstd=None) -> "TypeDPWeights":
        tv = np.asarray(type_vec, np.float64)
        base = DPWeights.random(cfg, seed=seed, davg=davg, dstd=dstd)
        rng = np.random.Generator(np.random.PCG64(seed + 1))
        W1, b1, r1 = base.emb[0][0]
        Wt = rng.normal(0.0, 1.0 / np.sqrt(1 + tv.shape[1]), size=(tv.shape[1], W1.shape[1]))
        emb = [(np.concatenate([W1, Wt], 0), b1, r1)] + list(base.emb[0][1:])
        return TypeDPWeights(cfg, tv, emb, base.fit, base.davg, base.dstd)


def fold(tw: TypeDPWeights) -> DPWeights:
    """The per-pair DPWeights of a type-embedding model (first-layer bias b1 + v(t_j) W1[1:])."""
    T = tw.cfg.ntypes
    (W1, b1, r1), rest = tw.emb[0], list(tw.emb[1:])
    b1 = np.zeros(W1.shape[1]) if b1 is None else np.asarray(b1, np.float64)
    per_tj = [(W1[:1].copy(), b1 + tw.type_vec[tj] @ W1[1:], r1) for tj in range(T)]
    emb = [[per_tj[p % T]] + rest for p in range(T * T)]
    return DPWeights(tw.cfg, emb, tw.fit, tw.davg, tw.dstd)


class TypeDPTorch(DPTorch):
    """TypeDP.forward on the HIP ops in the reference's form: the shared embedding net evaluated on
    [s_ij, v(t_j)] for every slot (dp_dp_typ_emb.py:388-394)."""

    def __init__(self, tw: TypeDPWeights, dtype=torch.float64, device="cuda", tall: bool = False):
        super().__init__(fold(tw), dtype, device, tall=tall)
        self.shared = _Net(tw.emb, 0, dtype, device)
        self.register_buffer("type_vec", torch.as_tensor(tw.type_vec, dtype=dtype, device=device))
        del self.emb

    def embed(self, ti: int, tj: int, Rt: torch.Tensor) -> torch.Tensor:
        s = Rt[..., 0:1]
        v = self.type_vec[tj].expand(*s.shape[:-1], self.type_vec.shape[1])
        return self.shared(torch.cat([s, v], -1), self.mm, self.add)

