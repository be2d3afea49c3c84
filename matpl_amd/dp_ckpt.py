"""DP model checkpoints in MatPL's layout, read into DPWeights and written back.

MatPL saves a DP model as one torch.save dict (dp_network.py:566-601):

    json_file        the run's parameters (input_param.py:276-327): model_type, atom_type,
                     max_neigh_num, model.descriptor {Rmax, Rmin, M2, network_size}
                     (nn_feature_type.py:271-283), model.fitting_net {network_size}
    epoch            int
    state_dict       DP.state_dict(): embedding_net.{ti*T+tj}.layers.{l}.{weight,bias,resnet_dt}
                     and fitting_net.{t}.layers.{l}.{weight,bias,resnet_dt} (dp_dp.py:69-85,
                     dp_embedding.py:30-38); weight [in, out], bias and resnet_dt [1, out]
    davg, dstd       numpy [T, S*4], the per-type mean and deviation tiled over the S slots
    energy_shift     per-type energy shifts (the fitting nets' initial output bias)
    atom_type_order  numpy array of atomic numbers, the type order of all of the above
    sij_max          largest normalized s seen in training (used by compression)
    [optimizer]      optimizer state (ignored here)

The file is read with torch.load(weights_only=True); numpy arrays and scalars are admitted
through torch's allow-list of numpy's array reconstruction, and nothing else is unpickled.

EmbeddingNet gives every layer a resnet_dt when the flag is set, but uses it only on layers whose
output width equals or doubles the input (dp_embedding.py:73-94); FittingNet uses it on hidden
layers after the first (dp_embedding.py:123-127, 140-152). The reader keeps resnet_dt exactly
where the forward uses it, which is the DPWeights convention.
"""
from __future__ import annotations

from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .dpnet import DPConfig, DPWeights, Layer
from .typedp import DEFAULT_PROPERTIES, TypeDPWeights, fold, type_vectors


def _numpy_globals() -> list:
    core = getattr(np, "_core", None) or np.core
    allowed = [np.ndarray, np.dtype, core.multiarray._reconstruct, core.multiarray.scalar]
    for t in (np.float64, np.float32, np.int64, np.int32, np.int16, np.int8, np.uint8, np.bool_, np.str_):
        allowed.append(type(np.dtype(t)))
    return allowed


def load_checkpoint(path: str) -> Dict[str, Any]:
    """torch.load with weights_only=True, admitting numpy arrays and scalars."""
    with torch.serialization.safe_globals(_numpy_globals()):
        return torch.load(path, map_location="cpu", weights_only=True)


def _np(t) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().double().numpy()
    return np.asarray(t, np.float64)


def _net(sd: Dict[str, torch.Tensor], prefix: str, kind: int) -> List[Layer]:
    layers: List[Layer] = []
    l = 0
    while f"{prefix}.layers.{l}.weight" in sd:
        W = _np(sd[f"{prefix}.layers.{l}.weight"])
        b = sd.get(f"{prefix}.layers.{l}.bias")
        r = sd.get(f"{prefix}.layers.{l}.resnet_dt")
        layers.append((W.copy(), None if b is None else _np(b).reshape(-1).copy(),
                       None if r is None else _np(r).reshape(-1).copy()))
        l += 1
    if not layers:
        raise ValueError(f"checkpoint has no layers under '{prefix}'")
    n = len(layers)
    out: List[Layer] = []
    for l, (W, b, r) in enumerate(layers):
        ni, no = W.shape
        used = (no == ni or no == 2 * ni) if kind == 0 else (0 < l < n - 1)
        out.append((W, b, r if used else None))
    return out


def read_dp_checkpoint(path: str) -> Tuple[DPWeights, List[int], Dict[str, Any]]:
    """(weights, atomic numbers in type order, {epoch, energy_shift, sij_max, compressed, typedp}).

    A type-embedding (TypeDP) checkpoint is read too: its shared embedding net on [s, v(t_j)]
    (src/model/dp_dp_typ_emb.py:66-97, 361-403) with the type vectors of its physical properties
    (json_file model.descriptor.type_embedding.physical_property, default nn_feature_type.py:35)
    comes back exactly folded into per-pair DPWeights (matpl_amd.typedp.fold), and unfolded as
    info["typedp"] (TypeDPWeights)."""
    ck = load_checkpoint(path)
    for key in ("json_file", "state_dict", "davg", "dstd", "atom_type_order"):
        if key not in ck:
            raise ValueError(f"{path}: not a MatPL DP checkpoint (no '{key}')")
    js = ck["json_file"]
    if str(js.get("model_type", "DP")).upper() != "DP":
        raise ValueError(f"{path}: model_type {js.get('model_type')} is not DP")
    sd = {k[7:] if k.startswith("module.") else k: v for k, v in ck["state_dict"].items()}
    order = np.asarray(ck["atom_type_order"]).reshape(-1).astype(np.int64).tolist()
    T = len(order)
    desc = js.get("model", {}).get("descriptor", {})
    nemb = 0
    while f"embedding_net.{nemb}.layers.0.weight" in sd:
        nemb += 1
    typedp = bool(desc.get("type_embedding")) or str(js.get("model_type", "")).upper() == "TYPEDP" or (
        nemb and _np(sd["embedding_net.0.layers.0.weight"]).shape[0] > 1)
    fit = [_net(sd, f"fitting_net.{t}", 1) for t in range(T)]
    if typedp:
        # TypeDP (dp_dp_typ_emb.py:66-97): the last embedding net is the shared one, on [s, v(t_j)]
        shared = _net(sd, f"embedding_net.{nemb - 1}", 0)
        props = list((desc.get("type_embedding") or {}).get("physical_property", DEFAULT_PROPERTIES))
        if shared[0][0].shape[0] != 1 + len(props):
            raise ValueError(f"{path}: shared embedding input width {shared[0][0].shape[0]}, expected "
                             f"1 + {len(props)} physical properties")
        emb = None
    else:
        emb = [_net(sd, f"embedding_net.{p}", 0) for p in range(T * T)]
        if emb[0][0][0].shape[0] != 1:
            raise ValueError(f"{path}: embedding input width {emb[0][0][0].shape[0]} without a type embedding")
    first = shared if typedp else emb[0]
    E = first[-1][0].shape[1]
    M2 = fit[0][0][0].shape[0] // E
    m = int(js.get("max_neigh_num", 0))
    davg_full, dstd_full = _np(ck["davg"]), _np(ck["dstd"])
    if davg_full.shape[0] != T or davg_full.shape[1] % 4 or dstd_full.shape != davg_full.shape:
        raise ValueError(f"{path}: davg/dstd shape {davg_full.shape} does not match {T} types")
    S = davg_full.shape[1] // 4
    if m == 0:
        m = S // T
    if S != T * m:
        raise ValueError(f"{path}: davg covers {S} slots, max_neigh_num {m} x {T} types needs {T * m}")
    davg = davg_full.reshape(T, S, 4)
    dstd = dstd_full.reshape(T, S, 4)
    if not (np.all(davg == davg[:, :1]) and np.all(dstd == dstd[:, :1])):
        raise ValueError(f"{path}: davg/dstd vary over neighbor slots; this build keeps one [4] vector per type")
    cfg = DPConfig(ntypes=T, rcut=float(desc.get("Rmax", 6.0)), rcut_smooth=float(desc.get("Rmin", 0.5)), m=m,
                   M2=M2, emb_sizes=[L[0].shape[1] for L in first], fit_sizes=[L[0].shape[1] for L in fit[0]],
                   emb_bias=first[0][1] is not None,
                   emb_resnet_dt=any(L[2] is not None for L in first),
                   fit_bias=fit[0][0][1] is not None,
                   fit_resnet_dt=any(L[2] is not None for L in fit[0]))
    info = {"epoch": ck.get("epoch"), "energy_shift": ck.get("energy_shift"), "sij_max": ck.get("sij_max"),
            "compressed": "compress" in ck, "typedp": None}
    if typedp:
        tw = TypeDPWeights(cfg, type_vectors(order, props), shared, fit, davg[:, 0].copy(), dstd[:, 0].copy())
        info["typedp"] = tw
        info["physical_property"] = props
        return fold(tw), order, info
    w = DPWeights(cfg, emb, fit, davg[:, 0].copy(), dstd[:, 0].copy())
    return w, order, info


def write_dp_checkpoint(w: DPWeights, atom_types: Sequence[int], path: str,
                        energy_shift: Optional[Sequence[float]] = None, sij_max: Optional[float] = None,
                        epoch: int = 0) -> None:
    """MatPL's layout, loadable by its DP.load_state_dict and load_davg_dstd_from_checkpoint."""
    cfg = w.cfg
    T = cfg.ntypes
    if len(atom_types) != T:
        raise ValueError(f"{len(atom_types)} atom types for a {T}-type model")
    sd: Dict[str, torch.Tensor] = {}
    t = lambda a: torch.as_tensor(np.asarray(a, np.float64))
    for name, nets in (("embedding_net", w.emb), ("fitting_net", w.fit)):
        for p, layers in enumerate(nets):
            for l, (W, b, r) in enumerate(layers):
                sd[f"{name}.{p}.layers.{l}.weight"] = t(W)
                if b is not None:
                    sd[f"{name}.{p}.layers.{l}.bias"] = t(b).reshape(1, -1)
                # EmbeddingNet registers a resnet_dt on every layer when the flag is set
                # (dp_embedding.py:65-70), also where its forward never reads it (the 1 -> E layer):
                # those get ones, the mean of MatPL's initialisation there
                if r is None and name == "embedding_net" and cfg.emb_resnet_dt:
                    r = np.ones(W.shape[1])
                if r is not None:
                    sd[f"{name}.{p}.layers.{l}.resnet_dt"] = t(r).reshape(1, -1)
    if energy_shift is None:
        energy_shift = [float(np.asarray(net[-1][1]).reshape(-1)[0]) if net[-1][1] is not None else 0.0
                        for net in w.fit]
    js = {"model_type": "DP", "atom_type": [int(a) for a in atom_types], "max_neigh_num": cfg.m,
          "model": {"descriptor": {"Rmax": cfg.rcut, "Rmin": cfg.rcut_smooth, "M2": cfg.M2,
                                   "network_size": list(cfg.emb_sizes)},
                    "fitting_net": {"network_size": list(cfg.fit_sizes)}}}
    torch.save({"json_file": js, "epoch": epoch, "state_dict": sd,
                "davg": w.davg_full(), "dstd": w.dstd_full(),
                "energy_shift": [float(e) for e in energy_shift],
                "atom_type_order": np.array([int(a) for a in atom_types]),
                "sij_max": sij_max}, path)
