"""NEP model files: the nep.txt text format MatPL and GPUMD write, read into NEPWeights and written back.

Layout (restated from the reference reader, src/user/nep_param.py:94-210, and writer,
src/model/nep_net.py:76-93 with nep_fitting.py:107-133), one value per line after a header:

    nep4|nep5  T  el_1 .. el_T
    [zbl inner outer]                   (optional; not supported here)
    cutoff  rc_radial  rc_angular  max_NN_radial  max_NN_angular
    n_max  n_radial  n_angular
    basis_size  k_radial  k_angular
    l_max  l3  l4  l5
    ANN  hidden  0
    per type t: W0 [hidden][F] (row-major), -b0 [hidden], W1 [hidden]  (nep5: then -bias_t)
    nep4 written by MatPL: -bias_t for every type; by GPUMD: one common -bias; nep5: one 0
    c2 [n_radial+1][k_radial+1][T][T], c3 [n_angular+1][k_angular+1][T][T] (when l3 > 0)
    q_scaler [F]

The hidden layer is tanh(x W0 - b0_file) and the output h W1 - bias_file, so biases are stored
negated. A nep4 file is MatPL's (per-type output biases) when its line count matches that layout
exactly and GPUMD's (one common bias) otherwise, the reference's own test.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

from .nepnet import NEPConfig, NEPWeights


def _ints(line: str, n: int) -> List[int]:
    return [int(x) for x in line.split()[1:1 + n]]


def read_nep_txt(path: str) -> Tuple[NEPWeights, List[str], Tuple[int, int]]:
    """(weights, element symbols in type order, (max_NN_radial, max_NN_angular))."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    while lines and not lines[-1].strip():
        lines.pop()
    head = lines[0].split()
    version, T, elements = head[0], int(head[1]), head[2:2 + int(head[1])]
    if not version.startswith("nep") or ("4" not in version and "5" not in version):
        raise ValueError(f"{path}: not a nep4/nep5 file (header '{lines[0]}')")
    if lines[1].split()[0] == "zbl":
        raise NotImplementedError(f"{path}: ZBL-augmented NEP models are not supported by this build")
    cut = lines[1].split()
    rc_r, rc_a = float(cut[1]), float(cut[2])
    max_nn = (int(cut[3]), int(cut[4])) if len(cut) >= 5 else (100, 100)
    n_r, n_a = _ints(lines[2], 2)
    k_r, k_a = _ints(lines[3], 2)
    l3, l4, l5 = _ints(lines[4], 3)
    hidden = _ints(lines[5], 1)[0]
    cfg = NEPConfig(ntypes=T, rcut_radial=rc_r, rcut_angular=rc_a, n_max_radial=n_r, n_max_angular=n_a,
                    basis_radial=k_r, basis_angular=k_a, l_max=[l3, l4, l5], neuron=[hidden])
    F = cfg.nfeat
    n_c2 = T * T * cfg.nr * cfg.nbr
    n_c3 = T * T * cfg.na * cfg.nba if l3 > 0 else 0
    nep5 = "5" in version
    per_type = F * hidden + 2 * hidden + (1 if nep5 else 0)
    vals = lines[6:]
    base = T * per_type + n_c2 + n_c3 + F
    if nep5:
        bias_mode = "nep5"
        need = base + 1
    elif len(vals) == base + T:
        bias_mode, need = "per_type", base + T
    else:
        bias_mode, need = "common", base + 1
    if len(vals) != need:
        raise ValueError(f"{path}: {len(vals)} values after the header, the layout needs {need}")
    x = np.array([float(v) for v in vals], np.float64)
    p = 0
    fit, last = [], []
    for _ in range(T):
        W0 = x[p:p + F * hidden].reshape(hidden, F).T.copy()
        p += F * hidden
        b0 = -x[p:p + hidden]
        p += hidden
        W1 = x[p:p + hidden].reshape(hidden, 1).copy()
        p += hidden
        if nep5:
            last.append(-x[p])
            p += 1
        fit.append([W0, b0, W1])
    if bias_mode == "per_type":
        last = list(-x[p:p + T])
        p += T
    elif bias_mode == "common":
        last = [-x[p]] * T
        p += 1
    else:
        p += 1                                    # nep5: the common bias slot holds 0
    c2 = x[p:p + n_c2].reshape(cfg.nr, cfg.nbr, T, T).transpose(2, 3, 0, 1).copy()
    p += n_c2
    if l3 > 0:
        c3 = x[p:p + n_c3].reshape(cfg.na, cfg.nba, T, T).transpose(2, 3, 0, 1).copy()
        p += n_c3
    else:
        c3 = np.zeros((T, T, cfg.na, cfg.nba))
    q_scaler = x[p:p + F].copy()
    p += F
    assert p == len(x)
    nets = [[(W0, b0, None), (W1, np.array([bl], np.float64), None)] for (W0, b0, W1), bl in zip(fit, last)]
    return NEPWeights(cfg, c2, c3, q_scaler, nets), list(elements), max_nn


def write_nep_txt(w: NEPWeights, elements: List[str], path: str, max_nn=(100, 100)) -> None:
    """MatPL's nep4 layout (per-type output biases); one hidden tanh layer per type."""
    cfg = w.cfg
    T = cfg.ntypes
    if len(cfg.neuron) != 1 or any(len(net) != 2 for net in w.fit):
        raise ValueError("nep.txt holds one hidden layer per type")
    out = [f"nep4   {T} " + " ".join(elements),
           f"cutoff {cfg.rcut_radial} {cfg.rcut_angular} {max_nn[0]} {max_nn[1]}",
           f"n_max  {cfg.n_max_radial} {cfg.n_max_angular}",
           f"basis_size {cfg.basis_radial} {cfg.basis_angular}",
           "l_max  " + " ".join(str(v) for v in cfg.l_max),
           f"ANN    {cfg.neuron[0]} 0"]
    vals: List[float] = []
    for (W0, b0, _), (W1, _b1, _) in w.fit:
        vals += list(np.asarray(W0).T.reshape(-1))
        vals += list(-np.asarray(b0).reshape(-1))
        vals += list(np.asarray(W1).reshape(-1))
    vals += [-float(np.asarray(net[1][1]).reshape(-1)[0]) for net in w.fit]
    vals += list(np.asarray(w.c2).transpose(2, 3, 0, 1).reshape(-1))
    if cfg.l_max[0] > 0:
        vals += list(np.asarray(w.c3).transpose(2, 3, 0, 1).reshape(-1))
    vals += list(np.asarray(w.q_scaler).reshape(-1))
    with open(path, "w") as f:
        f.write("\n".join(out + [repr(float(v)) for v in vals]) + "\n")
