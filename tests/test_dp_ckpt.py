"""DP checkpoints in MatPL's layout (dp_network.py:566-601): a checkpoint saved from modules built
the way MatPL's DP builds them (ModuleLists embedding_net / fitting_net of nets whose ModuleList
`layers` holds weight [in, out], bias and resnet_dt [1, out]; dp_dp.py:69-85,
dp_embedding.py:30-128) is read into DPWeights: key names come from torch's own module naming,
embedding outputs of the read weights equal the modules' at sample inputs, the energy of the read
model equals the oracle's on the source weights, and write -> read is exact."""
import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dp_ckpt import load_checkpoint, read_dp_checkpoint, write_dp_checkpoint
from matpl_amd.dp_fused import _emb_forward_np
from matpl_amd.dpnet import DPConfig, DPWeights


class _Layer(torch.nn.Module):
    def __init__(self, W, b, r):
        super().__init__()
        self.weight = torch.nn.Parameter(torch.as_tensor(W))
        self.bias = torch.nn.Parameter(torch.as_tensor(b).reshape(1, -1)) if b is not None else None
        self.resnet_dt = torch.nn.Parameter(torch.as_tensor(r).reshape(1, -1)) if r is not None else None


class _Net(torch.nn.Module):
    def __init__(self, layers):
        super().__init__()
        self.layers = torch.nn.ModuleList([_Layer(*L) for L in layers])


class _DP(torch.nn.Module):
    def __init__(self, emb, fit):
        super().__init__()
        self.embedding_net = torch.nn.ModuleList([_Net(n) for n in emb])
        self.fitting_net = torch.nn.ModuleList([_Net(n) for n in fit])


def _matpl_layers(w: DPWeights, rng):
    """Weights as MatPL's modules hold them: embedding layers all own a resnet_dt when the flag is
    set (dp_embedding.py:65-68), fitting layers from the second to the last hidden one
    (dp_embedding.py:123-126)."""
    emb = [[(W, b, r if r is not None else rng.normal(1.0, 1e-3, W.shape[1])) for (W, b, r) in net]
           for net in w.emb]
    return emb, [list(net) for net in w.fit]


def _checkpoint(w, atom_types, path, rng):
    emb, fit = _matpl_layers(w, rng)
    model = _DP(emb, fit).double()
    torch.save({"json_file": {"model_type": "DP", "atom_type": atom_types, "max_neigh_num": w.cfg.m,
                              "model": {"descriptor": {"Rmax": w.cfg.rcut, "Rmin": w.cfg.rcut_smooth,
                                                       "M2": w.cfg.M2, "network_size": w.cfg.emb_sizes},
                                        "fitting_net": {"network_size": w.cfg.fit_sizes}}},
                "epoch": 7, "state_dict": model.state_dict(), "davg": w.davg_full(), "dstd": w.dstd_full(),
                "energy_shift": [-3.5, -7.25], "atom_type_order": np.array(atom_types), "sij_max": 1.75},
               path)
    return model


def _weights(seed=5):
    cfg = DPConfig(ntypes=2, m=120, emb_resnet_dt=True, fit_resnet_dt=True)
    rng = np.random.default_rng(seed)
    return DPWeights.random(cfg, seed=seed, davg=rng.normal(0.1, 0.05, (2, 4)),
                            dstd=rng.uniform(0.5, 1.5, (2, 4)))


def test_reads_matpl_layout(tmp_path):
    w = _weights()
    rng = np.random.default_rng(1)
    model = _checkpoint(w, [72, 8], str(tmp_path / "dp_model.ckpt"), rng)
    sd = model.state_dict()
    assert "embedding_net.3.layers.0.resnet_dt" in sd and "fitting_net.1.layers.3.weight" in sd
    assert tuple(sd["embedding_net.0.layers.1.bias"].shape) == (1, 25)
    r, order, info = read_dp_checkpoint(str(tmp_path / "dp_model.ckpt"))
    assert order == [72, 8] and info["epoch"] == 7 and info["sij_max"] == 1.75 and not info["compressed"]
    assert r.cfg.m == 120 and r.cfg.M2 == 16 and r.cfg.emb_sizes == [25, 25, 25] and r.cfg.fit_sizes == [50, 50, 50, 1]
    assert np.array_equal(r.davg, w.davg) and np.array_equal(r.dstd, w.dstd)
    # resnet_dt of the 1 -> 25 embedding layer is unused by MatPL's forward and dropped
    assert all(net[0][2] is None for net in r.emb)
    # the embedding nets as MatPL's forward computes them (dp_embedding.py:73-94) at sample inputs
    for p in range(4):
        layers = model.embedding_net[p].layers
        for s in (-0.3, 0.0, 0.7, 2.1):
            x = torch.tensor([[s]], dtype=torch.float64)
            for i, L in enumerate(layers):
                h = torch.tanh(x @ L.weight + L.bias)
                ni, no = L.weight.shape
                x = h * L.resnet_dt + x if no == ni else (torch.cat((x, x), -1) + h * L.resnet_dt if no == 2 * ni else h)
            assert np.allclose(_emb_forward_np(r.emb[p], s), x.detach().numpy().reshape(-1), rtol=0, atol=1e-15)
    # the model read back is the model written: energy and forces on the oracle
    f = st.fluorite_hfo2((1, 1, 1), seed=3)
    a, b = oracle.dp_evaluate(f, w), oracle.dp_evaluate(f, r)
    assert a["E"] == b["E"] and np.array_equal(a["F"], b["F"])


def test_write_read_round_trip(tmp_path):
    w = _weights(seed=9)
    path = str(tmp_path / "rt.ckpt")
    write_dp_checkpoint(w, [3, 16], path, sij_max=2.0)
    r, order, info = read_dp_checkpoint(path)
    assert order == [3, 16] and info["sij_max"] == 2.0
    for A, B in ((w.emb, r.emb), (w.fit, r.fit)):
        for na, nb in zip(A, B):
            for (Wa, ba, ra), (Wb, bb, rb) in zip(na, nb):
                assert np.array_equal(Wa, Wb) and np.array_equal(ba, bb)
                assert (ra is None) == (rb is None) and (ra is None or np.array_equal(ra, rb))
    ck = load_checkpoint(path)
    # the fields MatPL's loader reads (dp_param_extract.py:318-326)
    assert ck["atom_type_order"].size == 2 and ck["davg"].shape == (2, 2 * 120 * 4)


@pytest.mark.parametrize("emb_rdt,fit_rdt", [(True, True), (False, True), (True, False), (False, False)])
def test_written_state_dict_loads_strictly_into_matpl_modules(tmp_path, emb_rdt, fit_rdt):
    """MatPL loads a checkpoint with model.load_state_dict(checkpoint['state_dict']) in strict mode
    (dp_network.py:225), into modules that own a resnet_dt on every embedding layer when the flag is
    set (dp_embedding.py:65-70) and on fitting layers 1..n-2 (dp_embedding.py:123-126)."""
    cfg = DPConfig(ntypes=2, m=120, emb_resnet_dt=emb_rdt, fit_resnet_dt=fit_rdt)
    w = DPWeights.random(cfg, seed=4)
    path = str(tmp_path / "strict.ckpt")
    write_dp_checkpoint(w, [72, 8], path)
    ones = lambda n: np.ones(n)
    emb = [[(W, b, (r if r is not None else ones(W.shape[1])) if emb_rdt else None) for (W, b, r) in net]
           for net in w.emb]
    fit = [[(W, b, r if (fit_rdt and 0 < l < len(net) - 1) else None) for l, (W, b, r) in enumerate(net)]
           for net in w.fit]
    model = _DP(emb, fit).double()
    missing, unexpected = model.load_state_dict(load_checkpoint(path)["state_dict"], strict=True)
    assert not missing and not unexpected
    for p in range(4):
        for l, (W, b, r) in enumerate(w.emb[p]):
            assert np.array_equal(model.embedding_net[p].layers[l].weight.detach().numpy(), W)


def test_rejects_other_layouts(tmp_path):
    w = _weights()
    path = str(tmp_path / "bad.ckpt")
    write_dp_checkpoint(w, [72, 8], path)
    ck = load_checkpoint(path)
    ck["davg"] = ck["davg"] + np.arange(ck["davg"].shape[1])[None] * 1e-3
    torch.save(ck, path)
    with pytest.raises(ValueError, match="vary over neighbor slots"):
        read_dp_checkpoint(path)
    ck = load_checkpoint(str(tmp_path / "bad.ckpt"))
    ck["json_file"]["model_type"] = "NEP"
    torch.save(ck, path)
    with pytest.raises(ValueError, match="not DP"):
        read_dp_checkpoint(path)


def _typedp_checkpoint(path, props=None, T=2, atom_types=(72, 8), seed=3):
    """A checkpoint saved from modules shaped like MatPL's TypeDP (dp_dp_typ_emb.py:66-97: one shared
    embedding net whose first layer takes [s, v(t_j)], per-type fitting nets)."""
    from matpl_amd.typedp import DEFAULT_PROPERTIES, TypeDPWeights, type_vectors
    props = list(props or DEFAULT_PROPERTIES)
    cfg = DPConfig(ntypes=T, m=100)
    rng = np.random.default_rng(seed)
    tv = type_vectors(list(atom_types), props)
    tw = TypeDPWeights.random(cfg, tv, seed=seed, davg=rng.normal(0.1, 0.02, (T, 4)), dstd=rng.uniform(0.5, 1.0, (T, 4)))
    model = _DP([tw.emb], tw.fit).double()
    desc = {"Rmax": cfg.rcut, "Rmin": cfg.rcut_smooth, "M2": cfg.M2, "network_size": cfg.emb_sizes}
    desc["type_embedding"] = {"physical_property": props}
    torch.save({"json_file": {"model_type": "DP", "atom_type": list(atom_types), "max_neigh_num": cfg.m,
                              "model": {"descriptor": desc, "fitting_net": {"network_size": cfg.fit_sizes}}},
                "epoch": 2, "state_dict": model.state_dict(), "davg": DPWeights(cfg, [], [], tw.davg, tw.dstd).davg_full(),
                "dstd": DPWeights(cfg, [], [], tw.davg, tw.dstd).dstd_full(), "energy_shift": [0.0] * T,
                "atom_type_order": np.array(atom_types), "sij_max": 1.0}, path)
    return tw


def test_reads_typedp_checkpoint(tmp_path):
    """A default-config TypeDP checkpoint (five physical properties, nn_feature_type.py:35) is read: the
    shared net and the fitting nets as saved, type vectors from the element table, and the folded
    per-pair model gives the same energy on the oracle as folding the saved weights directly."""
    from matpl_amd.typedp import DEFAULT_PROPERTIES, fold
    path = str(tmp_path / "typedp.ckpt")
    tw = _typedp_checkpoint(path)
    w, order, info = read_dp_checkpoint(path)
    rt = info["typedp"]
    assert order == [72, 8] and info["physical_property"] == list(DEFAULT_PROPERTIES)
    assert rt.emb[0][0].shape == (1 + len(DEFAULT_PROPERTIES), 25)
    for (Wa, ba, _), (Wb, bb, _) in zip(rt.emb, tw.emb):
        assert np.array_equal(Wa, Wb) and np.array_equal(ba, bb)
    assert np.array_equal(rt.type_vec, tw.type_vec)
    f = st.fluorite_hfo2((1, 1, 1), seed=3)
    a, b = oracle.dp_evaluate(f, fold(tw)), oracle.dp_evaluate(f, w)
    assert a["E"] == b["E"] and np.array_equal(a["F"], b["F"])
