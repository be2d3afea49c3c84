"""GPU parity of the LAMMPS-side DP packing (PairMATPL::generate_neighdata on the device) against its
CPU restatement, bit for bit, and of a DP evaluation on local + ghost atoms (forces on N + nghost
rows, folded back onto their owners as LAMMPS's reverse communication does) against the periodic
FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
import oracle.lammps as ol
from matpl_amd import structures as st
from matpl_amd.dp import DPTorch
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.neighbor import calc_stat, find_neighbore_gpu, generate_neighdata

pytestmark = pytest.mark.gpu

RC, SKIN = 6.0, 1.0


def _system(name, seed=0):
    if name == "cu":
        f, T, m, t2m = st.fcc_cu(), 1, 100, [0]
    else:
        f, T, m, t2m = st.fluorite_hfo2((2, 2, 2), seed=7), 2, 120, [1, 0]   # LAMMPS type 1 = O, 2 = Hf
    x, owner = ol.ghost_system(f.lattice, f.pos, RC + SKIN)
    lmp_type = np.array([t2m.index(t) + 1 for t in f.types], np.int32)[owner]
    il, nn, fn, ng = ol.full_list(x, f.natoms, RC + SKIN)
    # LAMMPS lists come in bin order, not index order: permute each atom's neighbors
    rng = np.random.default_rng(seed)
    for i in range(f.natoms):
        seg = slice(fn[i], fn[i] + nn[i])
        ng[seg] = rng.permutation(ng[seg])
    return f, T, m, t2m, x, owner, lmp_type, (il, nn, fn, ng)


def _gpu(dev, x, lmp_type, csr, t2m, T, m, rc=RC, check=True):
    il, nn, fn, ng = csr
    t = lambda a: torch.as_tensor(a, device=dev)
    return generate_neighdata(t(x), t(lmp_type), t2m, t(il), t(nn), t(fn), t(ng), T, m, rc, check=check)


@pytest.mark.parametrize("name", ["cu", "hfo2"])
def test_neighdata_bit_exact(dev, name):
    f, T, m, t2m, x, owner, lmp_type, csr = _system(name)
    print(f"case {name}: nlocal={f.natoms} nall={len(x)} entries={len(csr[3])}", flush=True)
    nd = _gpu(dev, x, lmp_type, csr, t2m, T, m)
    it, lst, dR, cnt, rmin = ol.generate_neighdata(x, lmp_type, t2m, *csr, T, m, RC)
    assert np.array_equal(nd.imagetype_map.cpu().numpy(), it)
    assert np.array_equal(nd.neighbor_list[0].cpu().numpy(), lst)
    assert np.array_equal(nd.dR_neigh[0].cpu().numpy(), dR)
    assert np.array_equal(nd.counts.cpu().numpy(), cnt)
    assert nd.min_rsq == rmin


def test_neighdata_errors(dev):
    f, T, m, t2m, x, owner, lmp_type, csr = _system("cu")
    print("case errors: overflow, bad index, bad type, too close", flush=True)
    with pytest.raises(OverflowError):
        _gpu(dev, x, lmp_type, csr, t2m, T, 40)
    nd = _gpu(dev, x, lmp_type, csr, t2m, T, 40, check=False)     # no check: buckets clipped at m
    assert int(nd.counts.max()) > 40 and int((nd.neighbor_list > 0).sum(-1).max()) == 40
    il, nn, fn, ng = csr
    bad = ng.copy()
    bad[3] = len(x)
    with pytest.raises(ValueError):
        _gpu(dev, x, lmp_type, (il, nn, fn, bad), t2m, T, m)
    fbad = fn.copy()
    fbad[5] = len(ng)            # a range past the end of neigh
    with pytest.raises(ValueError):
        _gpu(dev, x, lmp_type, (il, nn, fbad, ng), t2m, T, m)
    with pytest.raises(ValueError):
        _gpu(dev, x, lmp_type, csr, [-1], T, m)
    close = x.copy()
    close[1] = close[0] + np.array([0.5, 0.3, 0.0])
    il2, nn2, fn2, ng2 = ol.full_list(close, f.natoms, RC + SKIN)
    with pytest.raises(RuntimeError, match="too close"):
        _gpu(dev, close, lmp_type, (il2, nn2, fn2, ng2), t2m, T, m)


@pytest.mark.parametrize("name", ["cu", "hfo2"])
def test_dp_with_ghost_atoms_matches_periodic_oracle(dev, name):
    f, T, m, t2m, x, owner, lmp_type, csr = _system(name, seed=3)
    print(f"case ghost DP {name}", flush=True)
    cfg = DPConfig(ntypes=T, m=m)
    tm = torch.as_tensor(f.types, device=dev)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, T, m, cfg.rcut, device=dev)
    davg, dstd = calc_stat(nl.list, nl.dR, tm, T, cfg.rcut_smooth, cfg.rcut)
    w = DPWeights.random(cfg, seed=11, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())
    ref = oracle.dp_evaluate(f, w)
    nd = _gpu(dev, x, lmp_type, csr, t2m, T, m)
    N, nghost = f.natoms, len(x) - f.natoms
    E, Ei, F, V = DPTorch(w, torch.float64, dev)(nd.neighbor_list, nd.imagetype_map, nd.dR_neigh, nghost=nghost)
    F = F[0].detach().cpu().numpy()
    assert F.shape == (N + nghost, 3)
    Ff = F[:N].copy()
    np.add.at(Ff, owner[N:], F[N:])
    assert abs(E.item() - ref["E"]) < 1e-10 * abs(ref["E"])
    assert np.abs(Ei.detach().cpu().numpy().reshape(-1) - ref["Ei"].reshape(-1)).max() < 1e-10
    assert np.abs(Ff - ref["F"]).max() < 1e-9
    assert np.abs(V.detach().cpu().numpy().reshape(-1) - np.asarray(ref["V"]).reshape(-1)).max() < 1e-8
