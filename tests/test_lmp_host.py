"""The LAMMPS-side packing restatement (oracle/lammps.py) against the periodic DP neighbor oracle:
a frame given as local + ghost atoms with a full list must give, per centre atom and neighbor type,
the same neighbors (by owner) at the same displacements as the periodic search of NeighConst.f90."""
import numpy as np
import pytest

import oracle
import oracle.lammps as ol
from matpl_amd import structures as st


@pytest.mark.parametrize("name", ["cu", "hfo2"])
def test_lammps_packing_matches_periodic_search(name):
    if name == "cu":
        f, T, m, t2m = st.fcc_cu(), 1, 100, [0]
    else:
        f, T, m, t2m = st.fluorite_hfo2((2, 2, 2), seed=7), 2, 120, [1, 0]   # LAMMPS 1 = O, 2 = Hf
    rc, skin = 6.0, 1.0
    x, owner = ol.ghost_system(f.lattice, f.pos, rc + skin)
    model_t = f.types.astype(np.int32)
    lmp_type = np.array([t2m.index(t) + 1 for t in model_t], np.int32)[owner]
    il, nn, fn, ng = ol.full_list(x, f.natoms, rc + skin)
    it, lst, dR, cnt, rmin = ol.generate_neighdata(x, lmp_type, t2m, il, nn, fn, ng, T, m, rc)
    assert np.array_equal(it, model_t)
    plst, pdR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, T, m, rc)
    plst = plst.reshape(f.natoms, T, m)
    pdR = pdR.reshape(f.natoms, T, m, 3)
    lst = lst.reshape(f.natoms, T, m)
    dR = dR.reshape(f.natoms, T, m, 4)
    for i in range(f.natoms):
        for t in range(T):
            n = int(cnt[i, t])
            assert n == int((plst[i, t] > 0).sum())
            got = sorted(zip(owner[lst[i, t, :n] - 1].tolist(), np.round(dR[i, t, :n, 1:], 9).tolist()))
            want = sorted(zip((plst[i, t, :n] - 1).tolist(), np.round(pdR[i, t, :n], 9).tolist()))
            assert got == want
            assert np.allclose(dR[i, t, :n, 0], np.linalg.norm(dR[i, t, :n, 1:], axis=1), rtol=1e-15)
    d = dR[lst > 0][:, 1:]
    assert rmin == float((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).min())
    assert rmin > 0.61
