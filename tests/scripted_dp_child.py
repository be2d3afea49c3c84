"""Child process of test_gpu_torchscript.py: a scripted DP model with only the C++ operator library
loaded (no Python CalcOps_cuda registration), saved, reloaded with torch.jit.load and evaluated on a
LAMMPS-packed local + ghost system; prints one JSON line of its errors against the FP64 oracle."""
import json
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import oracle                                                            # noqa: E402
import oracle.lammps as ol                                               # noqa: E402
from matpl_amd import dp_script, structures as st                        # noqa: E402
from matpl_amd.dpnet import DPConfig, DPWeights                          # noqa: E402
from matpl_amd.neighbor import calc_stat, find_neighbore_gpu, generate_neighdata   # noqa: E402


def main(path):
    dev = torch.device("cuda:0")
    dp_script.load_ops()
    co = sys.modules.get("matpl_amd.calcops")
    assert co is None or not co._REGISTERED, "Python ops registered"
    RC, SKIN = 6.0, 1.0
    f, T, m, t2m = st.fluorite_hfo2((2, 2, 2), seed=7), 2, 120, [1, 0]
    x, owner = ol.ghost_system(f.lattice, f.pos, RC + SKIN)
    lmp_type = np.array([t2m.index(t) + 1 for t in f.types], np.int32)[owner]
    il, nn, fn, ng = ol.full_list(x, f.natoms, RC + SKIN)
    cfg = DPConfig(ntypes=T, m=m)
    tm = torch.as_tensor(f.types, device=dev)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, T, m, cfg.rcut, device=dev)
    davg, dstd = calc_stat(nl.list, nl.dR, tm, T, cfg.rcut_smooth, cfg.rcut)
    w = DPWeights.random(cfg, seed=11, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())
    ref = oracle.dp_evaluate(f, w)

    t = lambda a: torch.as_tensor(a, device=dev)
    nd = generate_neighdata(t(x), t(lmp_type), t2m, t(il), t(nn), t(fn), t(ng), T, m, RC)
    torch.jit.script(dp_script.DPScript(w, torch.float64, "cuda")).save(path)
    model = torch.jit.load(path, map_location=dev)
    N, nghost = f.natoms, len(x) - f.natoms
    atom_type = torch.tensor([72, 8], dtype=torch.int32)
    Etot, Ei, F, Eg, V = model(nd.neighbor_list, nd.imagetype_map, atom_type, nd.dR_neigh, nghost)
    assert Eg is None and tuple(F.shape) == (1, N + nghost, 3) and tuple(V.shape) == (1, 9)
    Fn = F[0].detach().cpu().numpy()
    Ff = Fn[:N].copy()
    np.add.at(Ff, owner[N:], Fn[N:])
    print(json.dumps({
        "dE": abs(Etot.item() - ref["E"]) / abs(ref["E"]),
        "dEi": float(np.abs(Ei.detach().cpu().numpy().reshape(-1) - ref["Ei"].reshape(-1)).max()),
        "dF": float(np.abs(Ff - ref["F"]).max()),
        "dV": float(np.abs(V.detach().cpu().numpy().reshape(-1) - np.asarray(ref["V"]).reshape(-1)).max()),
    }), flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
