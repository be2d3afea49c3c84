"""HIP-graph-captured MD blocks (10 velocity-Verlet steps incl. the neighbor-list rebuild) give the
same trajectory as eager launches, replay after replay, with eager work and synchronizations
between the replays."""
import numpy as np
import pytest
import torch

from matpl_amd import structures as st
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.md import DPMD

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(dev):
    from matpl_amd.neighbor import calc_stat, find_neighbore_gpu
    f = st.fluorite_hfo2((5, 5, 5), a=5.08, sigma=0.1, seed=41)
    cfg = DPConfig(ntypes=2, m=120)
    nl = find_neighbore_gpu(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut, device=dev)
    davg, dstd = calc_stat(nl.list, nl.dR, torch.as_tensor(f.types, device=dev), 2, cfg.rcut_smooth, cfg.rcut)
    w = DPWeights.random(cfg, seed=7, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())
    vel = st.maxwell_velocities(f, 600.0, seed=8)
    eager = DPMD(f, w, vel, precision="bf16", skin=0.5, mnl=160, device=dev)
    graph = DPMD(f, w, vel, precision="bf16", skin=0.5, mnl=160, device=dev)
    for md in (eager, graph):
        for _ in range(10):
            md._step()
    graph.capture()
    for rep in range(3):
        for _ in range(10):
            eager._step()
        torch.cuda.synchronize()
        assert graph.step() == 10
        torch.cuda.synchronize()
        eager.check(); graph.check()
        assert graph.nstep == eager.nstep
        assert torch.equal(graph.pos, eager.pos), rep
        assert torch.equal(graph.vel, eager.vel), rep
        assert torch.equal(graph.F, eager.F), rep
        assert torch.equal(graph.Ei, eager.Ei), rep
        assert torch.allclose(graph.V, eager.V, rtol=1e-12, atol=1e-9), rep
        assert torch.equal(graph.nl.list, eager.nl.list), rep
