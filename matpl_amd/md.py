"""Velocity-Verlet MD on the GPU with the DP force field.

``DPMD`` drives the fused path: neighbor list with a skin rebuilt every
``rebuild_every`` steps (cell list, positions re-wrapped into the cell at each
rebuild so the stored image shifts stay valid), fused E/F/V every step.

``DPMDFP64`` is the plain drop-in path used as the baseline: every step it
rebuilds the reference-layout neighbor list at the model cutoff (bucket size
m, like find_neighbore / PairMATPL::generate_neighdata), forms ImageDR, and
evaluates the reference-shaped FP64 model (DPTorch: torch FP64 nets,
autograd dE/dRi, HIP FP64 calculateForce / calculateVirial).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .calcops import _ptr, _stream
from .dp import DPTorch
from .dp_fused import FusedDP
from .dpnet import DPWeights
from .neighbor import find_neighbore_gpu, image_dr, reverse_map
from .structures import MASS, Frame


def _affine(x: torch.Tensor, M: np.ndarray, out=None) -> torch.Tensor:
    out = torch.empty_like(x) if out is None else out
    Mh = np.ascontiguousarray(M, np.float64).reshape(9)
    _lib.call("matpl_md_affine", _ptr(x), _ptr(out), Mh.ctypes.data, x.shape[0], _stream())
    return out


class DPMD:
    def __init__(self, frame: Frame, w: DPWeights, velocities: np.ndarray, precision="bf16", skin=0.5,
                 mnl=None, dt=1.0, rebuild_every=10, device="cuda", compress=None, store=False):
        """compress: a matpl_amd.compress.CompressTable for the tabulated embedding.
        store: FusedDP store mode (G and dG/ds kept from the forward for the backward)."""
        self.dev = torch.device(device)
        self.cfg = w.cfg
        self.N = frame.natoms
        self.L = np.asarray(frame.lattice, np.float64)
        self.Linv = np.linalg.inv(self.L)
        self.L_dev = torch.as_tensor(self.L.reshape(1, 9), device=self.dev)
        self.types = torch.as_tensor(frame.types, dtype=torch.int32, device=self.dev)
        self.pos = torch.as_tensor(frame.pos, dtype=torch.float64, device=self.dev).contiguous()
        self.vel = torch.as_tensor(velocities, dtype=torch.float64, device=self.dev).contiguous()
        m = np.array([MASS[frame.elements[t]] for t in frame.types])
        self.inv_mass = torch.as_tensor(1.0 / m, dtype=torch.float64, device=self.dev)
        self.rc_nl = self.cfg.rcut + skin
        self.mnl = mnl or self.cfg.m
        self.dt = dt
        self.rebuild_every = rebuild_every
        self.model = FusedDP(w, precision, self.dev, compress=compress, store=store)
        tm = frame.types
        self.type_ranges = [(int(np.searchsorted(tm, t, "left")), int(np.searchsorted(tm, t, "right")))
                            for t in range(self.cfg.ntypes)]
        self.nstep = 0
        self.rebuild()
        self.compute()

    def rebuild(self):
        """Neighbor list at rc + skin from the current positions; positions re-wrapped into the cell.
        After the first call every result goes into the same buffers (stream-capturable)."""
        frac = _affine(self.pos, self.Linv)
        first = not hasattr(self, "nl")
        self.nl = find_neighbore_gpu(self.L, frac, self.types, self.cfg.ntypes, self.mnl, self.rc_nl,
                                     device=self.dev, want_dR=False, want_shift=True, check=False,
                                     out=None if first else self.nl)
        _affine(self.nl.frac_wrapped, self.L, out=self.pos)
        self.rev, self.rev_status = reverse_map(self.nl, self.types, self.mnl,
                                                out=None if first else (self.rev, self.rev_status))

    def compute(self):
        self.Ei, self.F, self.V, self.status = self.model(self.pos, self.L_dev, self.types, self.nl.list,
                                                          self.nl.shift, self.nl.counts, self.mnl,
                                                          type_ranges=self.type_ranges, rev=self.rev)

    def _step(self):
        _lib.call("matpl_md_kick", _ptr(self.vel), _ptr(self.pos), _ptr(self.F), _ptr(self.inv_mass), self.dt,
                  self.N, 1, _stream())
        self.nstep += 1
        if self.nstep % self.rebuild_every == 0:
            self.rebuild()
        self.compute()
        _lib.call("matpl_md_kick", _ptr(self.vel), None, _ptr(self.F), _ptr(self.inv_mass), self.dt, self.N, 0,
                  _stream())

    def step(self):
        """One velocity-Verlet step; with a captured graph and at a block boundary, one replay
        advances ``rebuild_every`` steps at once (returns the number of steps taken)."""
        if self.graph is not None and self.nstep % self.rebuild_every == 0:
            self.graph.replay()
            self.nstep += self.rebuild_every
            return self.rebuild_every
        self._step()
        return 1

    graph = None

    def capture(self):
        """Captures ``rebuild_every`` MD steps (the last one with the neighbor-list rebuild) in one
        HIP graph. State that crosses replays (positions, velocities, neighbor list, reverse map,
        forces, energies, virial, status flags) lives in fixed buffers. Must be called at a block
        boundary (nstep % rebuild_every == 0) after at least one eager step."""
        if self.nstep % self.rebuild_every != 0:
            raise RuntimeError("capture at a neighbor-list rebuild boundary")
        static = dict(F=self.F, Ei=self.Ei, V=self.V, status=self.status)
        nstep0 = self.nstep
        s = torch.cuda.Stream(device=self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for _ in range(self.rebuild_every):
                    self._step()
                for k, t in static.items():
                    t.copy_(getattr(self, k))
        torch.cuda.current_stream(self.dev).wait_stream(s)
        for k, t in static.items():
            setattr(self, k, t)
        # capture records work without running it: state is unchanged
        self.nstep = nstep0
        self.graph = g

    def check(self):
        """Raises if a neighbor bucket overflowed (syncs the stream)."""
        self.nl.check()
        if int(self.rev_status.item()) != 0:
            raise RuntimeError("neighbor list is not symmetric; the force gather needs a reverse map")
        if int(self.status.item()) != 0:
            raise OverflowError("more than maxNeighborNum neighbors of one type within rcut")

    def energy(self) -> float:
        return float(self.Ei.double().sum().item())


class DPMDFP64:
    """The plain FP64 drop-in path, one MD step = NL + ImageDR + DPTorch + kicks."""

    def __init__(self, frame: Frame, w: DPWeights, velocities: np.ndarray, dt=1.0, device="cuda"):
        self.dev = torch.device(device)
        self.cfg = w.cfg
        self.N = frame.natoms
        self.L = np.asarray(frame.lattice, np.float64)
        self.Linv = np.linalg.inv(self.L)
        self.types = torch.as_tensor(frame.types, dtype=torch.int32, device=self.dev)
        self.pos = torch.as_tensor(frame.pos, dtype=torch.float64, device=self.dev).contiguous()
        self.vel = torch.as_tensor(velocities, dtype=torch.float64, device=self.dev).contiguous()
        m = np.array([MASS[frame.elements[t]] for t in frame.types])
        self.inv_mass = torch.as_tensor(1.0 / m, dtype=torch.float64, device=self.dev)
        self.dt = dt
        self.model = DPTorch(w, torch.float64, self.dev)
        self.compute()

    def compute(self):
        frac = _affine(self.pos, self.Linv)
        nl = find_neighbore_gpu(self.L, frac, self.types, self.cfg.ntypes, self.cfg.m, self.cfg.rcut,
                                device=self.dev, want_dR=True, check=False)
        idr = image_dr(nl.list, nl.dR)
        Etot, Ei, F, V = self.model(nl.list.unsqueeze(0), self.types, idr.unsqueeze(0))
        self.E = Etot.detach()
        self.F64 = F[0].detach()
        self.F = self.F64.float().contiguous()
        self.V = V.detach()

    def step(self):
        _lib.call("matpl_md_kick", _ptr(self.vel), _ptr(self.pos), _ptr(self.F), _ptr(self.inv_mass), self.dt,
                  self.N, 1, _stream())
        self.compute()
        _lib.call("matpl_md_kick", _ptr(self.vel), None, _ptr(self.F), _ptr(self.inv_mass), self.dt, self.N, 0,
                  _stream())
