"""Trajectory-level checks of the MD driver (matpl_amd.md.DPMD: fused E/F/V, skin list rebuilt every
10 steps, velocity-Verlet kicks on the GPU), the caller the reference's LAMMPS pair style feeds
(src/lmps/lammps-libtorch/MATPL/pair_matpl.cpp:864-1000).

  * Against an FP64 velocity-Verlet loop on the CPU whose forces are the oracle's
    (oracle.dp_evaluate): positions and velocities after 20 steps of 1 fs. The unit of the
    acceleration is derived here from the CODATA amu and eV (not the kernel's constant), so a wrong
    unit, a kick with a stale force or a drift before the first half kick fails.
  * Energy conservation: the total energy sum(Ei) + kinetic energy (same CODATA unit) of a 2,592-atom
    HfO2 box over 100 steps: the drift between the first and last ten steps and the largest excursion
    from the start, per atom, within the stated bounds for FP32 and BF16 nets.
  * The BF16 force gate (force RMSE vs the FP64 path <= 1e-3 eV/A) at the end of the 100 steps as well
    as at the start.
"""
import numpy as np
import pytest
import torch

import oracle
from matpl_amd import structures as st
from matpl_amd.dpnet import DPConfig, DPWeights
from matpl_amd.md import DPMD, DPMDFP64

pytestmark = pytest.mark.gpu

# amu (A/fs)^2 in eV from CODATA 2018: 1.66053906660e-27 kg * 1e10 m^2/s^2 / 1.602176634e-19 J/eV
MV2_EV = 1.66053906660e-27 * 1e10 / 1.602176634e-19


def _model(f, seed=7):
    cfg = DPConfig(ntypes=2, m=120)
    lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, 2, cfg.m, cfg.rcut)
    davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, 2, cfg.rcut_smooth, cfg.rcut)
    return DPWeights.random(cfg, seed=seed, davg=davg, dstd=dstd)


def _masses(f):
    return np.array([st.MASS[f.elements[t]] for t in f.types])


def _min_image(d, L):
    fr = d @ np.linalg.inv(L)
    return (fr - np.round(fr)) @ L


def test_trajectory_matches_fp64_oracle_integrator(dev):
    f = st.fluorite_hfo2((3, 3, 3), seed=51)
    w = _model(f)
    vel = st.maxwell_velocities(f, 300.0, seed=52)
    dt, nsteps = 1.0, 20
    print(f"case MD trajectory vs oracle integrator: {f.natoms} atoms, {nsteps} steps", flush=True)
    md = DPMD(f, w, vel, precision="fp32", skin=0.5, mnl=160, dt=dt, rebuild_every=10, device=dev)
    for _ in range(nsteps):
        md._step()
    torch.cuda.synchronize()
    md.check()
    # FP64 velocity Verlet with the oracle's forces
    acc = 1.0 / (_masses(f) * MV2_EV)[:, None]
    x, v = f.pos.copy(), vel.copy()
    F = oracle.dp_evaluate(f, w)["F"]
    for _ in range(nsteps):
        v = v + 0.5 * dt * F * acc
        x = x + dt * v
        F = oracle.dp_evaluate(st.Frame(f.lattice, x, f.types, f.elements), w)["F"]
        v = v + 0.5 * dt * F * acc
    dx = np.abs(_min_image(md.pos.cpu().numpy() - x, f.lattice)).max()
    dv = np.abs(md.vel.cpu().numpy() - v).max()
    dF = np.abs(md.F.double().cpu().numpy() - F).max()
    print(f"max |dx| {dx:.2e} A, max |dv| {dv:.2e} A/fs, max |dF| {dF:.2e} eV/A "
          f"(displacement over the run {np.abs(x - f.pos).max():.3f} A)", flush=True)
    assert dx < 1e-5 and dv < 1e-6 and dF < 1e-4


def _total_energy(md, mass):
    ke = 0.5 * float((torch.as_tensor(mass, device=md.vel.device)[:, None] * md.vel ** 2).sum().item()) * MV2_EV
    return md.energy() + ke


# largest |E(t) - E(0)| and |<E>_last10 - <E>_first10| over 100 steps, eV per atom
NVE_BOUNDS = {"fp32": (2e-4, 5e-5), "bf16": (2e-3, 1e-3)}


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_nve_energy_conservation_and_end_gate(dev, precision):
    f = st.fluorite_hfo2((6, 6, 6), seed=61)
    w = _model(f, seed=9)
    vel = st.maxwell_velocities(f, 300.0, seed=62)
    mass = _masses(f)
    print(f"case NVE {precision}: {f.natoms} atoms, 100 steps", flush=True)
    md = DPMD(f, w, vel, precision=precision, skin=0.5, mnl=160, dt=1.0, rebuild_every=10, device=dev)
    E = [_total_energy(md, mass)]
    for _ in range(100):
        md._step()
        E.append(_total_energy(md, mass))
    md.check()
    E = np.asarray(E) / f.natoms
    excursion = float(np.abs(E - E[0]).max())
    drift = float(abs(E[-10:].mean() - E[:10].mean()))
    print(f"{precision}: max |E(t)-E(0)| {excursion:.2e} eV/atom, drift {drift:.2e} eV/atom, "
          f"E(0) {E[0]:.6f} eV/atom", flush=True)
    ex_bound, drift_bound = NVE_BOUNDS[precision]
    assert excursion < ex_bound and drift < drift_bound
    if precision == "bf16":
        # the accuracy gate at the end of the trajectory, against the FP64 path at the same positions
        fe = st.Frame(f.lattice, md.pos.cpu().numpy(), f.types, f.elements)
        ref = DPMDFP64(fe, w, md.vel.cpu().numpy(), device=dev)
        rmse = float(torch.sqrt(((md.F.double() - ref.F64) ** 2).mean()).item())
        print(f"bf16 force RMSE vs FP64 after 100 steps: {rmse:.2e} eV/A", flush=True)
        assert rmse <= 1e-3
