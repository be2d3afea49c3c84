"""Benchmark: DP se_e2_a MD on fluorite HfO2, 19x19x23 cells = 99,636 atoms, one GPU.

One step = one velocity-Verlet MD step with energy + forces + virial
(neighbor list with a 0.5 A skin rebuilt every 10 steps, amortized inside
the timed region). Synthetic seeded structure and random-init weights of the
reference DP architecture (embedding 1-25-25-25, M2 16, fitting 400-50-50-50-1,
rcut 6.0 / smooth 5.5 A, 120 neighbors per type).

Prints one JSON line: the fused BF16/FP32 path's atom-steps/s (value), the
plain FP64 drop-in path measured on the same GPU, the force RMSE of the
fused path vs FP64 on this workload, a roofline entry for the dominant
kernel, and a CPU baseline (the FP64 C oracle, OpenMP) on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (energy+force+virial) vs plain FP64 path, at force RMSE ≤1e-3 eV/Å"
REPS = (19, 19, 23)
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per launch of dp_fused_bwd_kernel on the 19x19x23 workload: rocprofv3 FETCH_SIZE + WRITE_SIZE,
# each scaled by a copy kernel of known size (profiles/dp_md_hfo2_99636_hbm_traffic.md; re-measure after
# kernel changes). Counters cannot be read from inside the timed run. The XCD-aware block order (the
# autotuner's choice here) keeps partner-slot writes within one XCD's L2; the plain order writes more.
BWD_TRAFFIC_BYTES = {"xcd": 459.1e6, "plain": 587.1e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(reps=REPS, device="cuda"):
    import torch
    from matpl_amd import structures as st
    from matpl_amd.dpnet import DPConfig, DPWeights
    from matpl_amd.neighbor import calc_stat, find_neighbore_gpu
    frame = st.fluorite_hfo2(reps, a=5.08, sigma=0.1, seed=1003)
    cfg = DPConfig(ntypes=2, rcut=6.0, rcut_smooth=5.5, m=120)
    nl = find_neighbore_gpu(frame.lattice, frame.frac(), frame.types, 2, cfg.m, cfg.rcut, device=device)
    davg, dstd = calc_stat(nl.list, nl.dR, torch.as_tensor(frame.types, device=device), 2, cfg.rcut_smooth,
                           cfg.rcut)
    w = DPWeights.random(cfg, seed=2023, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())
    vel = st.maxwell_velocities(frame, 300.0, seed=2003)
    del nl
    return frame, w, vel


def flops_per_neighbor_bwd(E=25):
    """Algorithmic flops of the backward kernel per in-cutoff neighbor: the recomputed embedding forward
    (E + 2E^2 MACs) and the A contraction (4E), then dG (4E), dRi (4E), the two E x E layer backprops (2E^2)
    and the layer-1 input gradient (E); 2 flops per MAC (5,700 for E = 25)."""
    fwd = E + 2 * E * E + 4 * E          # embedding MACs (1->E, E->E, E->E) + contraction
    bwd = 4 * E + 2 * E * E + E + 4 * E  # dG, dRi, two E->E backprops, layer-1 input grad
    return 2 * (fwd + bwd)


def _cpu_info():
    """CPU model and counts of the host the baseline runs on (lscpu's fields, read from /proc/cpuinfo)."""
    model, phys = None, set()
    pid = cid = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    cid = v
                elif not k and pid is not None:
                    phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    if pid is not None:
        phys.add((pid, cid))
    return {"cpu_model": model, "host_physical_cores": len(phys) or None, "host_logical_cpus": os.cpu_count(),
            "cpus_allowed": len(os.sched_getaffinity(0))}


def _time_oracle(frame, w, seconds):
    import oracle
    t0 = time.perf_counter()
    nev = 0
    while True:
        oracle.dp_evaluate(frame, w)
        nev += 1
        if time.perf_counter() - t0 > seconds:
            break
    return frame.natoms * nev / (time.perf_counter() - t0), nev


def cpu_baseline():
    """The FP64 C oracle (a restatement of the reference's CPU path: NeighConst.f90 search, dp_dp.py
    env mat, nets and scatter; OpenMP over atoms) timed for full E+F+V evaluations, neighbor search
    included, on bounded samples of two workloads: HfO2 8x8x8 cells (6,144 atoms; a sample of the
    99,636-atom MD workload, same model) and fcc Cu 3x3x3 (108 atoms), each with all the threads the
    process may use and with one thread. About 20 s of CPU work in all."""
    import oracle
    from matpl_amd import structures as st
    from matpl_amd.dpnet import DPConfig, DPWeights
    L = oracle.lib()
    nthreads = L.oracle_num_threads()
    out = {"unit": "atom-steps/s", "kind": "port", **_cpu_info()}
    samples = {}
    for name, f, cfg in (("hfo2_6144", st.fluorite_hfo2((8, 8, 8), a=5.08, sigma=0.1, seed=1003),
                          DPConfig(ntypes=2, rcut=6.0, rcut_smooth=5.5, m=120)),
                         ("cu_108", st.fcc_cu((3, 3, 3), seed=1000), DPConfig(ntypes=1, rcut=6.0, rcut_smooth=5.5,
                                                                                m=100))):
        T = cfg.ntypes
        lst, dR = oracle.dp_find_neighbor(f.lattice, f.frac(), f.types, T, cfg.m, cfg.rcut)
        davg, dstd = oracle.dp_calc_stat(lst, dR, f.types, T, cfg.rcut_smooth, cfg.rcut)
        w = DPWeights.random(cfg, seed=2023, davg=davg, dstd=dstd)
        r = {"natoms": f.natoms}
        for key, nt, secs in (("all_threads", nthreads, 6.0), ("one_thread", 1, 4.0)):
            L.oracle_set_num_threads(nt)
            v, nev = _time_oracle(f, w, secs)
            r[key] = {"value": round(v, 1), "threads": nt, "evaluations": nev}
        L.oracle_set_num_threads(nthreads)
        samples[name] = r
    h = samples["hfo2_6144"]
    out.update({"value": h["all_threads"]["value"], "cores": nthreads,
                "sample": "fluorite HfO2 8x8x8 cells (6144 atoms), the MD workload's model, full E+F+V "
                          "evaluations incl. neighbor search, FP64 C oracle with OpenMP; one-thread and "
                          "fcc Cu 108-atom (the CPU-runnable DP configuration) figures under 'samples'",
                "samples": samples})
    log(f"CPU baseline ({out['cpu_model']}, {nthreads} threads): HfO2 6144 {h['all_threads']['value']:.4g} / "
        f"1 thread {h['one_thread']['value']:.4g} atom-steps/s; Cu 108 "
        f"{samples['cu_108']['all_threads']['value']:.4g} / {samples['cu_108']['one_thread']['value']:.4g}")
    return out


def nep_leg(steps=20):
    """NEP E+F+V on zincblende SiC 6x6x6 (1,728 atoms), radial/angular cutoff 6/4 A, n_max 4/4,
    basis 12, l_max [4,2,0], 30 hidden units per type: one step = neighbor search + features +
    fitting net + dE/dRi + force/virial ops, in float32 and in float64 (the drop-in path)."""
    import torch
    from matpl_amd import nep as mn
    from matpl_amd import nep_ops
    from matpl_amd import structures as st
    from matpl_amd.nepnet import NEPConfig, NEPWeights
    dev = torch.device("cuda:0")
    f = st.zincblende_sic((6, 6, 6), a=4.36, sigma=0.08, seed=1001)
    cfg = NEPConfig(ntypes=2, l_max=[4, 2, 0])
    w = NEPWeights.random(cfg, seed=2021)
    gp = mn.pack_frames([f], cfg.rcut_radial, dev)
    NNr, NNa = nep_ops.calculate_maxneigh(
        gp["num_atoms"], gp["box"], gp["box_orig"], gp["num_cell"], gp["position"], cfg.rcut_radial,
        cfg.rcut_angular, 2, gp["atom_type_map"], False)
    maxr, maxa = int(NNr.max().item()), int(NNa.max().item())
    out = {"workload": f"NEP zincblende SiC 6x6x6 ({f.natoms} atoms), rc 6/4 A, n_max 4/4, basis 12, "
                       f"l_max [4,2,0], 30 hidden, E+F+V incl. neighbor search", "natoms": f.natoms}
    res = {}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        model = mn.NEPTorch(w, dt, dev)

        def step():
            nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, 2, maxr, maxa, ri_dtype=dt)
            return model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"], gp["num_atoms"],
                         gp["atom_type_map"], create_graph=False)
        for _ in range(5):
            r = step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = step()
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        res[name] = r
        out[name] = {"value": round(f.natoms * steps / dt_s, 1), "unit": "atom-steps/s",
                     "ms_per_step": round(dt_s * 1e3 / steps, 3)}
        # the same step captured as one HIP graph (no host synchronisation inside), replayed
        g = mn.NEPGraph(model, gp, 2, maxr, maxa)
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            rg = g.replay()
        torch.cuda.synchronize()
        dg = time.perf_counter() - t0
        g.check()
        dF = float((rg[2].detach() - r[2].detach()).abs().max().item())
        out[name + "_graph"] = {"value": round(f.natoms * steps / dg, 1), "unit": "atom-steps/s",
                                "ms_per_step": round(dg * 1e3 / steps, 3), "max_abs_dF_vs_eager": dF}
        del g
    F32, F64 = res["fp32"][2].detach().double(), res["fp64"][2].detach()
    out["force_rmse_fp32_vs_fp64_eV_per_A"] = float(torch.sqrt(((F32 - F64) ** 2).mean()).item())
    out["speedup_fp32_vs_fp64"] = round(out["fp32"]["value"] / out["fp64"]["value"], 2)
    log(f"NEP: fp32 {out['fp32']['ms_per_step']} ms/step, fp64 {out['fp64']['ms_per_step']} ms/step, "
        f"HIP graph fp32 {out['fp32_graph']['ms_per_step']} / fp64 {out['fp64_graph']['ms_per_step']} ms/step, "
        f"force RMSE {out['force_rmse_fp32_vs_fp64_eV_per_A']:.2e}")
    return out


def nep_train_leg(steps=3):
    """NEP training gradient on 128 frames of a 4-element Ag/Au/Si/C alloy (N_f uniform in [32, 1024],
    0.06 atoms/A^3, 62,713 atoms in all), descriptor as the SiC leg with 4 types: one step = neighbor
    search + features + fitting nets + forces/virials (create_graph) + the backward of an
    energy + force + virial loss to the fitting-net and c2/c3 coefficient gradients, in float32 and
    float64."""
    import torch
    from matpl_amd import nep as mn
    from matpl_amd import structures as st
    from matpl_amd.nepnet import NEPConfig, NEPWeights
    dev = torch.device("cuda:0")
    frames = st.agausic_frames(nframes=128, nmin=32, nmax=1024, density=0.06, dmin=1.9, seed=1004)
    natoms = sum(f.natoms for f in frames)
    cfg = NEPConfig(ntypes=4, l_max=[4, 2, 0])
    w = NEPWeights.random(cfg, seed=2024)
    gp = mn.pack_frames(frames, cfg.rcut_radial, dev)
    out = {"workload": f"NEP training gradient, 128 Ag-Au-Si-C frames ({natoms} atoms), rc 6/4 A, n_max 4/4, "
                       "basis 12, l_max [4,2,0], 30 hidden, E+F+V loss backward to all parameters incl. c2/c3",
           "natoms": natoms}
    grads = {}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        model = mn.NEPTorch(w, dt, dev)
        params = [q for q in model.parameters() if q.requires_grad]

        def step():
            nb = mn.neighbors(gp, cfg.rcut_radial, cfg.rcut_angular, cfg.ntypes, ri_dtype=dt)
            E, Ei, F, V = model(nb["NN_r"], nb["NL_r"], nb["Ri_r"], nb["NN_a"], nb["NL_a"], nb["Ri_a"],
                                gp["num_atoms"], gp["atom_type_map"], create_graph=True)
            loss = (E ** 2).sum() + (F ** 2).sum() + (V ** 2).sum()
            return torch.autograd.grad(loss, params)
        grads[name] = [g.detach().double() for g in step()]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        out[name] = {"value": round(natoms * steps / dt_s, 1), "unit": "atom-steps/s",
                     "ms_per_step": round(dt_s * 1e3 / steps, 2)}
    out["speedup_fp32_vs_fp64"] = round(out["fp32"]["value"] / out["fp64"]["value"], 2)
    worst = max(float((a - b).norm() / b.norm()) for a, b in zip(grads["fp32"], grads["fp64"]) if b.norm() > 0)
    out["grad_rel_err_fp32_vs_fp64"] = float(f"{worst:.3g}")
    log(f"NEP train: fp32 {out['fp32']['ms_per_step']} ms/step, fp64 {out['fp64']['ms_per_step']} ms/step, "
        f"gradient rel. error fp32 vs fp64 {out['grad_rel_err_fp32_vs_fp64']}")
    return out


def _dp_train_workload(dev):
    """256 LGPS frames, 4 types, m = 150; weights with davg/dstd from the batch's own statistics."""
    from matpl_amd import structures as st
    from matpl_amd.dp_train import pack_batch
    from matpl_amd.dpnet import DPConfig, DPWeights
    from matpl_amd.neighbor import calc_stat
    frames = st.lgps_frames(nframes=256, nmin=64, nmax=512, seed=1002)
    cfg = DPConfig(ntypes=4, m=150)
    batch = pack_batch(frames, cfg, dev, seed=7, keep_full=True)
    davg, dstd = calc_stat(batch.list_full[0], batch.image_dr_full[0, ..., 1:], batch.type_map, 4, cfg.rcut_smooth,
                           cfg.rcut)
    w = DPWeights.random(cfg, seed=1002, davg=davg.cpu().numpy(), dstd=dstd.cpu().numpy())
    return frames, cfg, w, batch


def dp_train_fused_leg(steps=5, precisions=("bf16", "fp32")):
    """The DP training workload of dp_train_leg on the hand-written kernels only (DPTrainFused: fused
    forward and force kernels, force-loss double backward of the embedding nets; fitting nets in
    torch), gradient step + Adam, timed per precision; no FP64 path (for profiling the step alone)."""
    import torch
    from matpl_amd.dp_train_fused import DPTrainFused, pack_fused_batch
    dev = torch.device("cuda:0")
    frames, cfg, w, batch = _dp_train_workload(dev)
    fb = pack_fused_batch(frames, cfg, dev, seed=7)
    out = {"natoms": fb.natoms}
    for prec in precisions:
        fm = DPTrainFused(w, prec, dev)
        fparams = list(fm.parameters())
        fopt = torch.optim.Adam(fparams, lr=1e-4, fused=True)

        def fstep():
            loss, g = fm.loss_and_grads(fb)
            for q, gq in zip(fparams, g):
                q.grad = gq
            fopt.step()
        for _ in range(2):
            fstep()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fstep()
        torch.cuda.synchronize()
        out[prec] = {"ms_per_step": round((time.perf_counter() - t0) * 1e3 / steps, 2)}
        log(f"DP train fused {prec}: {out[prec]['ms_per_step']} ms/step")
    return out


def dp_train_leg(steps=5):
    """DP training step on 256 Li10GeP2S12-stoichiometry frames (N_f uniform in [64, 512], 2.0 g/cm^3,
    73.7k atoms in all), rc 6 A, m = 150 per type (S = 600 slots), nets as the MD workload: one step =
    env mat + embedding and fitting nets + dE/dRi (create_graph) + forces + the backward of an energy +
    force loss to every net parameter + an Adam update. Batched path (matpl_amd.dp_train.DPTrain: all
    frames packed, compact buckets, FP32 descriptors, BF16 or FP32 products) against the plain FP64
    path (DPTorch over all 600 slots, frames in chunks of 32 with gradients accumulated)."""
    import torch
    from matpl_amd.dp import DPTorch
    from matpl_amd.dp_train import DPTrain, plain_grads, train_grads
    dev = torch.device("cuda:0")
    frames, cfg, w, batch = _dp_train_workload(dev)
    N = batch.natoms
    out = {"workload": f"DP training step, 256 Li-Ge-P-S frames ({N} atoms), rc 6 A, m=150 per type (S=600), "
                       "E+F loss backward to all net parameters + Adam", "natoms": N, "mc": batch.mc}
    grads = {}

    def timed(name, fn, nsteps):
        # two untimed steps: the first steps of a precision run far longer (library set-up, the
        # optimizer's state, allocator growth; 1.1 s and 0.14 s against 66 ms for BF16)
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            r = fn()
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        out[name] = {"value": round(N * nsteps / dt_s, 1), "unit": "atom-steps/s",
                     "ms_per_step": round(dt_s * 1e3 / nsteps, 2)}
        return r

    for prec in ("bf16", "fp32"):
        model = DPTrain(w, prec, dev)
        params = [q for q in model.parameters() if q.requires_grad]
        opt = torch.optim.Adam(params, lr=1e-4, foreach=True)

        def step(model=model, params=params, opt=opt):
            loss, g = train_grads(model, batch)
            for q, gq in zip(params, g):
                q.grad = gq
            opt.step()
            return g
        # gradients at the initial weights, for the comparison below
        grads[prec] = [None if x is None else x.detach().clone() for x in train_grads(model, batch)[1]]
        timed(prec, step, steps)
    ref = DPTorch(w, torch.float64, dev, tall=True)
    rparams = [q for q in ref.parameters() if q.requires_grad]
    ropt = torch.optim.Adam(rparams, lr=1e-4, foreach=True)
    grads["fp64"] = [None if x is None else x.detach().clone() for x in plain_grads(ref, batch, chunk_frames=32)[1]]

    def rstep():
        loss, g = plain_grads(ref, batch, chunk_frames=32)
        for q, gq in zip(rparams, g):
            q.grad = gq
        ropt.step()
    timed("fp64_plain", rstep, 1)
    # the same step on the hand-written kernels (fused forward and force kernels, force-loss double
    # backward of the embedding nets on the matrix cores; fitting nets in torch)
    from matpl_amd.dp_train_fused import DPTrainFused, pack_fused_batch
    fb = pack_fused_batch(frames, cfg, dev, seed=7)
    nin = int(fb.nl_counts.sum().item())         # list entries at rc (= in-cutoff neighbors)
    for prec in ("bf16", "fp32"):
        fm = DPTrainFused(w, prec, dev)
        fparams = list(fm.parameters())
        fopt = torch.optim.Adam(fparams, lr=1e-4, fused=True)
        grads[prec + "_fused"] = [None if x is None else x.detach().clone() for x in fm.loss_and_grads(fb)[1]]

        def fstep(fm=fm, fparams=fparams, fopt=fopt):
            loss, g = fm.loss_and_grads(fb)
            for q, gq in zip(fparams, g):
                q.grad = gq
            fopt.step()
        timed(prec + "_fused", fstep, steps)
        del fm, fopt
    for prec in ("bf16", "fp32", "bf16_fused", "fp32_fused"):
        worst = 0.0
        for a, b in zip(grads[prec], grads["fp64"]):
            if a is not None and b is not None and b.norm() > 0:
                worst = max(worst, float((a.double() - b).norm() / b.norm()))
        out[f"grad_rel_err_{prec}_vs_fp64"] = float(f"{worst:.3g}")
        out[f"speedup_{prec}_vs_fp64_plain"] = round(out[prec]["value"] / out["fp64_plain"]["value"], 1)
    # path-level work of a training step: 3 F_DP per atom (forward, dE/dRi, double backward),
    # F_DP = 4 (1375 n + 26,650) flop with n the atom's in-cutoff neighbors
    flops = 3 * 4 * (1375 * nin + 26650 * N)
    for prec in ("bf16_fused", "fp32_fused"):
        tf = flops / (out[prec]["ms_per_step"] * 1e-3) / 1e12
        out[prec]["path_tflops"] = round(tf, 3)
        out[prec]["path_frac_of_bf16_peak"] = round(tf / BF16_PEAK_TFLOPS, 5)
    out["in_cutoff_neighbors"] = nin
    log(f"DP train: bf16 {out['bf16']['ms_per_step']} ms, fp32 {out['fp32']['ms_per_step']} ms, "
        f"fused bf16 {out['bf16_fused']['ms_per_step']} ms, fused fp32 {out['fp32_fused']['ms_per_step']} ms, "
        f"fp64 plain {out['fp64_plain']['ms_per_step']} ms per step (mc={batch.mc})")
    return out


def compress_table(w, dev, order=3, dx=0.01, s_max_raw=1.0):
    """The reference's DP compress defaults (ckpt_compress.py:24-26: dx 0.01, order 3); s tabulated up to
    the normalized s of a 1.0 A bond (the shortest HfO2 bond is above 1.9 A), then 10x coarser."""
    import torch
    from matpl_amd.compress import build_table
    from matpl_amd.dp import DPTorch
    plain = DPTorch(w, torch.float64, dev)
    ct = build_table(list(plain.emb), w.davg, w.dstd, s_max_raw, dx=dx, order=order, device=dev)
    del plain
    return ct


def compressed_md_leg(frame, w, vel, dev, ct, precision, steps, warmup, tune=True):
    """The same MD workload with the embedding evaluated from the compression table (the reference's
    DP compress) inside the fused kernels: G and dG/ds by Horner instead of the three tanh layers."""
    import torch
    from matpl_amd import autotune
    from matpl_amd.md import DPMD
    N = frame.natoms
    md = DPMD(frame, w, vel, precision=precision, skin=0.5, mnl=160, dt=1.0, rebuild_every=10, device=dev,
              compress=ct)
    md.check()
    launch = None
    if tune:
        key = autotune.workload_key(w.cfg.emb_sizes[0], w.cfg.ntypes, md.mnl, w.cfg.m, N,
                                    f"{precision}-compress{ct.order}")
        launch = autotune.tune_fused(md.compute, key, log=log)
    for _ in range(warmup):
        md.step()
    while md.nstep % md.rebuild_every:
        md.step()
    torch.cuda.synchronize()
    md.check()
    md.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        done += md.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    md.check()
    ms = (t1 - t0) * 1e3 / done
    log(f"fused {precision} compressed (order {ct.order}, dx {ct.dx}): {ms:.3f} ms/step")
    del md
    torch.cuda.empty_cache()
    return {"value": round(N * done / (t1 - t0), 1), "ms_per_step": round(ms, 4), "steps": done,
            "order": ct.order, "dx": ct.dx, "table_rows": ct.n_fine + ct.n_coarse, "launch": launch}


def copy_calibration(mib):
    """A known-traffic kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE: copy a buffer of `mib` MiB 3 times."""
    import torch
    a = torch.empty(mib << 18, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 x {mib} MiB read + written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp32"])
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--reps", default=None, help="override cells, e.g. 8,8,8")
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of a captured HIP graph")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes instead of the autotuned ones")
    ap.add_argument("--launch", default=None, metavar="FB,FL,FX,BB,BL,BX",
                    help="fixed launch shapes of the fused forward/backward kernels (blocks, LDS weights -1/0/1, "
                         "XCD order 0/1), no tuning")
    ap.add_argument("--no-nep", action="store_true", help="skip the NEP legs (SiC 1,728 atoms; 128-frame training)")
    ap.add_argument("--no-train", action="store_true", help="skip the DP training leg (256 LGPS frames)")
    ap.add_argument("--compress", action="store_true",
                    help="also time the MD workload with the compressed embedding (the reference's DP compress, "
                         "tables read by the fused kernels)")
    ap.add_argument("--store", action="store_true",
                    help="MD with the fused store mode: the forward keeps G and dG/ds per neighbor and the backward "
                         "reads them instead of evaluating the embedding net again")
    ap.add_argument("--only", default=None, choices=["nep", "nep_train", "dp_train", "dp_train_fused"],
                    help="run one leg alone and print its JSON (for profiling that leg); no MD line")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MiB",
                    help="after the run, copy a buffer of this size 3 times (a known-traffic kernel for "
                         "calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU decomposition is "
              "design-only in this build", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.only:
        leg = {"nep": nep_leg, "nep_train": nep_train_leg, "dp_train": dp_train_leg,
               "dp_train_fused": dp_train_fused_leg}[args.only]
        given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
        res = leg(steps=args.steps) if given else leg()      # each leg has its own default step count
        if args.copy_calib:
            copy_calibration(args.copy_calib)
        print(json.dumps(res), flush=True)
        return
    from matpl_amd import calcops
    from matpl_amd.md import DPMD, DPMDFP64
    calcops.register()
    dev = torch.device("cuda:0")
    reps = tuple(int(x) for x in args.reps.split(",")) if args.reps else REPS
    frame, w, vel = build_workload(reps, dev)
    N = frame.natoms
    log(f"workload: HfO2 {reps} = {N} atoms")

    md = DPMD(frame, w, vel, precision=args.precision, skin=0.5, mnl=160, dt=1.0, rebuild_every=10, device=dev,
              store=args.store)
    md.check()
    launch = None
    from matpl_amd import autotune
    if args.launch:
        launch = autotune.parse_launch(args.launch)
    elif not args.no_tune:
        key = autotune.workload_key(w.cfg.emb_sizes[0], w.cfg.ntypes, md.mnl, w.cfg.m, N,
                                    args.precision + ("-store" if args.store else ""))
        launch = autotune.tune_fused(md.compute, key, log=log)
    for _ in range(args.warmup):
        md.step()
    while md.nstep % md.rebuild_every:
        md.step()
    torch.cuda.synchronize()
    md.check()
    use_graph = not args.no_graph
    if use_graph:
        md.capture()
    log(f"warmup done ({md.nstep} steps){', captured a 10-step HIP graph' if use_graph else ''}")

    def run(k):
        done = 0
        while done < k:
            if md.graph is not None and md.nstep % md.rebuild_every == 0 and k - done >= md.rebuild_every:
                done += md.step()
            else:
                md._step()
                done += 1

    # timed region
    ev = []
    if not use_graph:   # eager: HIP events on the launch stream around every backward launch
        md.model.events = ev
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    md.check()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = N * args.steps / (t1 - t0)
    log(f"fused {args.precision}: {ms_per_step:.3f} ms/step, {value:.4g} atom-steps/s")
    if use_graph:
        # kernel durations: one more block of eager steps with events around the backward launches
        md.model.events = ev
        for _ in range(md.rebuild_every):
            md._step()
        torch.cuda.synchronize()
    md.model.events = None
    md.check()

    # the accuracy gate at the end of the trajectory as well: the FP64 path at the final positions
    end_gate = None
    if not args.no_fp64:
        from matpl_amd.structures import Frame
        fe = Frame(frame.lattice, md.pos.cpu().numpy(), frame.types, frame.elements)
        ref_end = DPMDFP64(fe, w, md.vel.cpu().numpy(), dt=1.0, device=dev)
        end_gate = {"steps": md.nstep,
                    "force_rmse_vs_fp64_eV_per_A":
                        float(torch.sqrt(((md.F.double() - ref_end.F64) ** 2).mean()).item()),
                    "energy_rel_err_vs_fp64": abs(md.energy() - float(ref_end.E.sum().item())) /
                    abs(float(ref_end.E.sum().item()))}
        log(f"after {md.nstep} steps: force RMSE vs FP64 = {end_gate['force_rmse_vs_fp64_eV_per_A']:.3e} eV/A")
        del ref_end
        torch.cuda.empty_cache()

    # roofline of the dominant (backward) kernel from the recorded events
    bwd_ms = [a.elapsed_time(b) for a, b in ev]
    from matpl_amd.neighbor import find_neighbore_gpu
    nl_rc = find_neighbore_gpu(frame.lattice, frame.frac(), frame.types, 2, w.cfg.m, w.cfg.rcut, device=dev,
                               want_dR=False)
    nvalid = int(nl_rc.counts.sum().item())      # in-cutoff neighbors at the initial configuration
    del nl_rc
    roofline = None
    traffic = None
    if reps == REPS and args.precision == "bf16":
        traffic = BWD_TRAFFIC_BYTES["xcd" if launch and launch["backward"]["xcd"] else "plain"]
    if bwd_ms and nvalid:
        kt = float(np.mean(bwd_ms)) * 1e-3
        flops = flops_per_neighbor_bwd(w.cfg.emb_out) * nvalid
        achieved = flops / kt / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / BF16_PEAK_TFLOPS, 5),
                    "traffic": traffic,
                    "traffic_GBps": round(traffic / kt / 1e9, 1) if traffic else None,
                    "kernel": "dp_fused_bwd_kernel", "avg_ms": round(kt * 1e3, 4),
                    "in_cutoff_neighbors": nvalid,
                    # the kernel issues ~69 VALU per MFMA instruction: VALU issue, not the matrix
                    # cores or HBM, limits it (SQ counters in profiles/dp_md_hfo2_99636_instruction_mix.md)
                    "limiter": "valu"}
        # the whole MD step against the same peak: 4 (1375 n + 26,650) flop per atom and step
        # (forward, fitting, backward of the DP path; n = in-cutoff neighbors of the atom)
        path_flops = 4.0 * (1375.0 * nvalid + 26650.0 * N)
        roofline["path_tflops"] = round(path_flops / (ms_per_step * 1e-3) / 1e12, 3)
        roofline["path_frac"] = round(roofline["path_tflops"] / BF16_PEAK_TFLOPS, 5)

    compressed, ct = None, None
    if args.compress:
        ct = compress_table(w, dev)
        compressed = compressed_md_leg(frame, w, vel, dev, ct, args.precision, args.steps, args.warmup,
                                       tune=not (args.no_tune or args.launch))

    # FP64 drop-in baseline on the same GPU, and the accuracy gate on this workload
    fp64 = None
    force_rmse = None
    if not args.no_fp64:
        ref = DPMDFP64(frame, w, vel, dt=1.0, device=dev)
        chk = DPMD(frame, w, vel, precision=args.precision, skin=0.5, mnl=160, device=dev, store=args.store)
        F64 = ref.F64
        force_rmse = float(torch.sqrt(((chk.F.double() - F64) ** 2).mean()).item())
        e_rel = abs(chk.energy() - float(ref.E.sum().item())) / abs(float(ref.E.sum().item()))
        log(f"force RMSE vs FP64 = {force_rmse:.3e} eV/A, |dE|/E = {e_rel:.2e}")
        del chk
        if compressed is not None:
            chk = DPMD(frame, w, vel, precision=args.precision, skin=0.5, mnl=160, device=dev, compress=ct)
            compressed["force_rmse_vs_fp64_eV_per_A"] = float(torch.sqrt(((chk.F.double() - F64) ** 2).mean()).item())
            compressed["energy_rel_err_vs_fp64"] = abs(chk.energy() - float(ref.E.sum().item())) / abs(
                float(ref.E.sum().item()))
            log(f"compressed: force RMSE vs FP64 = {compressed['force_rmse_vs_fp64_eV_per_A']:.3e} eV/A")
            del chk
        ref.step()
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(args.fp64_steps):
            ref.step()
        torch.cuda.synchronize()
        tb = time.perf_counter()
        fp64_ms = (tb - ta) * 1e3 / args.fp64_steps
        fp64 = {"value": N * args.fp64_steps / (tb - ta), "ms_per_step": round(fp64_ms, 3), "steps": args.fp64_steps}
        if compressed is not None:
            compressed["speedup_vs_fp64_dropin"] = round(compressed["value"] / fp64["value"], 2)
        log(f"fp64 drop-in: {fp64_ms:.2f} ms/step")
        del ref
        torch.cuda.empty_cache()

    if args.copy_calib:
        copy_calibration(args.copy_calib)

    autotune.reset()        # the other legs run other workloads: default launch shapes
    nep = None if args.no_nep else nep_leg()
    nep_train = None if args.no_nep else nep_train_leg()
    dp_train = None if args.no_train else dp_train_leg()
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16" if args.precision == "bf16" else "fp32", "data": "synthetic",
        "config": {"workload": f"DP se_e2_a velocity-Verlet MD, fluorite HfO2 {reps[0]}x{reps[1]}x{reps[2]} cells "
                               f"({N} atoms), rcut 6.0 A, m 120/type, NL skin 0.5 A rebuilt every 10 steps, "
                               "E+F+V every step",
                   "natoms": N, "precision": args.precision + " nets / fp32 descriptors / fp64 min-image+reductions",
                   "hip_graph": use_graph, "store": args.store},
        "fp64_dropin": fp64,
        "speedup_vs_fp64_dropin": round(value / fp64["value"], 2) if fp64 else None,
        "force_rmse_vs_fp64_eV_per_A": force_rmse,
        "end_of_trajectory_gate": end_gate,
        "roofline": roofline,
        "launch": launch,
        "cpu_baseline": cpu,
        "nep": nep,
        "nep_train": nep_train,
        "dp_train": dp_train,
        "compressed_md": compressed,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
