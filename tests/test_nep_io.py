"""nep.txt reader/writer (matpl_amd.nep_io) on the trained HfO2 model the reference ships
(tests/golden/hfo2_nep, written by tests/golden/make_hfo2_nep.py):
  * the layout is consumed exactly (the reader checks the value count as the reference's does);
  * write -> read reproduces every parameter bit for bit, and the written file equals the original
    value for value;
  * the GPUMD common-bias and nep5 variants of the layout read to the same model;
  * pinned against data produced outside this build: the model, read here and evaluated by the
    FP64 oracle, reproduces the DFT forces of four of its training frames (RMSE 0.05 eV/A against
    an RMS of 0.97 eV/A) and their energies to within the model's own offset (< 50 meV/atom)."""
import os

import numpy as np
import pytest

import oracle.nep as on
from matpl_amd import structures as st
from matpl_amd.nep_io import read_nep_txt, write_nep_txt

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hfo2_nep")
MODEL = os.path.join(HERE, "nep.txt")


def _frames(elements):
    d = np.load(os.path.join(HERE, "frames.npz"))
    z2t = {8: elements.index("O"), 72: elements.index("Hf")}
    types = np.array([z2t[int(z)] for z in d["image_type"]], np.int32)
    frames = []
    for k in range(d["lattice"].shape[0]):
        L = d["lattice"][k].reshape(3, 3)
        frames.append(st.Frame(L, d["position"][k].reshape(-1, 3) @ L, types, list(elements)))
    return frames, d["energies"], d["forces"]


def test_header_and_round_trip(tmp_path):
    w, el, mx = read_nep_txt(MODEL)
    assert el == ["O", "Hf"] and mx == (100, 100)
    c = w.cfg
    assert (c.ntypes, c.rcut_radial, c.rcut_angular, c.n_max_radial, c.basis_radial, c.l_max, c.neuron) == \
        (2, 6.0, 6.0, 4, 12, [4, 2, 1], [100])
    assert c.nfeat == 35 and w.fit[0][0][0].shape == (35, 100)
    out = tmp_path / "rt.txt"
    write_nep_txt(w, el, str(out), mx)
    a = [float(v) for v in open(MODEL).read().splitlines()[6:] if v.strip()]
    b = [float(v) for v in open(out).read().splitlines()[6:] if v.strip()]
    assert a == b
    w2, _, _ = read_nep_txt(str(out))
    for n1, n2 in zip(w.fit, w2.fit):
        for l1, l2 in zip(n1, n2):
            assert np.array_equal(l1[0], l2[0]) and np.array_equal(l1[1], l2[1])
    assert np.array_equal(w.c2, w2.c2) and np.array_equal(w.c3, w2.c3) and np.array_equal(w.q_scaler, w2.q_scaler)


def test_layout_variants_and_errors(tmp_path):
    w, el, mx = read_nep_txt(MODEL)
    lines = open(MODEL).read().splitlines()
    T, F, H = 2, 35, 100
    per = F * H + 2 * H
    body = lines[6:]
    # GPUMD nep4: one common output bias instead of T
    w_same = [net[1][1][0] for net in w.fit]
    common = body[:T * per] + [lines[6 + T * per]] + body[T * per + T:]
    (tmp_path / "gpumd.txt").write_text("\n".join(lines[:6] + common) + "\n")
    wg, _, _ = read_nep_txt(str(tmp_path / "gpumd.txt"))
    assert all(net[1][1][0] == w_same[0] for net in wg.fit)
    assert np.array_equal(wg.c2, w.c2) and np.array_equal(wg.q_scaler, w.q_scaler)
    # nep5: each type's output bias after its W1, then a 0 in the common slot
    b5 = []
    for t in range(T):
        b5 += body[t * per:(t + 1) * per] + [body[T * per + t]]
    b5 += ["0"] + body[T * per + T:]
    (tmp_path / "nep5.txt").write_text("\n".join(["nep5" + lines[0][4:]] + lines[1:6] + b5) + "\n")
    w5, _, _ = read_nep_txt(str(tmp_path / "nep5.txt"))
    for n1, n2 in zip(w.fit, w5.fit):
        assert np.array_equal(n1[0][0], n2[0][0]) and np.array_equal(n1[1][1], n2[1][1])
    (tmp_path / "short.txt").write_text("\n".join(lines[:-3]) + "\n")
    with pytest.raises(ValueError):
        read_nep_txt(str(tmp_path / "short.txt"))
    (tmp_path / "zbl.txt").write_text("\n".join([lines[0], "zbl 1.0 2.0"] + lines[1:]) + "\n")
    with pytest.raises(NotImplementedError):
        read_nep_txt(str(tmp_path / "zbl.txt"))


def test_trained_model_reproduces_dft_labels():
    w, el, _ = read_nep_txt(MODEL)
    frames, E, F = _frames(el)
    r = on.nep_evaluate(frames, w)
    Fp = np.asarray(r["F"]).reshape(len(frames), -1)
    rmse = float(np.sqrt(((Fp - F) ** 2).mean()))
    assert rmse < 0.1, rmse
    assert float(np.sqrt((F ** 2).mean())) > 0.5                 # the labels are not trivially small
    per_atom = (np.asarray(r["E"]).reshape(-1) - E) / frames[0].natoms
    assert np.abs(per_atom).max() < 0.05, per_atom
