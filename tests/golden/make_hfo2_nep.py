"""Writes tests/golden/hfo2_nep/: the trained HfO2 NEP model the reference ships
(example/HfO2/nep_demo/nep_lmps/nep_to_lmps.txt, copied unchanged) and four of its DFT-labelled
training frames (example/HfO2/pwdata/init_001_50/train: fractional positions, lattice, atomic
numbers, energies, forces) as frames.npz. Run in a tree where the reference is at /root/reference."""
import os
import shutil

import numpy as np

REF = "/root/reference/example/HfO2"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hfo2_nep")


def main():
    os.makedirs(OUT, exist_ok=True)
    shutil.copyfile(os.path.join(REF, "nep_demo/nep_lmps/nep_to_lmps.txt"), os.path.join(OUT, "nep.txt"))
    d = os.path.join(REF, "pwdata/init_001_50/train")
    ld = lambda n: np.load(os.path.join(d, n), allow_pickle=False)
    np.savez(os.path.join(OUT, "frames.npz"), lattice=ld("lattice.npy")[:4], position=ld("position.npy")[:4],
             image_type=ld("image_type.npy")[0], energies=ld("energies.npy")[:4, 0], forces=ld("forces.npy")[:4])


if __name__ == "__main__":
    main()
