"""Host-side NEP packing (matpl_amd.nep.pack_frames / expand_box) against the oracle's C restatement of
expand_box, bit for bit, including skewed and small cells that need several replicas."""
import numpy as np

import oracle.nep as on
from matpl_amd import nep as mn
from matpl_amd import structures as st


def test_pack_frames_matches_oracle():
    frames = [st.zincblende_sic((2, 2, 2), seed=4), st.zincblende_sic((1, 1, 1), seed=1),
              st.agausic_frames(nframes=2, nmin=30, nmax=50, seed=5)[0], st.fcc_cu((2, 2, 2), seed=2)]
    rng = np.random.default_rng(0)
    skew = st.Frame(frames[0].lattice + rng.normal(scale=0.4, size=(3, 3)), frames[0].pos, frames[0].types,
                    frames[0].elements)
    frames.append(skew)
    for rc in (6.0, 4.0, 8.5):
        want = on.pack_frames(frames, rc)
        got = mn.pack_frames(frames, rc, "cpu")
        for k in want:
            assert np.array_equal(got[k].numpy().reshape(-1), np.asarray(want[k]).reshape(-1)), (k, rc)

