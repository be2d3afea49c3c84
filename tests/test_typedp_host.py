"""Type-embedding DP folding (matpl_amd.typedp) on the CPU: the folded per-pair first layer equals the
shared layer on [s, v(t_j)] for every neighbor type, and the other layers are shared unchanged;
atomic-number type vectors are min-max normalized over 1..109."""
import numpy as np

from matpl_amd.dpnet import DPConfig
from matpl_amd.typedp import TypeDPWeights, fold, type_vectors


def test_type_vectors():
    v = type_vectors([1, 8, 72, 109])
    assert np.allclose(v[:, 0], [0.0, 7 / 108, 71 / 108, 1.0])
    v2 = type_vectors([8, 72], ["atomic_number", "p"], values={"p": {8: 2.0, 72: 4.0}}, ranges={"p": (0.0, 8.0)})
    assert np.allclose(v2, [[7 / 108, 0.25], [71 / 108, 0.5]])


def test_fold_first_layer_and_shared_layers():
    cfg = DPConfig(ntypes=3, m=10)
    tv = np.random.default_rng(1).uniform(size=(3, 4))
    tw = TypeDPWeights.random(cfg, tv, seed=5)
    w = fold(tw)
    s = np.random.default_rng(2).normal(size=(7, 1))
    W, b, _ = tw.emb[0]
    for ti in range(3):
        for tj in range(3):
            layers = w.emb[ti * 3 + tj]
            z_shared = np.concatenate([s, np.repeat(tv[tj][None], 7, 0)], 1) @ W + b
            z_fold = s @ layers[0][0] + layers[0][1]
            assert np.allclose(z_fold, z_shared, rtol=1e-14, atol=1e-13)
            for (Wa, ba, _), (Wb, bb, _) in zip(layers[1:], tw.emb[1:]):
                assert Wa is Wb and ba is bb
    assert w.fit is tw.fit
