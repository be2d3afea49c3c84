"""Type-embedding DP folding (matpl_amd.typedp) on the CPU: the folded per-pair first layer equals the
shared layer on [s, v(t_j)] for every neighbor type, and the other layers are shared unchanged;
atomic-number type vectors are min-max normalized over 1..109."""
import numpy as np

from matpl_amd.dpnet import DPConfig
from matpl_amd.typedp import TypeDPWeights, fold, type_vectors


def test_type_vectors():
    v = type_vectors([1, 8, 72, 109], ["atomic_number"])
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


def test_default_property_vectors_match_reference_table():
    """The five default properties (nn_feature_type.py:35) for every element, against the golden vectors
    tests/golden/make_type_vectors.py computed from the reference's own table (atom_type_emb_dict.py);
    elements the reference refuses (NaN or no entry) are refused here too."""
    import json
    import os
    import pytest
    from matpl_amd.typedp import DEFAULT_PROPERTIES
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "type_vectors.json")))
    assert tuple(g["properties"]) == DEFAULT_PROPERTIES
    ok = [int(z) for z, row in g["vectors"].items() if all(x is not None for x in row)]
    bad = [int(z) for z, row in g["vectors"].items() if any(x is None for x in row)]
    assert len(ok) > 60 and bad
    want = np.array([g["vectors"][str(z)] for z in ok], np.float64)
    np.testing.assert_allclose(type_vectors(ok), want, rtol=0, atol=1e-15)
    for z in bad:
        with pytest.raises(ValueError):
            type_vectors([z])
