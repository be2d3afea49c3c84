"""Host-side MD helpers: the spatial atom order is a permutation that keeps atoms grouped by type
and places atoms of the same spatial bin next to each other."""
import numpy as np

from matpl_amd import structures as st
from matpl_amd.md import spatial_order


def test_spatial_order_is_type_grouped_permutation():
    f = st.fluorite_hfo2((6, 6, 7), seed=5)
    o = spatial_order(f)
    assert np.array_equal(np.sort(o), np.arange(f.natoms))
    t = f.types[o]
    assert np.all(np.diff(t) >= 0)
    # the neighbors of an atom sit in a narrower index window than in lattice order
    from scipy.spatial import cKDTree
    inv = np.argsort(o)
    pairs = cKDTree(f.pos).query_pairs(4.0, output_type="ndarray")
    span0 = np.abs(pairs[:, 0] - pairs[:, 1])
    span1 = np.abs(inv[pairs[:, 0]] - inv[pairs[:, 1]])
    same = f.types[pairs[:, 0]] == f.types[pairs[:, 1]]
    assert np.median(span1[same]) < np.median(span0[same])
