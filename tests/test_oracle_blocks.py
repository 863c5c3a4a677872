"""The oracle's block-subset flux_div (used by the full-size GPU parity test)
computes exactly what the whole-grid flux_div computes on those blocks and
leaves the other blocks untouched."""
import numpy as np

import golden_io as gi
from cases import FLUX_CASES, G1, GAMMA, R
from oracle import pyoracle as po


def test_block_subset_equals_whole_grid():
    g = po.make_grid(G1["nb"], G1["nc"], G1["ng"], (0, 0, 0), (gi.TWO_PI_REF,) * 3)
    q = gi.load("g1_prim_ghost")
    fl = po.make_flux(*[po.make_term(**t) for t in FLUX_CASES["keep2_visc_pow"]])
    whole = po.flux_div(g, po.Gas(GAMMA, R), fl, q)
    blocks = [0, 5, 11]
    sub = po.flux_div_blocks(g, po.Gas(GAMMA, R), fl, q, blocks)
    for lb in range(q.shape[0]):
        if lb in blocks:
            assert np.array_equal(gi.interior(sub, 2)[lb], gi.interior(whole, 2)[lb])
        else:
            assert not np.any(sub[lb])
