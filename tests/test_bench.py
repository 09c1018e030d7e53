"""bench.py's algorithmic work count (no GPU)."""
import numpy as np
import torch

import bench
import oracle as O
from cm_halofinder_amd import workloads as W


def test_candidate_pairs_match_reference_sweep():
    # the count is the reference sweep's candidate set (src_v2/FOF.c:94-161),
    # which the oracle's union-find walks pair by pair
    bx = W.plain(24, seed=5)
    _, P, vols = W.host_domain(bx)
    d = vols[0].domain
    od = O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)
    _, pairs = O.link_uf(P["Pos"], od)
    recs = torch.from_numpy(P.view(np.uint8).reshape(-1).copy())
    for T in (1, 12, 96, 1 << 30):
        by, nd = bench.candidate_pairs(recs, vols, T)
        assert by["big"] + by["small"] + by["sparse"] == pairs
        if T == 1 << 30:
            assert by["big"] == by["small"] == 0 and nd == 0
    by1, nd1 = bench.candidate_pairs(recs, vols, 1)
    assert nd1 > 0 and by1["sparse"] == 0  # every occupied cell is dense at T = 1
    # a pair touching a dense cell is charged to a dense kernel
    by12, _ = bench.candidate_pairs(recs, vols, 12)
    byinf, _ = bench.candidate_pairs(recs, vols, 1 << 30)
    assert by12["sparse"] < byinf["sparse"]
