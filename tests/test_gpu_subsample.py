"""Subsampling on the GPU's ring: the C1 box (64^3, clumps, periodic ghosts)
linked through the C ABI, then the library's fof_subsample over the returned
next_out against the oracle's restatement of src_v2/FOF.c:172-284 over the
oracle's own ring of the same groups, with the same draws; and the files
written from the lists."""
import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import formats as F
from cm_halofinder_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("samp_in,samp_out", [(0.1, 0.01), (1.0, 0.0), (0.5, 0.5)])
def test_subsample_on_gpu_ring(ctx, tmp_path, samp_in, samp_out):
    bx = W.plain(64)
    _, P, vols = W.host_domain(bx)
    d = vols[0].domain
    n = len(P)
    print(f"case subsample C1 n={n} in={samp_in} out={samp_out}", flush=True)
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    rc, nh = ctx.link(L.particles(P), d, L.opts(0), nxt, lab, None, 0, None)
    L.check(rc, "fof_link_gpu")
    od = O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)
    olab, _ = O.link_uf(P["Pos"], od)
    oring = O.ring_from_labels(olab)
    assert np.array_equal(nxt.astype(np.int64) - 1, oring.astype(np.int64))
    draws = np.random.default_rng(11).random(n)
    k = [0]

    def uniform():
        v = draws[k[0]]
        k[0] += 1
        return v

    rmin, rmax = list(d.rmin), list(d.rmax)
    rc, ih, oh = F.subsample(P, nxt, samp_in, samp_out, rmin, rmax, d.nphalomin, uniform)
    assert rc == 0
    oih, ooh, used = O.subsample(oring, np.ascontiguousarray(P["Pos"]), samp_in, samp_out, rmin, rmax,
                                 d.nphalomin, draws)
    assert np.array_equal(ih, oih) and np.array_equal(oh, ooh) and k[0] == used
    # kept particles lie in the rank box; halo members only in the in-halo list
    for lst in (ih, oh):
        pos = P["Pos"][lst - 1].astype(np.float64)
        assert np.all(pos >= np.array(rmin)) and np.all(pos < np.array(rmax))
    sizes = np.bincount(lab, minlength=n)
    assert np.all(sizes[lab[ih - 1]] >= d.nphalomin)
    assert np.all(sizes[lab[oh - 1]] < d.nphalomin)
    p = tmp_path / "sub_inhalo.0"
    assert F.write_subsample(p, F.OUT_PARTICLE_ID, P, ih) == 0
    if len(ih):
        ids = [int(ln.split()[0]) for ln in open(p).read().splitlines()]
        assert ids == P["ID"][ih - 1].astype(np.int64).tolist()
