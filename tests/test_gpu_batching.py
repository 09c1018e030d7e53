"""Batched execution paths through the C ABI: concurrent pipelines, captured
graphs replayed, autotuning.  Every configuration must give the outputs of
the plain eager call bit for bit (and those match the oracle, checked in
test_gpu_parity / test_gpu_outputs)."""
import os

import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import workloads as W

pytestmark = pytest.mark.gpu


def batch_case(seed=51, scale=40, g=2):
    bx = W.plain(scale, seed=seed)
    recs = L.gen_host(bx.gen())
    P, _, vols = L.ghosts_host(bx.ghost(g), recs)
    return P, vols


def host_call(ctx, P, vols, o):
    n = len(P)
    out = dict(nxt=np.empty(n, np.uint32), lab=np.empty(n, np.uint32), mid=np.empty(n, np.uint64),
               hal=np.empty(n, L.HALO_DTYPE), mom=np.empty(n, L.MOMENTS_DTYPE),
               hoff=np.zeros(len(vols) + 1, np.uint64))
    st = L.Stats()
    rc, nh = ctx.link_ex(L.particles(P), vols, o, out["nxt"], out["lab"], out["mid"], out["hal"],
                         out["mom"], n, out["hoff"], st)
    L.check(rc)
    out["hal"], out["mom"] = out["hal"][:nh], out["mom"][:nh]
    return out, st


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


@pytest.mark.parametrize("pipes", [2, 3, 4, 8])
def test_pipelines_match_single(ctx, pipes):
    P, vols = batch_case()
    ref, st1 = host_call(ctx, P, vols, L.opts())
    got, st = host_call(ctx, P, vols, L.opts(0, 0.0, 0, pipes))
    assert st.pipelines == min(pipes, len(vols)) and st1.pipelines == 1
    assert same(ref, got)
    # pair-test counts vary run to run (merges found concurrently prune
    # tests); the groups do not
    assert st.nhalos == st1.nhalos and st.ngroups == st1.ngroups


def test_pipelines_device_outputs_gather_halos(ctx):
    import torch

    P, vols = batch_case(seed=52, g=3)
    ref, _ = host_call(ctx, P, vols, L.opts())
    n = len(P)
    dP = torch.from_numpy(P.view(np.uint8)).cuda()
    dn = torch.empty(n, dtype=torch.int32, device="cuda")
    dl = torch.empty(n, dtype=torch.int32, device="cuda")
    dh = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    dm = torch.empty(n * 192, dtype=torch.uint8, device="cuda")
    hoff = np.zeros(len(vols) + 1, np.uint64)
    rc, nh = ctx.link_ex(L.particles(dP, n=n), vols, L.opts(0, 0.0, 0, 4), dn, dl, None, dh, dm, n,
                         hoff)
    L.check(rc)
    assert nh == len(ref["hal"]) and np.array_equal(hoff, ref["hoff"])
    assert np.array_equal(dn.cpu().numpy().view(np.uint32), ref["nxt"])
    assert np.array_equal(dl.cpu().numpy().view(np.uint32), ref["lab"])
    assert np.array_equal(dh[: nh * 64].cpu().numpy().view(L.HALO_DTYPE), ref["hal"])
    assert np.array_equal(dm[: nh * 192].cpu().numpy().view(L.MOMENTS_DTYPE), ref["mom"])


@pytest.mark.parametrize("pipes", [1, 4])
def test_graph_replays_match_eager(ctx, pipes):
    # the first call captures, later ones replay; eager calls on other inputs
    # and host synchronisations run between the replays, and every replay is
    # checked against the eager result
    import torch

    P, vols = batch_case(seed=53)
    Q, qvols = batch_case(seed=54, scale=24)
    n = len(P)
    dP = torch.from_numpy(P.view(np.uint8)).cuda()
    dn = torch.empty(n, dtype=torch.int32, device="cuda")
    dl = torch.empty(n, dtype=torch.int32, device="cuda")
    dmid = torch.empty(n, dtype=torch.int64, device="cuda")
    dh = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    hoff = np.zeros(len(vols) + 1, np.uint64)
    ref, _ = host_call(ctx, P, vols, L.opts())
    o = L.opts(L.OPT_GRAPH, 0.0, 0, pipes)
    states = []
    for rep in range(4):
        for t in (dn, dl, dmid, dh):
            t.fill_(-1)
        st = L.Stats()
        rc, nh = ctx.link_ex(L.particles(dP, n=n), vols, o, dn, dl, dmid, dh, None, n, hoff, st)
        L.check(rc)
        torch.cuda.synchronize()
        states.append(st.graph)
        assert nh == len(ref["hal"]) and np.array_equal(hoff, ref["hoff"]), rep
        assert np.array_equal(dn.cpu().numpy().view(np.uint32), ref["nxt"]), rep
        assert np.array_equal(dl.cpu().numpy().view(np.uint32), ref["lab"]), rep
        assert np.array_equal(dmid.cpu().numpy().view(np.uint64), ref["mid"]), rep
        assert np.array_equal(dh[: nh * 64].cpu().numpy().view(L.HALO_DTYPE), ref["hal"]), rep
        host_call(ctx, Q, qvols, L.opts())  # eager work on other data in between
    assert states == [2, 1, 1, 1]


def test_graph_recaptures_on_new_arguments(ctx):
    import torch

    P, vols = batch_case(seed=55, scale=24)
    n = len(P)
    dP = torch.from_numpy(P.view(np.uint8)).cuda()
    dl = torch.empty(n, dtype=torch.int32, device="cuda")
    ref, _ = host_call(ctx, P, vols, L.opts())
    seen = []
    for T in (24, 24, 16, 16):
        st = L.Stats()
        rc, _ = ctx.link_ex(L.particles(dP, n=n), vols, L.opts(L.OPT_GRAPH, 0.0, T), None, dl, None,
                            None, None, 0, None, st)
        L.check(rc)
        torch.cuda.synchronize()
        seen.append(st.graph)
        assert np.array_equal(dl.cpu().numpy().view(np.uint32), ref["lab"])
    assert seen == [2, 1, 2, 1]


def test_autotune_matches_and_caches(ctx, tmp_path, monkeypatch):
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("FOF_TUNE_CACHE", str(path))
    P, vols = batch_case(seed=56)
    ref, _ = host_call(ctx, P, vols, L.opts())
    c2 = L.Context(0)  # a fresh context: its own in-memory cache
    try:
        got, st = host_call(c2, P, vols, L.opts(L.OPT_AUTOTUNE))
        assert same(ref, got)
        lines = path.read_text().strip().splitlines()
        assert len(lines) == 1 and "->" in lines[0]
        T, K, gb = (int(x) for x in lines[0].split("->")[1].split())
        assert st.dense_threshold == T and st.pipelines == max(1, K)
        got2, st2 = host_call(c2, P, vols, L.opts(L.OPT_AUTOTUNE))  # cached: no new entry
        assert same(ref, got2) and len(path.read_text().strip().splitlines()) == 1
    finally:
        c2.close()
    c3 = L.Context(0)  # another context reads the file
    try:
        got3, st3 = host_call(c3, P, vols, L.opts(L.OPT_AUTOTUNE))
        assert same(ref, got3) and st3.dense_threshold == T
        assert len(path.read_text().strip().splitlines()) == 1
    finally:
        c3.close()


def test_pipelines_against_oracle(ctx):
    # many small volumes, four pipelines: every volume against the oracle
    P, vols = batch_case(seed=57, scale=32, g=4)
    got, _ = host_call(ctx, P, vols, L.opts(0, 0.0, 0, 4))
    for v, vol in enumerate(vols):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        d = vol.domain
        od = O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)
        olab, _ = O.link_uf(P["Pos"][a:b], od)
        assert np.array_equal(got["lab"][a:b], olab), v
        oh = O.retrieve(O.ring_from_labels(olab), P["Pos"][a:b], P["Vel"][a:b], od)
        h = got["hal"][int(got["hoff"][v]):int(got["hoff"][v + 1])]
        assert np.array_equal(h["head"], oh["head"]) and np.array_equal(h["sum_pos"], oh["sum_pos"])


@pytest.mark.parametrize("flags", [L.OPT_WRITE_PAIRS, L.OPT_WRITE_DIRECT,
                                   L.OPT_WRITE_PAIRS | L.OPT_WRITE_DIRECT])
def test_label_and_ring_write_modes(ctx, flags):
    # labels and ring written straight by index (the default at this size)
    # or through the index partition (the default for volumes above 2^25
    # particles): the same outputs
    P, vols = batch_case(seed=58)
    ref, _ = host_call(ctx, P, vols, L.opts())
    got, _ = host_call(ctx, P, vols, L.opts(flags))
    assert same(ref, got)
