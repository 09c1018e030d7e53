"""The allocation-free entry point (fof_link_gpu_noalloc) against
fof_link_gpu_ex through the C ABI: bit-exact labels, ring, minimum-ID labels,
halos and moments, on two calls in a row (the second on a smaller input, so
that the fixed workspace holds the first call's contents), and the size
limit reported as FOF_E_RANGE."""
import numpy as np
import pytest

from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import workloads as W

pytestmark = pytest.mark.gpu


def device_outputs(torch, n):
    return dict(nxt=torch.empty(n, dtype=torch.int32, device="cuda"),
                lab=torch.empty(n, dtype=torch.int32, device="cuda"),
                mid=torch.empty(n, dtype=torch.int64, device="cuda"),
                hal=torch.empty(n * 64, dtype=torch.uint8, device="cuda"),
                mom=torch.empty(n * 192, dtype=torch.uint8, device="cuda"))


def fetch(torch, o, nh):
    torch.cuda.synchronize()
    return (o["nxt"].cpu().numpy().view(np.uint32), o["lab"].cpu().numpy().view(np.uint32),
            o["mid"].cpu().numpy().view(np.uint64), o["hal"][: nh * 64].cpu().numpy().view(L.HALO_DTYPE),
            o["mom"][: nh * 192].cpu().numpy().view(L.MOMENTS_DTYPE))


@pytest.fixture(scope="module")
def nctx():
    c = L.Context(0)
    yield c
    c.close()


def test_noalloc_matches_ex_twice(nctx):
    import torch

    cases = []
    for scale, seed in ((48, 61), (24, 62)):  # the second input is smaller
        bx = W.plain(scale, seed=seed)
        _, P, vols = W.host_domain(bx)
        cases.append((P, vols))
    for P, vols in cases:
        n = len(P)
        dP = torch.from_numpy(P.view(np.uint8)).cuda()
        parts = L.particles(dP, n=n)
        ref = device_outputs(torch, n)
        rc, nh = nctx.link_ex(parts, vols, L.opts(), ref["nxt"], ref["lab"], ref["mid"], ref["hal"],
                              ref["mom"], n)
        L.check(rc, "fof_link_gpu_ex")
        want = fetch(torch, ref, nh)
        got = device_outputs(torch, n)
        for t in got.values():
            t.fill_(-1)
        rc, nh2 = nctx.link_noalloc(parts, vols[0].domain, L.opts(), got["nxt"], got["lab"], got["mid"],
                                    got["hal"], got["mom"], n)
        L.check(rc, "fof_link_gpu_noalloc")
        assert nh2 == nh
        for a, b in zip(fetch(torch, got, nh), want):
            assert np.array_equal(a, b)


def test_noalloc_size_limit(nctx):
    import torch

    # a domain whose workspace exceeds the fixed array is refused before any
    # device work; fof_plan predicts it
    n = 5_000_000
    dP = torch.zeros(n * 32, dtype=torch.uint8, device="cuda")
    parts = L.particles(dP, n=n)
    dom = L.domain([0] * 3, [171] * 3, [0] * 3, [171] * 3, [1.0] * 3, 0.2, 20)
    vol = L.Volume()
    vol.offset, vol.n, vol.domain = 0, n, dom
    assert L.plan(parts, [vol]).workspace_bytes > L.NOALLOC_BYTES
    o = device_outputs(torch, n)
    rc, _ = nctx.link_noalloc(parts, dom, L.opts(), o["nxt"], o["lab"], None, o["hal"], None, n)
    assert rc == L.E_RANGE
    # host-resident input is not accepted
    P = np.zeros(100, L.PART_DTYPE)
    rc, _ = nctx.link_noalloc(L.particles(P), dom, L.opts(), o["nxt"], o["lab"])
    assert rc == L.E_ARG
