"""GPU parity of the full output set (fof_link_gpu_ex) and of the edge cases
the reference's arithmetic fixes, through the C ABI against the oracle.

Bar: labels, ring, canonical minimum-ID labels, halo table and raw second
moments bit-exact against the oracle on the same ring (the sums are taken in
the same order); centred second moments within 1e-12 of the Cauchy-Schwarz
scale sqrt(c_aa c_bb) of the compensated two-pass oracle.
"""
import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import workloads as W

pytestmark = pytest.mark.gpu


def odomain(v):
    d = v.domain if hasattr(v, "domain") else v
    return O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)


def link_ex(ctx, P, vols, flags=0, T=0, cap=None, moments=True):
    n = len(P)
    cap = n if cap is None else cap
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    mid = np.empty(n, np.uint64)
    halos = np.empty(max(cap, 1), L.HALO_DTYPE)
    mom = np.empty(max(cap, 1), L.MOMENTS_DTYPE) if moments else None
    hoff = np.zeros(len(vols) + 1, np.uint64)
    st = L.Stats()
    rc, nh = ctx.link_ex(L.particles(P), vols, L.opts(flags, 0.0, T), nxt, lab, mid, halos, mom, cap,
                         hoff, st)
    return rc, nh, nxt, lab, mid, halos, mom, hoff, st


def check_volume(P, vol, lab, nxt, mid, halos, mom):
    """One volume's outputs against the oracle (volume-local values)."""
    od = odomain(vol)
    olab, _ = O.link_uf(P["Pos"], od)
    assert np.array_equal(lab, olab), f"labels differ at {np.flatnonzero(lab != olab)[:10]}"
    oring = O.ring_from_labels(olab)
    assert np.array_equal(nxt.astype(np.int64) - 1, oring.astype(np.int64))
    assert np.array_equal(mid, O.minid_labels(olab, P["ID"]))
    oh = O.retrieve(oring, P["Pos"], P["Vel"], od)
    assert len(oh) == len(halos)
    for f in ("head", "n", "owned", "edge_warn", "sum_pos", "sum_vel"):
        assert np.array_equal(halos[f], oh[f]), f
    if mom is not None and len(oh):
        raw, cen = O.moments(oring, P["Pos"], P["Vel"], oh["head"])
        assert np.array_equal(mom["pos2"], raw[:, :6]) and np.array_equal(mom["vel2"], raw[:, 6:])
        gc = np.concatenate([mom["cpos2"], mom["cvel2"]], 1)
        diag = np.abs(cen[:, [0, 1, 2, 0, 0, 1, 6, 7, 8, 6, 6, 7]] * cen[:, [0, 1, 2, 1, 2, 2, 6, 7, 8, 7, 8, 8]])
        scale = np.sqrt(diag) + 1e-300
        err = np.abs(gc - cen) / scale
        assert err.max() < 1e-12, err.max()
    return olab


def test_ex_all_outputs_host(ctx):
    bx = W.plain(40, seed=31)
    _, P, vols = W.host_domain(bx)
    rc, nh, nxt, lab, mid, halos, mom, hoff, st = link_ex(ctx, P, vols, L.OPT_TIMING)
    L.check(rc)
    check_volume(P, vols[0], lab, nxt, mid, halos[:nh], mom[:nh])
    assert hoff[0] == 0 and hoff[1] == nh
    assert st.pair_tests == sum(st.pair_tests_kernel) and st.pair_tests > 0
    assert st.pair_tests_kernel[1] + st.pair_tests_kernel[2] > 0
    print("stats", st.as_dict())


def test_ex_device_outputs_match_host(ctx):
    import torch

    bx = W.plain(32, seed=32)
    _, P, vols = W.host_domain(bx)
    rc, nh, nxt, lab, mid, halos, mom, _, _ = link_ex(ctx, P, vols)
    L.check(rc)
    n = len(P)
    dP = torch.from_numpy(P.view(np.uint8)).cuda()
    dn = torch.empty(n, dtype=torch.int32, device="cuda")
    dl = torch.empty(n, dtype=torch.int32, device="cuda")
    dm = torch.empty(n, dtype=torch.int64, device="cuda")
    dh = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    dmo = torch.empty(n * 192, dtype=torch.uint8, device="cuda")
    hoff = np.zeros(2, np.uint64)
    rc, nh2 = ctx.link_ex(L.particles(dP, n=n), vols, L.opts(), dn, dl, dm, dh, dmo, n, hoff)
    L.check(rc)
    torch.cuda.synchronize()
    assert nh2 == nh
    assert np.array_equal(dn.cpu().numpy().view(np.uint32), nxt)
    assert np.array_equal(dl.cpu().numpy().view(np.uint32), lab)
    assert np.array_equal(dm.cpu().numpy().view(np.uint64), mid)
    assert np.array_equal(dh[: nh * 64].cpu().numpy().view(L.HALO_DTYPE), halos[:nh])
    assert np.array_equal(dmo[: nh * 192].cpu().numpy().view(L.MOMENTS_DTYPE), mom[:nh])


def test_batch_halo_cap_too_small_reports_need(ctx):
    bx = W.plain(32, seed=33)
    recs = L.gen_host(bx.gen())
    P, _, vols = L.ghosts_host(bx.ghost(2), recs)
    rc, nh, *_ = link_ex(ctx, P, vols)
    L.check(rc)
    assert nh > 8
    rc2, nh2, _, _, _, _, _, hoff2, st2 = link_ex(ctx, P, vols, cap=3)
    assert rc2 == L.E_RANGE
    assert nh2 == nh and hoff2[-1] == nh and st2.nhalos == nh


@pytest.fixture(scope="module")
def c2_box():
    """The 1024^3 workload's density and clump law at 128^3 (box 125 Mpc/h,
    l = 0.1953125, 3906 clumps, no planted pairs)."""
    bx = W.Box(seed=7, pside=128, box=125.0)
    recs = L.gen_host(bx.gen())
    return bx, recs


@pytest.mark.timeout(300)
def test_c2_density_plain(ctx, c2_box):
    bx, recs = c2_box
    P, _, vols = L.ghosts_host(bx.ghost(1), recs)
    print(f"case C2 density 128^3 plain n={len(P)}", flush=True)
    rc, nh, nxt, lab, mid, halos, mom, _, st = link_ex(ctx, P, vols, moments=False)
    L.check(rc)
    check_volume(P, vols[0], lab, nxt, mid, halos[:nh], None)
    assert st.ndense > 0


@pytest.mark.timeout(300)
def test_c2_density_batched_2cubed(ctx, c2_box):
    bx, recs = c2_box
    P, _, vols = L.ghosts_host(bx.ghost(2), recs)
    print(f"case C2 density 128^3 as 2^3 volumes n={len(P)}", flush=True)
    rc, nh, nxt, lab, mid, halos, mom, hoff, st = link_ex(ctx, P, vols, moments=False)
    L.check(rc)
    for v, vol in enumerate(vols):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        h0, h1 = int(hoff[v]), int(hoff[v + 1])
        check_volume(P[a:b], vol, lab[a:b], nxt[a:b], mid[a:b], halos[h0:h1], None)


# ---- edge cases of the reference arithmetic ---------------------------------
def run_edge(ctx, P, dom):
    vol = L.Volume()
    vol.offset, vol.n, vol.domain = 0, len(P), dom
    rc, nh, nxt, lab, mid, halos, mom, _, st = link_ex(ctx, P, [vol])
    L.check(rc)
    return check_volume(P, vol, lab, nxt, mid, halos[:nh], mom[:nh]), st


def records(pos, seed=0):
    rng = np.random.default_rng(seed)
    P = np.zeros(len(pos), L.PART_DTYPE)
    P["Pos"] = np.asarray(pos, np.float32)
    P["Vel"] = rng.normal(0, 1, (len(pos), 3)).astype(np.float32)
    P["ID"] = rng.permutation(len(pos) * 7)[: len(pos)].astype(np.uint64)
    return P


def test_coincident_duplicates(ctx):
    # stacks of identical positions (d^2 = 0 links), some on cell faces,
    # enough to make dense cells
    rng = np.random.default_rng(41)
    base = rng.random((60, 3)) * 8.0
    base[:20, 0] = np.floor(base[:20, 0])  # on x faces (lcell = 1)
    pos = np.repeat(base, rng.integers(1, 60, 60), 0)
    P = records(pos, 41)
    dom = L.domain([0] * 3, [8] * 3, [0] * 3, [8] * 3, [1.0] * 3, 0.05, 2)
    olab, _ = run_edge(ctx, P, dom)
    for k in range(60):
        m = np.all(P["Pos"] == base[k].astype(np.float32), 1)
        assert len(np.unique(olab[m])) == 1


def test_zero_linking_length(ctx):
    # l = 0: only exact duplicates link (distsq = 0 is not > 0)
    rng = np.random.default_rng(42)
    pos = rng.random((4000, 3)) * 10.0
    pos[2000:2600] = pos[:600]
    P = records(pos, 42)
    dom = L.domain([0] * 3, [10] * 3, [0] * 3, [10] * 3, [10 / np.cbrt(4000)] * 3, 0.0, 2)
    olab, _ = run_edge(ctx, P, dom)
    assert len(np.unique(olab)) == 4000 - 600


def test_faces_and_buffer_max(ctx):
    # coordinates exactly on reference-cell faces (x = k lcell) and at
    # rmax_buff itself (the last cell, src_v2/FOF.c:79-85 with the ceil()+1 grid)
    rng = np.random.default_rng(43)
    lc = 0.5
    n = 20000
    pos = rng.random((n, 3)) * 16.0
    face = rng.random((n, 3)) < 0.5
    pos[face] = np.round(pos[face] / lc) * lc
    pos[:300, 0] = 16.0
    pos[300:600, 1] = 16.0
    pos[600:700] = 16.0
    P = records(pos, 43)
    dom = L.domain([0] * 3, [16] * 3, [0] * 3, [16] * 3, [lc] * 3, 0.125, 2)
    run_edge(ctx, P, dom)


@pytest.mark.parametrize("below", [False, True])
def test_lattice_at_linking_length(ctx, below):
    # a cubic lattice of spacing a = l exactly: every neighbour pair is an
    # exact tie d^2 = l^2 and links (one group); with l one ulp below a,
    # nothing links
    a = 0.25
    g = np.arange(24) * a + 1.0
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    P = records(pos, 44)
    l = np.nextafter(a, 0.0) if below else a
    dom = L.domain([0] * 3, [8] * 3, [0] * 3, [8] * 3, [0.75] * 3, l, 2)
    olab, st = run_edge(ctx, P, dom)
    assert len(np.unique(olab)) == (len(P) if below else 1)
    if not below:
        assert st.rechecks > 0


@pytest.mark.parametrize("side", [12, 18, 28])
def test_group_sizes_across_sort_paths(ctx, side):
    # one lattice group (a = l) of 1728, 5832 or 21952 members, beside a
    # randomly permuted background of pairs and small groups: every segment
    # sort path runs (lanes, one wave, a workgroup in LDS, the radix sort)
    a = 0.25
    g = np.arange(side) * a + 1.0
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(side)
    far = rng.random((500, 3)) * 5.5 + 9.5
    pairs = np.concatenate([far, far + np.array([0.1, 0.0, 0.0])])
    pos = np.concatenate([lat, pairs])[rng.permutation(len(lat) + len(pairs))]
    P = records(pos, side)
    dom = L.domain([0] * 3, [16] * 3, [0] * 3, [16] * 3, [0.75] * 3, a, 2)
    olab, st = run_edge(ctx, P, dom)
    assert np.bincount(np.unique(olab, return_inverse=True)[1]).max() == side ** 3


@pytest.mark.parametrize("fmt", ["ascii", "gadget1", "picola"])
def test_read_format_ghosts_link(ctx, fmt):
    # a fixture in each of the reference's input formats: read on the host,
    # ghost shell built on the device, linked, checked against the oracle
    import os

    import torch

    from cm_halofinder_amd import formats as F

    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    path = os.path.join(gold, f"particles.{fmt}")
    recs = {"ascii": lambda: F.read_ascii(path), "gadget1": lambda: F.read_gadget1(path)[0],
            "picola": lambda: F.read_picola(path)}[fmt]()
    assert recs.tobytes() == np.load(os.path.join(gold, "expected.npy"), allow_pickle=False).tobytes()
    bx = W.plain(10, seed=71)
    gh = L.ghost_params(bx.box, 2.0, 1, 1, 1, bx.pside, bx.linklength, 5)
    drecs = torch.from_numpy(recs.view(np.uint8).copy()).cuda()
    counts, vols = L.ghosts_device(gh, drecs, len(recs))
    tot = int(counts.sum())
    dout = torch.empty(tot * 32, dtype=torch.uint8, device="cuda")
    counts, vols = L.ghosts_device(gh, drecs, len(recs), dout, tot)
    P = dout.cpu().numpy().view(L.PART_DTYPE)
    hP, _, _ = L.ghosts_host(gh, recs)
    assert np.array_equal(np.sort(P, order=["ID", "Pos"]), np.sort(hP, order=["ID", "Pos"]))
    rc, nh, nxt, lab, mid, halos, mom, _, _ = link_ex(ctx, P, vols)
    L.check(rc)
    check_volume(P, vols[0], lab, nxt, mid, halos[:nh], mom[:nh])
