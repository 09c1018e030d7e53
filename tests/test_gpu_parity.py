"""GPU parity: the HIP linker against the CPU oracle through the C ABI.

Bar: bit-exact labels (group minimum index), bit-exact successor ring
(ascending index order inside each group) and bit-exact halo table (the
oracle's retrieval / Checkhalo / Output_Halos arithmetic run over that same
ring); halo sums within 1e-6 relative of compensated FP64 sums.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import workloads as W

pytestmark = pytest.mark.gpu


def odomain(v):
    d = v.domain if hasattr(v, "domain") else v
    return O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)


def gpu_link(ctx, P, dom, flags=0, band=0.0, T=0):
    n = len(P)
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    halos = np.empty(max(n, 1), L.HALO_DTYPE)
    st = L.Stats()
    rc, nh = ctx.link(L.particles(P), dom, L.opts(flags, band, T), nxt, lab, halos, len(halos), st)
    L.check(rc, "fof_link_gpu")
    return lab, nxt, halos[:nh], st


def check_against_oracle(ctx, P, dom, T=0, vl=None, ovl=None, flags=0):
    print(f"case n={len(P)} l={dom.linklength} T={T} flags={flags}", flush=True)
    lab, nxt, halos, st = gpu_link(ctx, P, dom, flags=flags, T=T)
    od = odomain(dom)
    olab, pairs = O.link_uf(P["Pos"], od, ovl)
    assert np.array_equal(lab, olab), f"labels differ at {np.flatnonzero(lab != olab)[:10]}"
    oring = O.ring_from_labels(olab)
    assert np.array_equal(nxt.astype(np.int64) - 1, oring.astype(np.int64))
    oh = O.retrieve(oring, P["Pos"], P["Vel"], od)
    assert len(oh) == len(halos)
    for f in ("head", "n", "owned", "edge_warn"):
        assert np.array_equal(halos[f], oh[f]), f
    assert np.array_equal(halos["sum_pos"], oh["sum_pos"])
    assert np.array_equal(halos["sum_vel"], oh["sum_vel"])
    if len(oh):
        sp, sv = O.props_compensated(olab, P["Pos"], P["Vel"], oh["head"])
        rel = np.abs(halos["sum_pos"] - sp) / np.maximum(np.abs(sp), 1e-30)
        assert rel.max() < 1e-6
    return lab, st


def test_c1_plain(ctx):
    bx = W.plain(64)
    _, P, vols = W.host_domain(bx)
    lab, st = check_against_oracle(ctx, P, vols[0].domain)
    assert st.ndense > 0 and st.pair_tests > 0
    print("stats", st.as_dict())


@pytest.mark.parametrize("T", [1, 4, 1 << 30])
def test_dense_threshold_invariance(ctx, T):
    bx = W.plain(24, seed=3)
    _, P, vols = W.host_domain(bx)
    check_against_oracle(ctx, P, vols[0].domain, T=T)


@pytest.mark.parametrize("b", [0.05, 0.15, 0.3, 0.6, 1.2])
def test_linking_lengths(ctx, b):
    bx = W.plain(20, seed=4, b=b)
    _, P, vols = W.host_domain(bx)
    check_against_oracle(ctx, P, vols[0].domain)


def test_random_dense_blobs(ctx):
    rng = np.random.default_rng(9)
    n = 20000
    P = np.zeros(n, L.PART_DTYPE)
    c = rng.random((30, 3)) * 20
    pos = c[rng.integers(0, 30, n)] + rng.normal(0, 0.08, (n, 3))
    P["Pos"] = np.clip(pos, 0, 20).astype(np.float32)
    P["Vel"] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    P["ID"] = np.arange(n)
    dom = L.domain([0] * 3, [20] * 3, [0] * 3, [20] * 3, [1.0] * 3, 0.05, 5)
    check_against_oracle(ctx, P, dom)


def test_many_big_cells_in_one_bucket(ctx):
    # 640 cells of 260 particles each, all among the first 2^14 cell ids:
    # one sort bucket lists more big dense cells (>= 256 particles) than its
    # LDS list holds, so part of the list is appended directly
    rng = np.random.default_rng(21)
    nx, ny, per, lc = 32, 20, 260, 0.4
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    corner = np.stack([ix.ravel() * lc, iy.ravel() * lc, np.zeros(nx * ny)], 1)
    pos = np.repeat(corner, per, 0) + rng.random((nx * ny * per, 3)) * lc * 0.999
    n = len(pos)
    P = np.zeros(n, L.PART_DTYPE)
    P["Pos"] = pos.astype(np.float32)
    P["Vel"] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    dom = L.domain([0] * 3, [nx * lc, ny * lc, lc], [0] * 3, [nx * lc, ny * lc, lc], [lc] * 3, 0.08, 5)
    lab, st = check_against_oracle(ctx, P, dom)
    assert st.ndense >= nx * ny


def test_exact_ties(ctx):
    # d^2 == l^2 exactly links (src_v2/FOF.c:357); one ulp further does not
    l = 25.0 / 128.0
    pts = []
    base = np.array([512.0, 256.0, 300.0], np.float32)
    for k in range(50):
        o = base + np.float32(3.0 * k)
        pts.append(o)
        pts.append(o + np.array([15 / 128, 20 / 128, 0], np.float32))  # 3-4-5 triangle: d = l
        far = o + np.array([0, 0, 1.5], np.float32)
        pts.append(far)
        q = far + np.array([l, 0, 0], np.float32)
        q[0] = np.nextafter(q[0], np.float32(1e9))
        pts.append(q)
    P = np.zeros(len(pts), L.PART_DTYPE)
    P["Pos"] = np.array(pts, np.float32)
    dom = L.domain([500] * 3, [700] * 3, [500, 240, 290], [700, 450, 500], [0.9765625] * 3, l, 2)
    lab, st = check_against_oracle(ctx, P, dom)
    lab = lab.reshape(-1, 4)
    assert np.all(lab[:, 1] == lab[:, 0]) and np.all(lab[:, 3] != lab[:, 2])
    # the ties are re-decided in FP64; a one-ulp offset at |x| ~ 512 changes
    # d^2 by ~6e-4 relative, outside the band, and is decided in FP32
    assert st.rechecks == 50


def test_near_ties_in_band(ctx):
    # at |x| ~ 1 one ulp moves d^2 by ~1e-6 relative: inside the 1e-5 band
    l = 25.0 / 128.0
    pts, expect = [], []
    for k in range(40):
        o = np.array([1.0 + 0.5 * (k % 4), 1.0 + 0.5 * ((k // 4) % 4), 1.0 + 0.5 * (k // 16)], np.float32)
        q = o + np.array([15 / 128, 20 / 128, 0], np.float32)
        steps = (k % 5) - 2  # -2..2 ulps on x
        for _ in range(abs(steps)):
            q[0] = np.nextafter(q[0], np.float32(9.0 if steps > 0 else -9.0))
        pts += [o, q]
        expect.append(O.pair_linked(o, q, l * l))
    P = np.zeros(len(pts), L.PART_DTYPE)
    P["Pos"] = np.array(pts, np.float32)
    dom = L.domain([0] * 3, [4] * 3, [0] * 3, [4] * 3, [0.9765625] * 3, l, 2)
    lab, st = check_against_oracle(ctx, P, dom)
    lab = lab.reshape(-1, 2)
    assert np.array_equal(lab[:, 0] == lab[:, 1], np.array(expect))
    assert 0 < sum(expect) < len(expect)
    assert st.rechecks >= 40


def test_precision_planted_pairs(ctx):
    bx = W.precision(pside=64)
    _, P, vols = W.host_domain(bx)
    lab, st = check_against_oracle(ctx, P, vols[0].domain)
    assert st.rechecks > 0.9 * bx.n * bx.pair_frac / 2


def test_device_resident_matches_host(ctx):
    import torch

    bx = W.plain(32, seed=6)
    _, P, vols = W.host_domain(bx)
    dom = vols[0].domain
    lab, nxt, halos, _ = gpu_link(ctx, P, dom)
    dP = torch.from_numpy(P.view(np.uint8)).cuda()
    n = len(P)
    dnext = torch.empty(n, dtype=torch.int32, device="cuda")
    dlab = torch.empty(n, dtype=torch.int32, device="cuda")
    dh = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    parts = L.particles(dP, n=n)
    rc, nh = ctx.link(parts, dom, L.opts(L.OPT_OUT_DEVICE), dnext, dlab, dh, n)
    L.check(rc)
    torch.cuda.synchronize()
    assert np.array_equal(dlab.cpu().numpy().view(np.uint32), lab)
    assert np.array_equal(dnext.cpu().numpy().view(np.uint32), nxt)
    hh = dh[: nh * 64].cpu().numpy().view(L.HALO_DTYPE)
    assert np.array_equal(hh, halos)


def test_batch_subvolumes(ctx):
    bx = W.plain(40, seed=8)
    recs = L.gen_host(bx.gen())
    P, counts, vols = L.ghosts_host(bx.ghost(2), recs)
    n = len(P)
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    halos = np.empty(n, L.HALO_DTYPE)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    st = L.Stats()
    rc = ctx.link_batch(L.particles(P), vols, L.opts(), nxt, lab, halos, n, hoff, st)
    L.check(rc)
    for v, vol in enumerate(vols):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        od = odomain(vol)
        olab, _ = O.link_uf(P["Pos"][a:b], od)
        assert np.array_equal(lab[a:b], olab), v
        oring = O.ring_from_labels(olab)
        assert np.array_equal(nxt[a:b].astype(np.int64) - 1, oring)
        oh = O.retrieve(oring, P["Pos"][a:b], P["Vel"][a:b], od)
        hv = halos[int(hoff[v]):int(hoff[v + 1])]
        assert np.array_equal(hv["head"], oh["head"]) and np.array_equal(hv["sum_pos"], oh["sum_pos"])
        assert np.array_equal(hv["owned"], oh["owned"])


def test_domain_error(ctx):
    P = np.zeros(10, L.PART_DTYPE)
    P["Pos"] = 1.0
    P["Pos"][3] = [5.0, 1.0, 1.0]
    dom = L.domain([0] * 3, [2] * 3, [0] * 3, [2] * 3, [1.0] * 3, 0.1, 2)
    rc, _ = ctx.link(L.particles(P), dom, None, None, None, None, 0)
    assert rc == L.E_DOMAIN
    P["Pos"][3] = [np.nan, 1.0, 1.0]
    rc, _ = ctx.link(L.particles(P), dom, None, None, None, None, 0)
    assert rc == L.E_DOMAIN
    P["Pos"][3] = [1.5, 1.0, 1.0]
    rc, _ = ctx.link(L.particles(P), dom, None, None, None, None, 0)
    assert rc == 0


def test_deterministic(ctx):
    bx = W.plain(32, seed=12)
    _, P, vols = W.host_domain(bx)
    a = gpu_link(ctx, P, vols[0].domain)
    b = gpu_link(ctx, P, vols[0].domain)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def test_generator_and_ghosts_device_match_host():
    import torch

    bx = W.plain(32, seed=13)
    g = bx.gen()
    h = L.gen_host(g)
    d = torch.empty(len(h) * 32, dtype=torch.uint8, device="cuda")
    L.gen_device(g, d)
    dh = d.cpu().numpy().view(L.PART_DTYPE)
    assert np.array_equal(dh["ID"], h["ID"])
    # device libm may round transcendental results differently in the last bit
    assert np.abs(dh["Pos"] - h["Pos"]).max() <= 1e-4
    hP, hc, hv = L.ghosts_host(bx.ghost(2), dh)
    out = torch.empty(len(hP) * 32, dtype=torch.uint8, device="cuda")
    dc, dv = L.ghosts_device(bx.ghost(2), d, len(h), out, len(hP))
    assert np.array_equal(dc, hc)
    assert np.array_equal(out.cpu().numpy().view(L.PART_DTYPE), hP)


def lightcone_like(n, R, seed):
    """Particles in a ball with density falling with radius, plus clumps."""
    rng = np.random.default_rng(seed)
    r = R * rng.random(n) ** (1 / 2.2)
    v = rng.normal(size=(n, 3))
    pos = v / np.linalg.norm(v, axis=1)[:, None] * r[:, None]
    k = n // 3
    cen = rng.normal(size=(60, 3))
    cen = cen / np.linalg.norm(cen, axis=1)[:, None] * (R * rng.random(60))[:, None]
    pos[:k] = cen[rng.integers(0, 60, k)] + rng.normal(0, 0.05 + 0.02 * np.abs(cen[rng.integers(0, 60, k)]).max(1)[:, None] / R, (k, 3))
    P = np.zeros(n, L.PART_DTYPE)
    P["Pos"] = np.clip(pos, -R, R).astype(np.float32)
    P["Vel"] = rng.normal(0, 300, (n, 3)).astype(np.float32)
    P["ID"] = np.arange(n)
    return P


@pytest.mark.parametrize("seed", [1, 2])
def test_variable_linking_length(ctx, seed):
    R = 20.0
    P = lightcone_like(40000, R, seed)
    lc = 2 * R / np.cbrt(len(P))
    # l(r) = 0.2 n(r)^(-1/3) growing with radius, tabulated on r^2; the
    # smallest l keeps the sub-cell path on (<= 12 sub-cells per axis)
    r2max = 3 * R * R
    r = np.sqrt(np.linspace(0, r2max, 257))
    l = 0.2 * lc * (0.8 + (r / R) ** 0.8 * 0.45)
    tab = l ** 2
    vl = L.VarLinkTable([0.0, 0.0, 0.0], r2max, tab)
    dom = L.domain([-R] * 3, [R] * 3, [-R] * 3, [R] * 3, [lc] * 3, 0.0, 5, vl)
    ovl = O.varlink([0.0, 0.0, 0.0], r2max, tab)
    od = odomain(dom)
    lab, nxt, halos, st = gpu_link(ctx, P, dom, T=12)  # small clumps: low threshold
    olab, _ = O.link_uf(P["Pos"], od, ovl)
    assert np.array_equal(lab, olab)
    assert np.array_equal(O.link_brute(P["Pos"][:6000], od, ovl),
                          O.link_uf(P["Pos"][:6000], od, ovl)[0])
    assert st.ndense > 0 and len(np.unique(lab)) < len(lab) * 0.95


def test_many_volumes(ctx):
    P, vols = W.volumes_host(48, median=3000, lo=300, hi=30000, seed=3)
    n = len(P)
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    halos = np.empty(n, L.HALO_DTYPE)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    st = L.Stats()
    L.check(ctx.link_batch(L.particles(P), vols, L.opts(), nxt, lab, halos, n, hoff, st))
    for v, vol in enumerate(vols):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        od = odomain(vol)
        olab, _ = O.link_uf(P["Pos"][a:b], od)
        assert np.array_equal(lab[a:b], olab), v
        oh = O.retrieve(O.ring_from_labels(olab), P["Pos"][a:b], P["Vel"][a:b], od)
        hv = halos[int(hoff[v]):int(hoff[v + 1])]
        assert np.array_equal(hv["head"], oh["head"]) and np.array_equal(hv["sum_vel"], oh["sum_vel"])


@pytest.mark.parametrize("flags", [0, L.OPT_SPARSE])
def test_tiny_volumes(ctx, flags):
    # 70 volumes of 1-90 particles: one wave of 64 consecutive particles
    # spans several volumes, so every per-wave volume lookup searches
    rng = np.random.default_rng(17)
    sizes = rng.integers(1, 91, 70)
    parts, vols, off = [], [], 0
    for k, m in enumerate(sizes):
        box = 3.0 + (k % 5)
        pos = rng.random((m, 3)) * box
        pos[: m // 2] = box / 2 + rng.normal(0, 0.15, (m // 2, 3))  # a clump in half of it
        pos = np.clip(pos, 0, box * 0.999)
        q = np.zeros(m, L.PART_DTYPE)
        q["Pos"] = pos.astype(np.float32)
        q["Vel"] = rng.normal(0, 1, (m, 3)).astype(np.float32)
        parts.append(q)
        v = L.Volume()
        v.offset, v.n = off, m
        v.domain = L.domain([0] * 3, [box] * 3, [0] * 3, [box] * 3, [0.5] * 3, 0.25 + 0.01 * (k % 3), 2)
        vols.append(v)
        off += m
    P = np.concatenate(parts)
    n = len(P)
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    halos = np.empty(n, L.HALO_DTYPE)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    L.check(ctx.link_batch(L.particles(P), vols, L.opts(flags, 0.0, 4), nxt, lab, halos, n, hoff))
    for v, vol in enumerate(vols):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        od = odomain(vol)
        olab, _ = O.link_uf(P["Pos"][a:b], od)
        assert np.array_equal(lab[a:b], olab), v
        oring = O.ring_from_labels(olab)
        assert np.array_equal(nxt[a:b].astype(np.int64) - 1, oring.astype(np.int64)), v
        oh = O.retrieve(oring, P["Pos"][a:b], P["Vel"][a:b], od)
        hv = halos[int(hoff[v]):int(hoff[v + 1])]
        assert len(hv) == len(oh), v
        assert np.array_equal(hv["head"], oh["head"]) and np.array_equal(hv["sum_pos"], oh["sum_pos"])


@pytest.mark.parametrize("scale,b", [(32, 0.2), (20, 0.6), (24, 0.1)])
def test_sparse_cells(ctx, scale, b):
    bx = W.plain(scale, seed=15, b=b)
    _, P, vols = W.host_domain(bx)
    lab, st = check_against_oracle(ctx, P, vols[0].domain, flags=L.OPT_SPARSE)
    assert st.ncells > 0


def test_sparse_variable_linking_length(ctx):
    R = 20.0
    P = lightcone_like(30000, R, 5)
    lc = 2 * R / np.cbrt(len(P))
    r2max = 3 * R * R
    r = np.sqrt(np.linspace(0, r2max, 129))
    tab = (0.2 * lc * (0.8 + (r / R) ** 0.8 * 0.45)) ** 2
    vl = L.VarLinkTable([0.0, 0.0, 0.0], r2max, tab)
    dom = L.domain([-R] * 3, [R] * 3, [-R] * 3, [R] * 3, [lc] * 3, 0.0, 5, vl)
    lab, _, _, st = gpu_link(ctx, P, dom, flags=L.OPT_SPARSE, T=12)
    olab, _ = O.link_uf(P["Pos"], odomain(dom), O.varlink([0.0, 0.0, 0.0], r2max, tab))
    assert np.array_equal(lab, olab)
    assert st.ndense > 0


def test_sparse_batch(ctx):
    P, vols = W.volumes_host(24, median=2000, lo=300, hi=20000, seed=9)
    n = len(P)
    outs = []
    for fl in (0, L.OPT_SPARSE):
        lab = np.empty(n, np.uint32)
        nxt = np.empty(n, np.uint32)
        halos = np.empty(n, L.HALO_DTYPE)
        hoff = np.zeros(len(vols) + 1, np.uint64)
        L.check(ctx.link_batch(L.particles(P), vols, L.opts(fl), nxt, lab, halos, n, hoff))
        outs.append((lab, nxt, halos[: int(hoff[-1])], hoff))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
    for v, vol in enumerate(vols[:6]):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        olab, _ = O.link_uf(P["Pos"][a:b], odomain(vol))
        assert np.array_equal(outs[1][0][a:b], olab)


@pytest.mark.parametrize("flags", [0, L.OPT_SPARSE])
def test_lightcone_shells(ctx, flags):
    lc = W.Lightcone(n=150000, nshell=8, seed=3)
    P0 = W.lightcone_particles(lc)
    P, vols, bounds, vt = W.lightcone_shells(lc, P0)
    n = len(P)
    lab = np.empty(n, np.uint32)
    halos = np.empty(n, L.HALO_DTYPE)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    st = L.Stats()
    L.check(ctx.link_batch(L.particles(P), vols, L.opts(flags), None, lab, halos, n, hoff, st))
    assert st.ncells > 0
    r2max, tab = lc.varlink_table()
    ovl = O.varlink([0.0, 0.0, 0.0], r2max, tab)
    for v, vol in enumerate(vols):
        a, b = int(vol.offset), int(vol.offset + vol.n)
        olab, _ = O.link_uf(P["Pos"][a:b], odomain(vol), ovl)
        assert np.array_equal(lab[a:b], olab), v


def test_varlink_subcells_follow_occupied_cells(ctx):
    # The outermost lightcone shell: its table reaches r = 0, where l is
    # smallest (about 1.6x below l at the shell), which would call for more
    # than 12 sub-cells per axis; only occupied cells are home cells, so the
    # sub-cell size follows their smallest l and dense cells keep sub-cells.
    lc = W.Lightcone(n=150000, nshell=8, seed=3)
    P0 = W.lightcone_particles(lc)
    P, vols, bounds, vt = W.lightcone_shells(lc, P0)
    v = vols[-1]
    a, b = int(v.offset), int(v.offset + v.n)
    Q = P[a:b].copy()
    lab, nxt, halos, st = gpu_link(ctx, Q, v.domain, T=12)
    assert st.ndense > 0
    r2max, tab = lc.varlink_table()
    olab, _ = O.link_uf(Q["Pos"], odomain(v), O.varlink([0.0, 0.0, 0.0], r2max, tab))
    assert np.array_equal(lab, olab)


@pytest.mark.parametrize("T,flags,l", [(16, 0, 0.2), (24, 0, 0.2), (16, L.OPT_SPARSE, 0.2),
                                       (16, 0, 0.05), (16, L.OPT_SPARSE, 0.05)])
def test_small_dense_cell_sizes(ctx, T, flags, l):
    # one cell per size around the boundaries of the all-pairs kernel for
    # dense cells below 256 particles (64-particle tiles, 8-particle blocks)
    # and of the sub-cell run kernel from 256: half of each cell's particles
    # in tight blobs (cliques), half uniform at about one linking length
    # apart, so cells hold several components that neighbours merge and the
    # neighbour queues overflow; with sparse addressing the smaller dense
    # cells are sub-sorted for the sparse kernel's run scan.  At l = 0.05,
    # given as a (constant) variable-l table so that the reference grid is
    # kept, the volume has no sub-cells (m > 12): the cells below 256
    # particles still go to the all-pairs kernel, larger ones to the sparse
    # kernel
    rng = np.random.default_rng(31)
    sizes = [16, 23, 24, 25, 63, 64, 65, 71, 127, 128, 129, 191, 192, 193, 254, 255, 256, 257]
    nx, ny, nz, lc = 4, 4, 3, 1.0
    cells = [(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)]
    pos = []
    for c, (i, j, k) in enumerate(cells):
        n = sizes[c % len(sizes)]
        lo = np.array([i, j, k], np.float64) * lc
        nb = n // 2
        centres = lo + 0.2 + rng.random((3, 3)) * 0.6
        blob = centres[rng.integers(0, 3, nb)] + rng.normal(0, 0.03, (nb, 3))
        uni = lo + rng.random((n - nb, 3)) * lc
        pos.append(np.clip(np.concatenate([blob, uni]), lo, lo + lc * 0.999))
    pos = np.concatenate(pos)
    n = len(pos)
    P = np.zeros(n, L.PART_DTYPE)
    P["Pos"] = pos[rng.permutation(n)].astype(np.float32)
    P["Vel"] = rng.normal(0, 1, (n, 3)).astype(np.float32)
    P["ID"] = np.arange(n)
    box = [nx * lc, ny * lc, nz * lc]
    vl = ovl = None
    if l < 0.1:
        r2max = float(np.sum(np.square(box)))
        tab = np.full(8, l * l)
        vl = L.VarLinkTable([0.0, 0.0, 0.0], r2max, tab)
        ovl = O.varlink([0.0, 0.0, 0.0], r2max, tab)
    dom = L.domain([0] * 3, box, [0] * 3, box, [lc] * 3, l, 5, vl)
    lab, st = check_against_oracle(ctx, P, dom, T=T, vl=vl, ovl=ovl, flags=flags)
    assert st.ndense > 0 and st.pair_tests_kernel[2] > 0
    assert (st.pair_tests_kernel[1] > 0) == (l > 0.1)  # runs only with sub-cells


@pytest.mark.parametrize("coarse", [2.5, 6.0])
def test_coarse_grid_refined(ctx, coarse):
    # reference cells several linking lengths across (m > 12 on the
    # reference grid): with a constant l the engine links on a finer grid of
    # its own; the oracle sweeps the reference grid, and the groups agree
    bx = W.plain(24, seed=8)
    _, P, vols = W.host_domain(bx)
    d = vols[0].domain
    lc = [d.lcell[k] * coarse for k in range(3)]
    dom = L.domain(list(d.rmin), list(d.rmax), list(d.rmin_buff), list(d.rmax_buff), lc,
                   d.linklength, d.nphalomin)
    lab, st = check_against_oracle(ctx, P, dom)
    ref_cells = np.prod([np.ceil((d.rmax_buff[k] - d.rmin_buff[k]) / lc[k]) + 1 for k in range(3)])
    assert st.ncells > ref_cells and st.ndense > 0


@pytest.mark.parametrize("nph", [0.5, 1.0, 2.0])
def test_size_cut_of_one(ctx, nph):
    """Nphalomin <= 1 keeps groups of one: each is a halo and its own ring."""
    bx = W.plain(20, seed=11, nphalomin=nph)
    _, P, vols = W.host_domain(bx)
    check_against_oracle(ctx, P, vols[0].domain)


def test_size_cut_of_one_batched(ctx):
    """Volumes keeping groups of one beside volumes that do not, in one call."""
    bx = W.plain(24, seed=53)
    recs = L.gen_host(bx.gen())
    P, _, vols = L.ghosts_host(bx.ghost(2), recs)
    for k, v in enumerate(vols):
        v.domain.nphalomin = 1.0 if k % 2 == 0 else 20.0
    n = len(P)
    nxt, lab = np.empty(n, np.uint32), np.empty(n, np.uint32)
    hal = np.empty(n, L.HALO_DTYPE)
    hoff = np.zeros(len(vols) + 1, np.uint64)
    L.check(ctx.link_batch(L.particles(P), vols, L.opts(), nxt, lab, hal, n, hoff))
    for k, v in enumerate(vols):
        a, b = v.offset, v.offset + v.n
        Pv = P[a:b]
        od = odomain(v.domain)
        olab, _ = O.link_uf(Pv["Pos"], od, None)
        assert np.array_equal(lab[a:b], olab), k
        oring = O.ring_from_labels(olab)
        assert np.array_equal(nxt[a:b].astype(np.int64) - 1, oring.astype(np.int64)), k
        oh = O.retrieve(oring, Pv["Pos"], Pv["Vel"], od)
        h = hal[int(hoff[k]):int(hoff[k + 1])]
        assert len(h) == len(oh), k
        for f in ("head", "n", "owned", "edge_warn", "sum_pos", "sum_vel"):
            assert np.array_equal(h[f], oh[f]), (k, f)
