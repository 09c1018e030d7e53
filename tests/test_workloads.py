"""Synthetic inputs and ghost shells (host implementations, no GPU)."""
import numpy as np

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import workloads as W


def test_generator_deterministic_and_shaped():
    bx = W.plain(16, seed=3)
    a = L.gen_host(bx.gen())
    b = L.gen_host(bx.gen())
    assert np.array_equal(a, b)
    assert np.array_equal(np.sort(a["ID"]), np.arange(bx.n))  # IDs: a permutation
    assert not np.array_equal(a["ID"], np.arange(bx.n))       # shuffled order
    assert a["Pos"].min() >= 0 and a["Pos"].max() < bx.box
    c = L.gen_host(W.plain(16, seed=4).gen())
    assert not np.array_equal(a["Pos"], c["Pos"])
    v = a["Vel"].reshape(-1)
    assert 250 < v.std() < 350 and abs(v.mean()) < 20


def test_clump_fraction_and_sizes():
    bx = W.plain(32, seed=1)
    recs = L.gen_host(bx.gen())
    # half of the generation indices are clump members: they sit in compact clumps
    ids = recs["ID"]
    clumped = recs[ids < bx.n // 2]["Pos"]
    bg = recs[ids >= bx.n // 2]["Pos"]
    d = O.make_domain([0] * 3, [bx.box] * 3, [0] * 3, [bx.box] * 3, [bx.box / 32] * 3, bx.linklength, 20)
    lab_c, _ = O.link_uf(clumped, d)
    lab_b, _ = O.link_uf(bg, d)
    sizes = np.bincount(lab_c)
    assert (sizes >= 20).sum() > 0.8 * bx.nclumps  # most clumps are halos
    assert np.bincount(lab_b).max() < 10           # the background does not percolate


def test_borderline_pairs():
    bx = W.precision(pside=16)
    recs = L.gen_host(bx.gen())
    g = bx.gen()
    npairs = int(g.pair_frac * g.n / 2)
    base = g.n - 2 * npairs
    by_id = np.empty(g.n, np.int64)
    by_id[recs["ID"]] = np.arange(g.n)
    l2 = bx.linklength ** 2
    rel = []
    for q in range(npairs):
        a, b = recs[by_id[base + 2 * q]]["Pos"], recs[by_id[base + 2 * q + 1]]["Pos"]
        fx = (a - b).astype(np.float32).astype(np.float64)
        d2 = (fx[0] * fx[0] + fx[1] * fx[1]) + fx[2] * fx[2]
        rel.append(d2 / l2 - 1)
    rel = np.abs(np.array(rel))
    assert npairs > 10 and rel.max() < bx.pair_rel


def test_ghost_shell_semantics():
    bx = W.plain(16, seed=2)
    recs = L.gen_host(bx.gen())
    for g in (1, 2, 3):
        out, counts, vols = L.ghosts_host(bx.ghost(g), recs)
        assert counts.sum() == len(out)
        owned_total = 0
        for v in vols:
            P = out[int(v.offset):int(v.offset + v.n)]
            p = P["Pos"].astype(np.float64)
            lo = np.array(v.domain.rmin_buff[:]), np.array(v.domain.rmax_buff[:])
            assert np.all(p >= lo[0]) and np.all(p <= lo[1])
            own = np.all((p >= np.array(v.domain.rmin[:])) & (p < np.array(v.domain.rmax[:])), 1)
            owned_total += own.sum()
            # every image is the owner's position shifted by 0 or +-box per axis
            src = recs[np.argsort(recs["ID"])][P["ID"]]["Pos"].astype(np.float64)
            s = (p - src) / bx.box
            assert np.all(np.isin(np.round(s, 6), [-1.0, 0.0, 1.0]))
            # a particle never appears twice at the same place in one volume
            key = np.rec.fromarrays([P["ID"], P["Pos"][:, 0], P["Pos"][:, 1], P["Pos"][:, 2]])
            assert len(np.unique(key)) == len(P)
        assert owned_total == len(recs)  # each particle owned exactly once


def test_comoving_distance_closed_form():
    # Einstein-de Sitter (om = 1): r(z) = 2 c/H0 (1 - 1/sqrt(1+z))
    z, r = W.comoving_distance_table(om=1.0, zmax=3.0, nz=30001)
    exact = 2 * W.C_OVER_H0 * (1 - 1 / np.sqrt(1 + z))
    assert np.abs(r - exact).max() < 1e-6 * exact.max()


def test_lightcone_density_and_shells():
    lc = W.Lightcone(n=100000, nshell=8, seed=2)
    rr = np.linspace(0, lc.rmax, 20001)
    integral = np.trapz(4 * np.pi * rr ** 2 * lc.density(rr), rr)
    assert abs(integral / lc.n - 1) < 1e-4
    assert lc.z_of_r(0.0) == 0.0 and 1.4 < lc.z_of_r(3000.0) < 1.55  # flat LCDM, om 0.31
    P = W.lightcone_particles(lc)
    r = np.linalg.norm(P["Pos"].astype(np.float64), axis=1)
    assert r.max() < lc.rmax and np.array_equal(np.sort(P["ID"]), np.arange(lc.n))
    Q, vols, bounds, vt = W.lightcone_shells(lc, P)
    owned = 0
    for v, (lo, hi) in zip(vols, bounds):
        rq = np.linalg.norm(Q["Pos"][int(v.offset):int(v.offset + v.n)].astype(np.float64), axis=1)
        assert rq.min() >= lo - lc.ghost and rq.max() < hi + lc.ghost
        owned += ((rq >= lo) & (rq < hi)).sum()
        assert v.domain.linklength < v.domain.lcell[0]  # l < lcell inside every shell
    assert owned == lc.n


def test_lightcone_device_generator_and_shells():
    # the device generator (torch, here on the CPU) follows the host model:
    # IDs a permutation, background radii distributed as n(r) r^2, velocities
    # N(0, 300); the device shell builder selects what the host one does
    import torch

    lc = W.Lightcone(n=400000, nshell=8, seed=3)
    R = W.lightcone_particles_device(lc, torch.device("cpu"))
    P = R.numpy().view(L.PART_DTYPE)
    assert np.array_equal(np.sort(P["ID"]), np.arange(lc.n))
    r = np.linalg.norm(P["Pos"].astype(np.float64), axis=1)
    assert r.max() < lc.rmax
    bg = np.sort(r[P["ID"] < lc.n // 2])
    u = (np.arange(len(bg)) + 0.5) / len(bg)
    assert np.abs(bg - lc.sample_r(u)).max() < 0.02 * lc.rmax  # Kolmogorov-type bound
    assert abs(P["Vel"].std() - 300) < 3
    D, vols, bounds, _ = W.lightcone_shells_device(lc, R)
    Q, vh, bh, _ = W.lightcone_shells(lc, P)
    assert [int(v.n) for v in vols] == [int(v.n) for v in vh] and bounds == bh
    Dq = D.numpy().view(L.PART_DTYPE)
    for v, w in zip(vols, vh):
        a = np.sort(Dq[int(v.offset):int(v.offset + v.n)]["ID"])
        b = np.sort(Q[int(w.offset):int(w.offset + w.n)]["ID"])
        assert np.array_equal(a, b)


def test_ghost_shell_complete():
    # every periodic image (x + s box, s in {-1,0,1} per axis, rounded to
    # float as src_v2/read_data.c:792-797 does) that lies in a volume's
    # buffered box is present in that volume, and nothing else is: the
    # builder's output equals a brute-force enumeration, as multisets
    bx = W.plain(12, seed=3)
    recs = L.gen_host(bx.gen())
    pos = recs["Pos"].astype(np.float64)
    shifts = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], np.float64)
    imgs = (pos[:, None, :] + shifts[None, :, :] * bx.box).astype(np.float32)  # (n, 27, 3)
    ids = np.repeat(recs["ID"], 27)
    flat = imgs.reshape(-1, 3).astype(np.float64)
    for g in (1, 2, 3):
        out, counts, vols = L.ghosts_host(bx.ghost(g), recs)
        for v in vols:
            lo, hi = np.array(v.domain.rmin_buff[:]), np.array(v.domain.rmax_buff[:])
            m = np.all((flat >= lo) & (flat <= hi), 1)
            want = np.rec.fromarrays([ids[m], flat[m, 0].astype(np.float32), flat[m, 1].astype(np.float32),
                                      flat[m, 2].astype(np.float32)])
            P = out[int(v.offset):int(v.offset + v.n)]
            got = np.rec.fromarrays([P["ID"], P["Pos"][:, 0], P["Pos"][:, 1], P["Pos"][:, 2]])
            assert len(got) == len(want), (g, len(got), len(want))
            assert np.array_equal(np.sort(got), np.sort(want))
