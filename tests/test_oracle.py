"""Pin the CPU oracle independently of its own formulas.

* predicate: exact ties from Pythagorean triples/quadruples at l = 5/16 and
  25/128 (d^2 == l^2 must link, src_v2/FOF.c:357), one-ulp offsets, and a
  brute-force evaluation in exact rational arithmetic;
* linkers: brute force O(n^2) vs cell union-find vs the faithful ring
  restatement; permutation invariance; lattices (a <= l links, a > l does
  not); chains; the reference's candidate restriction when l >= lcell;
* retrieval: heads, size cut, ownership and sums against direct numpy /
  math.fsum computations;
* error paths: particles outside the buffered box and non-finite positions.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def dom(lo, hi, lcell, l, nmin=2, owned=None):
    lo3, hi3 = [lo] * 3, [hi] * 3
    o0, o1 = owned if owned else (lo3, hi3)
    return O.make_domain(o0, o1, lo3, hi3, [lcell] * 3, l, nmin)


def exact_linked(pi, pj, l):
    """Reference predicate in exact rationals: dx = fl32(xi - xj) then the two
    FP64 additions are the only roundings; evaluate them exactly rounded."""
    d = [float(np.float32(np.float32(a) - np.float32(b))) for a, b in zip(pi, pj)]
    s = float(Fraction(d[0]) ** 2 + Fraction(d[1]) ** 2)  # round-to-nearest of exact sum
    s = float(Fraction(s) + Fraction(d[2]) ** 2)
    return not (s > l * l)


@pytest.mark.parametrize("l,tri", [(5 / 16, (3, 4, 0, 5)), (25 / 128, (3, 4, 0, 5)), (25 / 128, (15, 20, 0, 25)),
                                   (0.375, (2, 2, 1, 3)), (0.75, (1, 2, 2, 3))])
def test_pair_exact_ties(l, tri):
    a, b, c, h = tri
    for base in (0.0, 1.0, 100.0, 512.0, 999.0):
        pi = np.array([base, base, base], np.float32)
        pj = pi + np.array([a * l / h, b * l / h, c * l / h], np.float32)
        d2 = sum(float(np.float32(x - y)) ** 2 for x, y in zip(pi, pj))
        if d2 != l * l:
            continue  # offset not representable at this base
        assert O.pair_linked(pi, pj, l * l)
        q = pj.copy()
        q[0] = np.nextafter(q[0], np.float32(1e9))
        assert not O.pair_linked(pi, q, l * l)
        assert O.pair_linked(pi, q, l * l) == exact_linked(pi, q, l)


def test_pair_random_vs_exact():
    rng = np.random.default_rng(0)
    l = 25 / 128
    for _ in range(3000):
        pi = (rng.random(3) * 1000).astype(np.float32)
        dirn = rng.normal(size=3)
        pj = (pi + dirn / np.linalg.norm(dirn) * l * (1 + rng.normal() * 1e-6)).astype(np.float32)
        assert O.pair_linked(pi, pj, l * l) == exact_linked(pi, pj, l)


def test_pair_nan_and_duplicates():
    # the reference predicate links NaN separations (!(NaN > linksq)) ...
    assert O.pair_linked(np.array([np.nan, 0, 0], np.float32), np.zeros(3, np.float32), 0.01)
    # ... and coincident particles at l = 0
    assert O.pair_linked(np.ones(3, np.float32), np.ones(3, np.float32), 0.0)


def random_set(seed, n=1500, L=10.0, nclump=10, sig=0.15):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * L
    c = rng.random((nclump, 3)) * L
    k = n // 2
    pos[:k] = c[rng.integers(0, nclump, k)] + rng.normal(0, sig, (k, 3))
    return np.clip(pos, 0, L).astype(np.float32)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("lfac", [0.2, 0.5, 1.0])
def test_linkers_agree(seed, lfac):
    pos = random_set(seed)
    L, n = 10.0, len(pos)
    lcell = L / np.cbrt(n)
    d = dom(0.0, L, lcell, lfac * lcell)
    lb = O.link_brute(pos, d)
    lu, pu = O.link_uf(pos, d)
    nx, pr = O.link_ring(pos, d)
    assert np.array_equal(lb, lu)
    assert np.array_equal(O.labels_from_ring(nx), lu)
    assert pu == pr  # same candidate set, same number of predicate calls


def test_candidate_restriction_l_above_lcell():
    # with l > lcell the reference only tests adjacent cells: two particles
    # at distance 1.5 lcell < l in cells two apart stay unlinked
    pos = np.array([[0.1, 0.5, 0.5], [1.6, 0.5, 0.5]], np.float32)
    d = dom(0.0, 4.0, 0.75, 2.0)
    assert O.link_uf(pos, d)[0][1] == 1
    assert O.link_brute(pos, d)[1] == 1
    assert O.labels_from_ring(O.link_ring(pos, d)[0])[1] == 1
    d2 = dom(0.0, 4.0, 2.0, 2.0)
    assert O.link_uf(pos, d2)[0][1] == 0


def canonical_partition(lab):
    groups = {}
    for i, g in enumerate(lab):
        groups.setdefault(int(g), []).append(i)
    return sorted(tuple(sorted(v)) for v in groups.values())


def test_permutation_invariance():
    pos = random_set(7)
    d = dom(0.0, 10.0, 10 / np.cbrt(len(pos)), 0.15)
    lab, _ = O.link_uf(pos, d)
    perm = np.random.default_rng(1).permutation(len(pos))
    lab2, _ = O.link_uf(pos[perm], d)
    # map back: group of perm[i] in the first run == group of i in the second
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    p1 = canonical_partition(lab)
    p2 = canonical_partition(lab2[inv])
    assert p1 == p2
    # labels are group minima
    for g in p1:
        assert all(lab[i] == g[0] for i in g)


@pytest.mark.parametrize("a,linked", [(0.25, True), (0.3125, True), (0.3126, False)])
def test_lattice(a, linked):
    k = 6
    g = np.arange(k) * a + 1.0
    pos = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.astype(np.float32)
    l = 0.3125
    d = dom(0.0, 4.0, 1.5625, l)
    lab, _ = O.link_uf(pos, d)
    spacing_exact = all(np.float32(x + a) - np.float32(x) == np.float32(a) for x in g[:-1])
    if linked and spacing_exact:
        assert len(np.unique(lab)) == 1
    if not linked:
        assert len(np.unique(lab)) == len(pos)
    assert np.array_equal(lab, O.link_brute(pos, d))


def test_chain_links_transitively():
    x = np.arange(200) * 0.19 + 0.5
    pos = np.stack([x, np.full_like(x, 5.0), np.full_like(x, 5.0)], 1).astype(np.float32)
    d = dom(0.0, 50.0, 1.0, 0.2)
    lab, _ = O.link_uf(pos, d)
    assert np.all(lab == 0)
    pos[100, 1] += 0.5  # lift one link out of the line: 99 and 101 are 0.38 > l apart
    lab, _ = O.link_uf(pos, d)
    assert len(np.unique(lab)) == 3


def test_ring_from_labels_and_retrieve():
    pos = random_set(3)
    vel = np.random.default_rng(3).normal(0, 100, pos.shape).astype(np.float32)
    L = 10.0
    d = dom(0.0, L, L / np.cbrt(len(pos)), 0.12, nmin=5, owned=([2.0] * 3, [8.0] * 3))
    lab, _ = O.link_uf(pos, d)
    ring = O.ring_from_labels(lab)
    # ascending order inside each group, closing at the minimum
    for i in range(len(lab)):
        j = ring[i]
        assert lab[j] == lab[i]
        assert j > i or j == lab[i]
    h = O.retrieve(ring, pos, vel, d)
    sizes = np.bincount(lab, minlength=len(lab))
    heads = np.flatnonzero(sizes >= 5)
    assert np.array_equal(h["head"], heads)
    assert np.array_equal(h["n"], sizes[heads])
    for rec in h:
        mem = np.flatnonzero(lab == rec["head"])
        exact = np.array([math.fsum(pos[mem, k].astype(np.float64)) for k in range(3)])
        assert np.allclose(rec["sum_pos"], exact, rtol=1e-13, atol=0)
        com = rec["sum_pos"] / rec["n"]
        assert rec["owned"] == int(np.all((com >= 2.0) & (com < 8.0)))
        assert np.allclose(rec["out"][:3], (rec["sum_pos"] * (1.0 / rec["n"])).astype(np.float32))
    # the reference ring gives the same halos up to summation order
    nx, _ = O.link_ring(pos, d)
    h2 = O.retrieve(nx, pos, vel, d)
    assert np.array_equal(h2["head"], h["head"]) and np.array_equal(h2["n"], h["n"])
    assert np.allclose(h2["sum_pos"], h["sum_pos"], rtol=1e-13)


def test_props_compensated_vs_fsum():
    rng = np.random.default_rng(5)
    n = 5000
    lab = (rng.integers(0, 20, n) * 7).astype(np.uint32)
    lab = np.minimum(lab, np.arange(n, dtype=np.uint32))
    pos = (rng.random((n, 3)) * 1e3).astype(np.float32)
    vel = rng.normal(0, 300, (n, 3)).astype(np.float32)
    heads = np.unique(lab)
    sp, sv = O.props_compensated(lab, pos, vel, heads)
    for k, h in enumerate(heads):
        m = lab == h
        for q in range(3):
            assert sp[k, q] == math.fsum(pos[m, q].astype(np.float64))
            assert sv[k, q] == math.fsum(vel[m, q].astype(np.float64))


def test_domain_errors():
    pos = np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5]], np.float32)
    d = dom(0.0, 1.0, 0.5, 0.1)
    for fn in ("ring", "uf", "brute"):
        assert O.link_rc(fn, pos, d) == -2
    pos[1] = [np.nan, 0.5, 0.5]
    for fn in ("ring", "uf", "brute"):
        assert O.link_rc(fn, pos, d) == -2


def test_varlink_cell_rule():
    # l^2 table linear in r^2 is reproduced exactly at cell centres
    d = dom(-5.0, 5.0, 1.0, 0.1)
    r2max = 300.0
    tab = 0.01 + 1e-4 * np.linspace(0, r2max, 301)
    vl = O.varlink([0.0, 0.0, 0.0], r2max, tab)
    for ix, iy, iz in [(0, 0, 0), (5, 5, 5), (3, 7, 9), (11, 11, 11)]:
        c = np.array([-5.0 + ix + 0.5, -5.0 + iy + 0.5, -5.0 + iz + 0.5])
        r2 = float((c ** 2).sum())
        expect = tab[-1] if r2 >= r2max else 0.01 + 1e-4 * r2
        assert abs(O.cell_linksq(d, vl, ix, iy, iz) - expect) < 1e-15
    # the three linkers agree under the variable rule
    pos = random_set(11, n=800, L=10.0) - np.float32(5.0)
    tab2 = np.linspace(0.3, 0.05, 64) ** 2
    vl2 = O.varlink([0.0, 0.0, 0.0], 75.0, tab2)
    d2 = dom(-5.0, 5.0, 10 / np.cbrt(800), 0.3)
    lb = O.link_brute(pos, d2, vl2)
    lu, _ = O.link_uf(pos, d2, vl2)
    nx, _ = O.link_ring(pos, d2, vl2)
    assert np.array_equal(lb, lu) and np.array_equal(O.labels_from_ring(nx), lu)
    assert len(np.unique(lu)) < len(lu)


# ---- second moments (src_v2/FOF.c:811-824, :848-860) -----------------------
def test_moments_closed_form():
    # +-a along each axis about c: centred xx,yy,zz = 2a_k^2, cross terms 0;
    # raw sums = n c_a c_b + centred (c and a exact binary fractions)
    c = np.array([250.5, -17.25, 1000.125])
    a = [1.5, 0.25, 3.0]
    pos = []
    for k in range(3):
        for s in (1, -1):
            p = c.copy()
            p[k] += s * a[k]
            pos.append(p)
    pos = np.array(pos, np.float32)
    vel = (pos - c.astype(np.float32)) * np.float32(2.0)
    lab = np.zeros(6, np.uint32)
    raw, cen = O.moments(O.ring_from_labels(lab), pos, vel, [0])
    exp_c = [2 * a[0] ** 2, 2 * a[1] ** 2, 2 * a[2] ** 2, 0.0, 0.0, 0.0]
    assert np.array_equal(cen[0, :6], exp_c)
    assert np.array_equal(cen[0, 6:], 4 * np.array(exp_c))
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for q, (i, j) in enumerate(pairs):
        assert raw[0, q] == 6 * c[i] * c[j] + exp_c[q]


def test_moments_vs_exact_rationals():
    # raw: FP64 sums in ring order agree with the exact sums to FP64 rounding;
    # centred: the compensated two-pass value is the exact one to a few ulps
    from fractions import Fraction as Fr

    rng = np.random.default_rng(17)
    n = 600
    grp = rng.integers(0, 12, n)
    first = np.full(12, n)
    np.minimum.at(first, grp, np.arange(n))
    lab = first[grp].astype(np.uint32)  # group minimum index, as the linkers label
    pos = (500.0 + rng.normal(0, 0.3, (n, 3))).astype(np.float32)
    vel = rng.normal(0, 300, (n, 3)).astype(np.float32)
    heads = np.unique(lab)
    raw, cen = O.moments(O.ring_from_labels(lab), pos, vel, heads)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for k, h in enumerate(heads):
        m = np.flatnonzero(lab == h)
        for blk, arr in ((0, pos), (1, vel)):
            X = [[Fr(float(arr[i, d])) for d in range(3)] for i in m]
            mean = [sum(x[d] for x in X) / len(X) for d in range(3)]
            for q, (i, j) in enumerate(pairs):
                exact_raw = sum(x[i] * x[j] for x in X)
                exact_cen = sum((x[i] - mean[i]) * (x[j] - mean[j]) for x in X)
                scale = float(abs(exact_raw)) or 1.0
                assert abs(raw[k, 6 * blk + q] - float(exact_raw)) <= 4 * len(m) * 1.2e-16 * scale
                cs = float(sum(abs(x[i] - mean[i]) * abs(x[j] - mean[j]) for x in X)) or 1.0
                assert abs(cen[k, 6 * blk + q] - float(exact_cen)) <= 1e-14 * cs


def test_moments_reference_formulas():
    # two particles at x = 1 and 3 (y = z = 0), velocities 2 and 6 on x:
    # sums xh = 4, xh2 = 10 -> Iyy = Izz = xh2 - xh^2/2 = 2, Ixx = 0; dispersion
    # vx: (40 - 64/2)/2 = 4
    sp = np.array([4.0, 0, 0])
    sv = np.array([8.0, 0, 0])
    raw = np.zeros(12)
    raw[0] = 10.0
    raw[6] = 40.0
    ine, dis = O.moments_reference_out(2, sp, sv, raw)
    assert list(ine) == [0.0, 2.0, 2.0, 0.0, 0.0, 0.0]
    assert list(dis) == [4.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def test_minid_labels_order_free():
    # canonical labels (group minimum particle ID) do not depend on the input
    # order, while minimum-index labels do
    rng = np.random.default_rng(23)
    pos = random_set(23, n=3000, L=12.0)
    ids = rng.permutation(10 ** 6)[:3000].astype(np.uint64)
    d = dom(0.0, 12.0, 12 / np.cbrt(3000), 0.3)
    lab, _ = O.link_uf(pos, d)
    can = O.minid_labels(lab, ids)
    perm = rng.permutation(3000)
    lab2, _ = O.link_uf(pos[perm], d)
    can2 = O.minid_labels(lab2, ids[perm])
    assert np.array_equal(can[perm], can2)
    for g in np.unique(lab)[:50]:
        m = lab == g
        assert np.all(can[m] == ids[m].min())
