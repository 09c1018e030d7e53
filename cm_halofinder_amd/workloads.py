"""Synthetic workloads of the FoF engine (clumps + background, ghost shells).

Generator rule (one choice, fixed here): half of the particles lie in
``nclumps`` Plummer clumps whose sizes are quantiles of dN/dM ~ M^-1.9 on
[20, 5000] rescaled so that they sum to the clumped count exactly (the
power law alone has mean ~133, below the ~262 per clump that "half of 64^3
in 500 clumps" implies, so the clump count and the clumped fraction are kept
and the sizes are stretched).  Scale radius a = 0.05 (size/100)^(1/3) Mpc/h,
radii truncated at the Plummer mass fraction 0.99.  The other half is
uniform.  Velocities are Gaussian with 300 km/s per axis.  IDs are the
generation index; the array order is a keyed random permutation.

Workloads (sizes per side P, box L in Mpc/h, b = linklength / (L/P)):

* ``plain``      L=100, P=64 (scale 1): one periodic domain + 3 Mpc/h shell
* ``plain_1024`` L=1000, P=1024: the same at 1024^3, 2,048,000 clumps
* ``precision``  plain_1024 + 1% of particles in planted pairs with
                 |d/l - 1| < 1e-6 under the reference arithmetic
* ``batched``    the 1024^3 box cut into g^3 subvolumes with 3 Mpc/h shells
* ``volumes``    independent periodic volumes with log-normal counts
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _lib as L

CLUMPS_PER_PARTICLE = 500.0 / 262144.0  # 500 clumps in 64^3


@dataclasses.dataclass
class Box:
    seed: int
    pside: int
    box: float
    b: float = 0.2
    nphalomin: float = 20.0
    shell: float = 3.0
    pair_frac: float = 0.0
    pair_rel: float = 2e-6  # on d^2/l^2, i.e. |d/l - 1| < 1e-6

    @property
    def n(self) -> int:
        return self.pside ** 3

    @property
    def linklength(self) -> float:
        return self.b * self.box / self.pside

    @property
    def nclumps(self) -> int:
        return int(round(self.n * CLUMPS_PER_PARTICLE))

    def gen(self) -> L.Gen:
        return L.gen_params(self.seed, self.n, self.box, self.nclumps, pair_frac=self.pair_frac,
                            pair_rel=self.pair_rel, pair_l=self.linklength)

    def ghost(self, g: int = 1) -> L.Ghost:
        return L.ghost_params(self.box, self.shell, g, g, g, self.pside, self.linklength, self.nphalomin)


def plain(scale: int = 64, seed: int = 1, box: float | None = None, **kw) -> Box:
    """Config C1 (scale 64 = 64^3 particles in 100 Mpc/h); other scales keep
    the mean density."""
    if box is None:
        box = 100.0 * scale / 64.0
    return Box(seed=seed, pside=scale, box=box, **kw)


def plain_1024(seed: int = 7) -> Box:
    return Box(seed=seed, pside=1024, box=1000.0)


def precision(pside: int = 1024, seed: int = 11) -> Box:
    return Box(seed=seed, pside=pside, box=1000.0 * pside / 1024.0, pair_frac=0.01)


def host_domain(bx: Box, g: int = 1):
    """Records (owned + shell, reference layout) and volumes, built on the host."""
    recs = L.gen_host(bx.gen())
    out, counts, vols = L.ghosts_host(bx.ghost(g), recs)
    return recs, out, vols


def volumes_spec(nvol: int = 4096, median: float = 6e4, sigma: float = 0.63, lo: float = 1e3,
                 hi: float = 2e6, seed: int = 5, sep: float = 0.9765625):
    """Config C5: per-volume (n, box, b) with log-normal counts clipped to
    [lo, hi], boxes scaled for a fixed mean separation, b ~ U(0.15, 0.25)."""
    rng = np.random.default_rng(seed)
    n = np.clip(np.exp(np.log(median) + sigma * rng.standard_normal(nvol)), lo, hi).astype(np.int64)
    b = rng.uniform(0.15, 0.25, nvol)
    boxes = sep * np.cbrt(n.astype(np.float64))
    return n, boxes, b


def volume_box(k: int, n: int, box: float, b: float, seed: int = 5, shell: float = 3.0) -> Box:
    pside = int(round(n ** (1.0 / 3.0)))
    bx = Box(seed=seed * 100003 + k, pside=pside, box=box, b=b, shell=min(shell, box / 4.1))
    return bx


def _one_volume(k, n, box, b, seed):
    bx = volume_box(k, int(n), float(box), float(b), seed)
    g = L.gen_params(bx.seed, int(n), bx.box, max(1, int(round(n * CLUMPS_PER_PARTICLE))),
                     pair_l=bx.linklength)
    recs = L.gen_host(g)
    gh = L.ghost_params(bx.box, bx.shell, 1, 1, 1, 1, bx.linklength, bx.nphalomin)
    out, counts, vv = L.ghosts_host(gh, recs)
    v = vv[0]
    # the reference cell of a periodic volume of n particles: box / n^(1/3)
    for d in range(3):
        v.domain.lcell[d] = bx.box / np.cbrt(float(n))
    return out, v


def volumes_host(nvol: int, median: float = 6e4, lo: float = 1e3, hi: float = 2e6, seed: int = 5,
                 workers: int = 1):
    """Independent periodic volumes (config C5 at full size with the
    defaults), each with its periodic shell, built on the host; the C
    generator releases the GIL, so ``workers`` threads run in parallel."""
    n, boxes, b = volumes_spec(nvol, median, lo=lo, hi=hi, seed=seed)
    if workers > 1:
        import concurrent.futures as cf

        with cf.ThreadPoolExecutor(workers) as ex:
            res = list(ex.map(lambda k: _one_volume(k, n[k], boxes[k], b[k], seed), range(nvol)))
    else:
        res = [_one_volume(k, n[k], boxes[k], b[k], seed) for k in range(nvol)]
    parts, vols, off = [], [], 0
    for out, v in res:
        v.offset = off
        off += len(out)
        parts.append(out)
        vols.append(v)
    return np.concatenate(parts), vols


# ---------------------------------------------------------------------------
# Lightcone: full-sky ball, radial shells, variable linking length
# ---------------------------------------------------------------------------
C_OVER_H0 = 2997.92458  # Mpc/h


def comoving_distance_table(om: float = 0.31, zmax: float = 6.0, nz: int = 60001):
    """(z, r(z)) for flat LCDM: r = c/H0 * int_0^z dz / sqrt(om (1+z)^3 + 1 - om)."""
    z = np.linspace(0.0, zmax, nz)
    ez = 1.0 / np.sqrt(om * (1 + z) ** 3 + 1 - om)
    r = C_OVER_H0 * np.concatenate([[0.0], np.cumsum(0.5 * (ez[1:] + ez[:-1]) * np.diff(z))])
    return z, r


@dataclasses.dataclass
class Lightcone:
    n: int = 400_000_000
    rmax: float = 3000.0
    nshell: int = 64
    om: float = 0.31
    b: float = 0.2
    seed: int = 17
    nphalomin: float = 20.0
    ntab: int = 4096

    def __post_init__(self):
        z, r = comoving_distance_table(self.om)
        self._z, self._r = z, r
        rr = np.linspace(0.0, self.rmax, 200001)
        zz = np.interp(rr, r, z)
        w = rr ** 2 * (1 + zz) ** -1.5
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(rr))])
        self.n0 = self.n / (4 * np.pi * cdf[-1])  # n(r) = n0 (1+z)^-1.5
        self._rr, self._cdf = rr, cdf / cdf[-1]

    def z_of_r(self, r):
        return np.interp(r, self._r, self._z)

    def density(self, r):
        return self.n0 * (1 + self.z_of_r(r)) ** -1.5

    def link(self, r):
        """l(r) = b n(r)^(-1/3)."""
        return self.b * self.density(r) ** (-1.0 / 3.0)

    def sample_r(self, u):
        return np.interp(u, self._cdf, self._rr)

    @property
    def ghost(self):
        # wider than any clump: radius <= 12.2 a with a <= 0.12 l (5000*2/100)^(1/3) < 0.6 l
        return 8.0 * float(self.link(self.rmax))

    def varlink_table(self):
        r2max = (self.rmax + self.ghost) ** 2 * 3.0
        r = np.sqrt(np.linspace(0.0, r2max, self.ntab))
        return r2max, self.link(np.minimum(r, self.rmax * 1.5)) ** 2

    def shells(self):
        """Shell edges at equal expected particle counts."""
        return self.sample_r(np.linspace(0.0, 1.0, self.nshell + 1))


def lightcone_particles(lc: Lightcone) -> np.ndarray:
    """Host generator (numpy): half background following n(r), half in
    Plummer clumps (same mass function) whose centres follow n(r), so the
    clump abundance tracks n(r); scale radius a = 0.12 l(r) (size/100)^(1/3),
    i.e. compact relative to the local linking length at every redshift."""
    rng = np.random.default_rng(lc.seed)
    n = lc.n
    nbg = n // 2
    out = np.zeros(n, L.PART_DTYPE)

    def dirs(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1)[:, None]

    rb = lc.sample_r(rng.random(nbg))
    pos = dirs(nbg) * rb[:, None]
    # clumps: quantile sizes of M^-1.9 on [20, 5000] rescaled to the clumped count
    ncl = n - nbg
    K = max(1, int(round(n * CLUMPS_PER_PARTICLE)))
    u = (np.arange(K) + 0.5) / K
    a1 = 1 - 1.9
    m = (20.0 ** a1 + u * (5000.0 ** a1 - 20.0 ** a1)) ** (1 / a1)
    sz = np.maximum(1, np.floor(m * ncl / m.sum())).astype(np.int64)
    sz[: ncl - sz.sum()] += 1 if ncl > sz.sum() else 0
    while sz.sum() > ncl:
        sz[np.argmax(sz)] -= 1
    rc = lc.sample_r(rng.random(K))
    cen = dirs(K) * rc[:, None]
    a = 0.12 * lc.link(rc) * (sz / 100.0) ** (1 / 3)
    who = np.repeat(np.arange(K), sz)
    uu = 0.99 * rng.random(len(who))
    rr = a[who] / np.sqrt(uu ** (-2 / 3) - 1)
    pc = cen[who] + dirs(len(who)) * rr[:, None]
    allp = np.concatenate([pos, pc])
    rad = np.linalg.norm(allp, axis=1)
    keep = rad < lc.rmax
    allp[~keep] *= (lc.rmax * 0.999 / rad[~keep])[:, None]  # pull the few escapees inside
    perm = rng.permutation(n)
    out["Pos"] = allp[perm].astype(np.float32)
    out["Vel"] = rng.normal(0, 300, (n, 3)).astype(np.float32)
    out["ID"] = perm
    return out


def _interp_t(x, xp, fp):
    """np.interp for torch tensors (xp increasing, clamped at the ends)."""
    import torch

    i = torch.searchsorted(xp, x).clamp(1, len(xp) - 1)
    x0, x1, y0, y1 = xp[i - 1], xp[i], fp[i - 1], fp[i]
    t = ((x - x0) / (x1 - x0)).clamp(0.0, 1.0)
    return y0 + t * (y1 - y0)


def lightcone_particles_device(lc: Lightcone, device):
    """The model of lightcone_particles generated on the device with torch's
    RNG (another realization of the same model), for sizes where the host
    version would take minutes.  Returns the 32-byte records as a uint8
    tensor (ID, Pos, Vel as in the reference record)."""
    import torch

    f64 = torch.float64
    g = torch.Generator(device=device)
    g.manual_seed(lc.seed)
    cdf = torch.tensor(lc._cdf, dtype=f64, device=device)
    rr = torch.tensor(lc._rr, dtype=f64, device=device)
    zr = torch.tensor(lc._r, dtype=f64, device=device)
    zz = torch.tensor(lc._z, dtype=f64, device=device)

    def sample_r(k):
        return _interp_t(torch.rand(k, generator=g, device=device, dtype=f64), cdf, rr)

    def link(r):
        z = _interp_t(r, zr, zz)
        return lc.b * (lc.n0 * (1 + z) ** -1.5) ** (-1.0 / 3.0)

    def dirs(k):
        v = torch.randn(k, 3, generator=g, device=device, dtype=f64)
        return v / v.norm(dim=1, keepdim=True)

    n = lc.n
    nbg = n // 2
    ncl = n - nbg
    pos = torch.empty(n, 3, dtype=f64, device=device)
    pos[:nbg] = dirs(nbg) * sample_r(nbg)[:, None]
    # clump sizes: the quantile rule of the host generator
    K = max(1, int(round(n * CLUMPS_PER_PARTICLE)))
    u = (np.arange(K) + 0.5) / K
    a1 = 1 - 1.9
    m = (20.0 ** a1 + u * (5000.0 ** a1 - 20.0 ** a1)) ** (1 / a1)
    sz = np.maximum(1, np.floor(m * ncl / m.sum())).astype(np.int64)
    sz[: ncl - sz.sum()] += 1 if ncl > sz.sum() else 0
    while sz.sum() > ncl:
        sz[np.argmax(sz)] -= 1
    szt = torch.tensor(sz, device=device)
    rc = sample_r(K)
    cen = dirs(K) * rc[:, None]
    a = 0.12 * link(rc) * (szt.to(f64) / 100.0) ** (1.0 / 3.0)
    who = torch.repeat_interleave(torch.arange(K, device=device), szt)
    uu = 0.99 * torch.rand(ncl, generator=g, device=device, dtype=f64)
    rad = a[who] / torch.sqrt(uu ** (-2.0 / 3.0) - 1)
    pos[nbg:] = cen[who] + dirs(ncl) * rad[:, None]
    del who, uu, rad, cen
    rad = pos.norm(dim=1)
    out = rad >= lc.rmax
    pos[out] *= (lc.rmax * 0.999 / rad[out])[:, None]  # pull the few escapees inside
    del rad, out
    perm = torch.randperm(n, generator=g, device=device)
    rec = torch.empty(n, 8, dtype=torch.float32, device=device)
    rec[:, 2:5] = pos[perm].to(torch.float32)
    del pos
    rec[:, 5:8] = 300.0 * torch.randn(n, 3, generator=g, device=device, dtype=torch.float32)
    rec.view(torch.int64).view(n, 4)[:, 0] = perm
    return rec.view(torch.uint8).view(-1)


def lightcone_shells_device(lc: Lightcone, recs):
    """lightcone_shells on the device: one sort by radius, then each shell
    (with its radial ghosts) is a contiguous slice, shuffled.  Returns
    (records, volumes, owned-radius bounds, table)."""
    import torch

    n = recs.numel() // 32
    dev = recs.device
    pos = recs.view(torch.float32).view(n, 8)[:, 2:5]
    r = pos.to(torch.float64).norm(dim=1)
    rs, order = torch.sort(r)
    del r
    edges = lc.shells()
    gh = lc.ghost
    r2max, tab = lc.varlink_table()
    vt = L.VarLinkTable([0.0, 0.0, 0.0], r2max, tab)
    gen = torch.Generator(device=dev)
    gen.manual_seed(lc.seed + 1)
    rows = recs.view(n, 32)
    parts, vols, bounds, off = [], [], [], 0
    for k in range(lc.nshell):
        lo, hi = float(edges[k]), float(edges[k + 1])
        q = torch.tensor([lo - gh, hi + gh], dtype=torch.float64, device=dev)
        i0, i1 = (int(x) for x in torch.searchsorted(rs, q).tolist())
        idx = order[i0:i1]
        idx = idx[torch.randperm(len(idx), generator=gen, device=dev)]
        parts.append(rows[idx])
        R = hi + gh
        sep = float(lc.density(hi) ** (-1.0 / 3.0))
        v = L.Volume()
        v.offset, v.n = off, len(idx)
        v.domain = L.domain([-R] * 3, [R] * 3, [-R] * 3, [R] * 3, [sep] * 3, float(lc.link(hi)),
                            lc.nphalomin, vt)
        off += len(idx)
        vols.append(v)
        bounds.append((lo, hi))
    return torch.cat(parts).view(-1), vols, bounds, vt


def lightcone_shells(lc: Lightcone, P: np.ndarray):
    """Radial shells with radial ghost layers as one batch.  Shell k owns
    r in [r_k, r_k+1); its domain is the cube around the origin that holds
    r_k+1 + ghost, with lcell = the mean separation at r_k+1 (so b <= 0.2
    inside the shell).  Returns (records, volumes, owned-radius bounds)."""
    edges = lc.shells()
    r = np.linalg.norm(P["Pos"].astype(np.float64), axis=1)
    g = lc.ghost
    r2max, tab = lc.varlink_table()
    vt = L.VarLinkTable([0.0, 0.0, 0.0], r2max, tab)
    parts, vols, bounds, off = [], [], [], 0
    for k in range(lc.nshell):
        lo, hi = edges[k], edges[k + 1]
        sel = (r >= lo - g) & (r < hi + g)
        Pk = P[sel]
        R = hi + g
        sep = float(lc.density(hi) ** (-1.0 / 3.0))
        v = L.Volume()
        v.offset, v.n = off, len(Pk)
        v.domain = L.domain([-R] * 3, [R] * 3, [-R] * 3, [R] * 3, [sep] * 3, float(lc.link(hi)),
                            lc.nphalomin, vt)
        off += len(Pk)
        parts.append(Pk)
        vols.append(v)
        bounds.append((lo, hi))
    return np.concatenate(parts), vols, bounds, vt
