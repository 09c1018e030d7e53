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


def volumes_host(nvol: int, median: float, lo: float, hi: float, seed: int = 5):
    """Independent periodic volumes built on the host (scaled C5)."""
    n, boxes, b = volumes_spec(nvol, median, lo=lo, hi=hi, seed=seed)
    parts, vols, off = [], [], 0
    for k in range(nvol):
        bx = volume_box(k, int(n[k]), float(boxes[k]), float(b[k]), seed)
        g = L.gen_params(bx.seed, int(n[k]), bx.box, max(1, int(round(n[k] * CLUMPS_PER_PARTICLE))),
                         pair_l=bx.linklength)
        recs = L.gen_host(g)
        gh = L.ghost_params(bx.box, bx.shell, 1, 1, 1, 1, bx.linklength, bx.nphalomin)
        out, counts, vv = L.ghosts_host(gh, recs)
        v = vv[0]
        # the reference cell for a volume of n particles: box / n^(1/3)
        for d in range(3):
            v.domain.lcell[d] = bx.box / np.cbrt(float(n[k]))
        v.offset = off
        off += len(out)
        parts.append(out)
        vols.append(v)
    return np.concatenate(parts), vols
