"""Host-side mirror of the reference's FOF() interface around the GPU linker.

The reference keeps all state in globals and runs, per MPI rank,
``Read_Parameterfile`` (src_v2/startup.c:109-314), ``Set_Params``
(src_v2/startup.c:3-107), ``Read_Data`` and ``FOF()`` (src_v2/FOF.c:23-339).
This module keeps those names and their argument meaning:

* :func:`Read_Parameterfile` parses ``run_parameters.dat`` with the same tag
  rules (every tag mandatory, unknown or repeated tags fatal, ``%`` comments).
* :func:`Set_Params` computes the rank's owned and buffered boxes.
* :func:`FOF` links one rank's particles with ``fof_link_gpu`` (replacing
  src_v2/FOF.c:44-161) and returns the successor ring ``next`` (1-based, as
  the reference's ``next[1..n]``), the heads ``ihalo`` / sizes ``nparthalo``
  of owned halos (Checkhalo, src_v2/FOF.c:378-430) and their Output_Halos
  values (src_v2/FOF.c:797-847).
* :func:`Output_Halos` writes the reference's catalogue layouts.

Errors that the reference turns into ``FatalError`` raise :class:`FatalError`.
"""
from __future__ import annotations

import dataclasses
import math
import struct

import numpy as np

from . import _lib as L


class FatalError(RuntimeError):
    """What src_v2/main.c:107-112 reports before MPI_Abort."""


# Tag table of src_v2/startup.c:121-256: (tag, kind).
_TAGS = [
    ("Nx", int), ("Ny", int), ("Nz", int),
    ("Px", int), ("Py", int), ("Pz", int),
    ("Lxmin", float), ("Lymin", float), ("Lzmin", float),
    ("Lxmax", float), ("Lymax", float), ("Lzmax", float),
    ("Seed", int), ("Buffer", float), ("Boundarysize", float), ("Linklength", float),
    ("Nphalomin", float), ("SampInHalos", float), ("SampOutHalos", float),
    ("InputStyle", int), ("Ninputfiles", int), ("Starting_file", int),
    ("Nread", int), ("Nwrite", int), ("InputFileBase", str), ("OutputFileBase", str),
]
_TAGS_LIGHTCONE_UNFORMATTED = [("InputInfoFile", str)]
# Origin_* are declared INT in the reference (src_v2/startup.c:245-255).
_TAGS_VARLINK = [("Omega", float), ("Origin_x", int), ("Origin_y", int), ("Origin_z", int)]


def _atoi(s: str) -> int:
    """C atoi: leading integer prefix, 0 if none."""
    s = s.strip()
    i, sign = 0, 1
    if i < len(s) and s[i] in "+-":
        sign = -1 if s[i] == "-" else 1
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return sign * int(s[i:j]) if j > i else 0


def _atof(s: str) -> float:
    """C atof: longest float prefix, 0.0 if none."""
    s = s.strip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return 0.0


def Read_Parameterfile(fname: str, varlink: bool = False, lightcone_unformatted: bool = False) -> dict:
    """Parse run_parameters.dat (src_v2/startup.c:109-314)."""
    tags = list(_TAGS)
    if lightcone_unformatted:
        tags += _TAGS_LIGHTCONE_UNFORMATTED
    if varlink:
        tags += _TAGS_VARLINK
    kinds = dict(tags)
    seen: dict = {}
    errors = []
    try:
        fh = open(fname, "r")
    except OSError:
        raise FatalError(f"ERROR: Parameter file {fname} not found.")
    with fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 2:
                continue
            tag, val = parts[0], parts[1]
            if tag.startswith("%"):
                continue
            if tag in kinds and tag not in seen:
                k = kinds[tag]
                seen[tag] = _atoi(val) if k is int else (_atof(val) if k is float else val)
            else:
                errors.append(f"Error: in run parameters file {fname}:  Tag '{tag}' not allowed or multiple defined.")
    for tag, _ in tags:
        if tag not in seen:
            errors.append(f"Error: I miss a value for tag '{tag}' in parameter file '{fname}'.")
    if errors:
        raise FatalError("\n".join(errors))
    return seen


@dataclasses.dataclass
class Geometry:
    local: tuple
    rmin: tuple
    rmax: tuple
    rmin_buff: tuple
    rmax_buff: tuple


def Set_Params(p: dict, ThisTask: int, NTask: int) -> Geometry:
    """Rank coordinates and boxes (src_v2/startup.c:13-47)."""
    if NTask != p["Nx"] * p["Ny"] * p["Nz"]:
        raise FatalError("ERROR: The number of cells does not match the number of processors... Whoops!!")
    nz = ThisTask // (p["Nx"] * p["Ny"])
    ny = (ThisTask - p["Nx"] * p["Ny"] * nz) // p["Nx"]
    nx = ThisTask - p["Nx"] * ny - p["Nx"] * p["Ny"] * nz
    rmin, rmax, rb0, rb1 = [], [], [], []
    for loc, N, lo, hi in ((nx, p["Nx"], p["Lxmin"], p["Lxmax"]), (ny, p["Ny"], p["Lymin"], p["Lymax"]),
                           (nz, p["Nz"], p["Lzmin"], p["Lzmax"])):
        a = loc * ((hi - lo) / float(N)) + lo
        b = a + ((hi - lo) / float(N))
        rmin.append(a)
        rmax.append(b)
        rb0.append(a - p["Boundarysize"])
        rb1.append(b + p["Boundarysize"])
    return Geometry((nx, ny, nz), tuple(rmin), tuple(rmax), tuple(rb0), tuple(rb1))


def lcell(p: dict, periodic: bool = True) -> tuple:
    """Reference cell size (src_v2/FOF.c:45-53)."""
    bs = p["Boundarysize"]
    ext = ((p["Lxmax"] - p["Lxmin"]), (p["Lymax"] - p["Lymin"]), (p["Lzmax"] - p["Lzmin"]))
    P = (p["Px"], p["Py"], p["Pz"])
    if periodic:
        return tuple(e / n for e, n in zip(ext, P))
    return tuple((e + 2.0 * bs) / n for e, n in zip(ext, P))


@dataclasses.dataclass
class FofResult:
    next: np.ndarray        # next[i-1] == reference next[i], 1-based successors
    labels: np.ndarray      # 0-based group minimum index
    halos: np.ndarray       # all groups passing the size cut (L.HALO_DTYPE)
    ihalo: np.ndarray       # 1-based heads of owned halos (src_v2/FOF.c:424)
    nparthalo: np.ndarray   # sizes of owned halos (src_v2/FOF.c:425)
    haloblock: np.ndarray   # float32 [nhalos, 6]: mean Pos, mean Vel (src_v2/FOF.c:842-847)
    stats: dict


def FOF(P: np.ndarray, p: dict, geom: Geometry, periodic: bool = True, ctx: L.Context | None = None,
        varlink: L.VarLinkTable | None = None, timing: bool = False) -> FofResult:
    """Link one rank's particles on the GPU and derive the owned-halo table.

    P is the rank's particle array (owned + ghost shell) with the reference's
    32-byte record (``L.PART_DTYPE``)."""
    P = np.ascontiguousarray(P)
    if P.dtype != L.PART_DTYPE:
        raise TypeError("P must use the reference record layout (PART_DTYPE)")
    n = len(P)
    ctx = ctx or _default_ctx()
    dom = L.domain(geom.rmin, geom.rmax, geom.rmin_buff, geom.rmax_buff, lcell(p, periodic),
                   p["Linklength"], p["Nphalomin"], varlink)
    nxt = np.empty(n, np.uint32)
    lab = np.empty(n, np.uint32)
    cap = n if p["Nphalomin"] <= 1 else n // max(1, int(math.ceil(p["Nphalomin"]))) + 1
    halos = np.empty(max(cap, 1), L.HALO_DTYPE)
    st = L.Stats()
    rc, nh = ctx.link(L.particles(P), dom, L.opts(L.OPT_TIMING if timing else 0), nxt, lab, halos,
                      len(halos), st)
    if rc == L.E_DOMAIN:
        raise FatalError("ERROR: particle outside the buffered region (src_v2/read_data.c:810-812)")
    L.check(rc, "fof_link_gpu")
    halos = halos[:nh]
    own = halos[halos["owned"] == 1]
    inv = 1.0 / own["n"].astype(np.float64)
    hb = np.empty((len(own), 6), np.float32)
    hb[:, :3] = (own["sum_pos"] * inv[:, None]).astype(np.float32)
    hb[:, 3:] = (own["sum_vel"] * inv[:, None]).astype(np.float32)
    return FofResult(nxt, lab, halos, own["head"].astype(np.int64) + 1, own["n"].astype(np.int64), hb,
                     st.as_dict())


_CTX = None


def _default_ctx():
    global _CTX
    if _CTX is None:
        _CTX = L.Context(0)
    return _CTX


def Output_Halos(path: str, res: FofResult, P: np.ndarray | None = None, unformatted: bool = False,
                 output_particles: bool = False, particle_id: bool = True) -> None:
    """Write the rank's halo file in the reference's layouts
    (src_v2/FOF.c:546-913): ASCII ``%15d %15.6f x6`` lines, Fortran-record
    binary, or OUTPUT_PARTICLES with the members in ring order."""
    nh = len(res.ihalo)
    if nh == 0:
        return  # the reference writes no file for a rank without halos (:588)
    if not output_particles:
        if unformatted:
            with open(path, "wb") as f:
                f.write(struct.pack("<iii", 4, nh, 4))
                blk = np.empty((nh, 7), np.float32)
                blk[:, 0] = res.nparthalo.astype(np.float32)
                blk[:, 1:] = res.haloblock
                f.write(struct.pack("<i", nh * 7 * 4))
                f.write(blk.tobytes())
                f.write(struct.pack("<i", nh * 7 * 4))
        else:
            with open(path, "w") as f:
                for k in range(nh):
                    f.write("%15d %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n"
                            % (res.nparthalo[k], *[float(v) for v in res.haloblock[k]]))
        return
    if P is None:
        raise ValueError("OUTPUT_PARTICLES needs the particle array")
    with open(path, "wb" if unformatted else "w") as f:
        if unformatted:
            f.write(struct.pack("<iii", 4, nh, 4))
        else:
            f.write("%15d\n" % nh)
        for k in range(nh):
            if unformatted:
                blk = np.empty(7, np.float32)
                blk[0] = np.float32(res.nparthalo[k])
                blk[1:] = res.haloblock[k]
                f.write(struct.pack("<i", 28) + blk.tobytes() + struct.pack("<i", 28))
            else:
                f.write("%15d %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n"
                        % (res.nparthalo[k], *[float(v) for v in res.haloblock[k]]))
            j = int(res.ihalo[k])
            while True:
                r = P[j - 1]
                vals = [float(v) for v in r["Pos"]] + [float(v) for v in r["Vel"]]
                if unformatted:
                    body = (struct.pack("<Q", int(r["ID"])) if particle_id else b"") + np.asarray(vals, np.float32).tobytes()
                    f.write(struct.pack("<i", len(body)) + body + struct.pack("<i", len(body)))
                elif particle_id:
                    f.write("%15d %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n" % (int(r["ID"]), *vals))
                else:
                    f.write("%15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n" % tuple(vals))
                j = int(res.next[j - 1])
                if j == int(res.ihalo[k]):
                    break
