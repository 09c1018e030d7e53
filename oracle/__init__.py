"""CPU oracle for the CM_Halofinder Friends-of-Friends path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg as the checker.  The product package
(``cm_halofinder_amd``) never imports it.

The C restatement lives in ``fof_oracle.c``; every function there cites the
reference lines it follows (src_v2/FOF.c, src_v2/read_data.c).  The
reference itself is not buildable in this image (its vars.h includes GSL
headers that are absent), so the oracle is pinned by brute force, exact
ties, closed forms and invariants in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libfof_oracle.so")


class Domain(C.Structure):
    _fields_ = [
        ("rmin", C.c_double * 3),
        ("rmax", C.c_double * 3),
        ("rmin_buff", C.c_double * 3),
        ("rmax_buff", C.c_double * 3),
        ("lcell", C.c_double * 3),
        ("linklength", C.c_double),
        ("nphalomin", C.c_double),
    ]


class VarLink(C.Structure):
    _fields_ = [
        ("origin", C.c_double * 3),
        ("r2max", C.c_double),
        ("ntab", C.c_uint32),
        ("l2tab", C.POINTER(C.c_double)),
    ]


class Halo(C.Structure):
    _fields_ = [
        ("head", C.c_uint32),
        ("n", C.c_uint32),
        ("owned", C.c_uint32),
        ("edge_warn", C.c_uint32),
        ("sum_pos", C.c_double * 3),
        ("sum_vel", C.c_double * 3),
        ("out", C.c_float * 6),
    ]


HALO_DTYPE = np.dtype(
    [
        ("head", "<u4"),
        ("n", "<u4"),
        ("owned", "<u4"),
        ("edge_warn", "<u4"),
        ("sum_pos", "<f8", 3),
        ("sum_vel", "<f8", 3),
        ("out", "<f4", 6),
    ]
)
assert HALO_DTYPE.itemsize == C.sizeof(Halo)

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        dp = C.POINTER(Domain)
        vp = C.c_void_p
        L.or_pair_linked.argtypes = [f32p, f32p, C.c_double]
        L.or_pair_linked.restype = C.c_int
        L.or_cell_linksq.argtypes = [dp, vp, C.c_int, C.c_int, C.c_int]
        L.or_cell_linksq.restype = C.c_double
        L.or_link_ring.argtypes = [C.c_uint32, f32p, dp, vp, u32p, C.POINTER(C.c_uint64)]
        L.or_link_uf.argtypes = [C.c_uint32, f32p, dp, vp, u32p, C.POINTER(C.c_uint64)]
        L.or_link_brute.argtypes = [C.c_uint32, f32p, dp, vp, u32p]
        for f in (L.or_link_ring, L.or_link_uf, L.or_link_brute):
            f.restype = C.c_int
        L.or_labels_from_ring.argtypes = [C.c_uint32, u32p, u32p]
        L.or_ring_from_labels.argtypes = [C.c_uint32, u32p, u32p]
        L.or_retrieve.argtypes = [C.c_uint32, u32p, f32p, f32p, dp, vp, C.c_uint64]
        L.or_retrieve.restype = C.c_int64
        L.or_props_comp.argtypes = [C.c_uint32, u32p, f32p, f32p, C.c_uint64, u32p, f64p, f64p]
        L.or_moments.argtypes = [C.c_uint32, u32p, f32p, f32p, C.c_uint64, u32p, f64p, f64p]
        L.or_moments_reference_out.argtypes = [C.c_uint32, f64p, f64p, f64p,
                                               np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"),
                                               np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")]
        L.or_subsample.argtypes = [C.c_uint32, u32p, f32p, C.POINTER(Subsamp), f64p, C.c_uint64,
                                   C.POINTER(C.c_uint64), u32p, C.POINTER(C.c_uint64), u32p,
                                   C.POINTER(C.c_uint64)]
        L.or_subsample.restype = C.c_int
        _lib = L
    return _lib


class Subsamp(C.Structure):
    _fields_ = [("samp_in", C.c_double), ("samp_out", C.c_double), ("rmin", C.c_double * 3),
                ("rmax", C.c_double * 3), ("nphalomin", C.c_double), ("closed_upper", C.c_uint32)]


def subsample(next0, pos, samp_in, samp_out, rmin, rmax, nphalomin, draws, closed_upper=0):
    """Subsampling (src_v2/FOF.c:172-284) of a 0-based ring with the draws
    given in order: (inhalo, outhalo, draws used), kept indices 1-based."""
    nxt = np.ascontiguousarray(next0, dtype=np.uint32)
    p = _pos(pos)
    u = np.ascontiguousarray(draws, dtype=np.float64)
    n = len(nxt)
    s = Subsamp(samp_in, samp_out, (C.c_double * 3)(*rmin), (C.c_double * 3)(*rmax), nphalomin,
                closed_upper)
    ih = np.zeros(max(n, 1), np.uint32)
    oh = np.zeros(max(n, 1), np.uint32)
    used, ni, no = C.c_uint64(), C.c_uint64(), C.c_uint64()
    rc = lib().or_subsample(n, nxt, p, C.byref(s), u, len(u), C.byref(used), ih, C.byref(ni), oh,
                            C.byref(no))
    if rc:
        raise ValueError("oracle subsample: the draws ran out")
    return ih[:ni.value].copy(), oh[:no.value].copy(), used.value


def make_domain(rmin, rmax, rmin_buff, rmax_buff, lcell, linklength, nphalomin) -> Domain:
    d = Domain()
    for k in range(3):
        d.rmin[k] = float(rmin[k])
        d.rmax[k] = float(rmax[k])
        d.rmin_buff[k] = float(rmin_buff[k])
        d.rmax_buff[k] = float(rmax_buff[k])
        d.lcell[k] = float(lcell[k])
    d.linklength = float(linklength)
    d.nphalomin = float(nphalomin)
    return d


class _VL:
    """Keeps the table alive next to the ctypes struct."""

    def __init__(self, origin, r2max, l2tab):
        self.tab = np.ascontiguousarray(l2tab, dtype=np.float64)
        self.s = VarLink()
        for k in range(3):
            self.s.origin[k] = float(origin[k])
        self.s.r2max = float(r2max)
        self.s.ntab = len(self.tab)
        self.s.l2tab = self.tab.ctypes.data_as(C.POINTER(C.c_double))

    @property
    def ptr(self):
        return C.cast(C.pointer(self.s), C.c_void_p)


def varlink(origin, r2max, l2tab) -> _VL:
    return _VL(origin, r2max, l2tab)


def _pos(pos):
    p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    assert p.size % 3 == 0
    return p


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed with code {rc}")


def pair_linked(pi, pj, linksq) -> bool:
    return bool(lib().or_pair_linked(_pos(pi), _pos(pj), float(linksq)))


def cell_linksq(d: Domain, vl, ix, iy, iz) -> float:
    return lib().or_cell_linksq(C.byref(d), vl.ptr if vl else None, ix, iy, iz)


def link_ring(pos, d: Domain, vl=None):
    """Faithful restatement of the reference's serial ring linker; returns
    (next0, pair_tests): next0[i] is the 0-based successor in the ring."""
    p = _pos(pos)
    n = p.size // 3
    nxt = np.empty(n, np.uint32)
    pairs = C.c_uint64(0)
    _check(lib().or_link_ring(n, p, C.byref(d), vl.ptr if vl else None, nxt, C.byref(pairs)), "ring")
    return nxt, int(pairs.value)


def link_uf(pos, d: Domain, vl=None):
    """Union-find over the reference candidate set; returns (label0, pair_tests)."""
    p = _pos(pos)
    n = p.size // 3
    lab = np.empty(n, np.uint32)
    pairs = C.c_uint64(0)
    _check(lib().or_link_uf(n, p, C.byref(d), vl.ptr if vl else None, lab, C.byref(pairs)), "uf")
    return lab, int(pairs.value)


def link_brute(pos, d: Domain, vl=None):
    p = _pos(pos)
    n = p.size // 3
    lab = np.empty(n, np.uint32)
    _check(lib().or_link_brute(n, p, C.byref(d), vl.ptr if vl else None, lab), "brute")
    return lab


def link_rc(fn, pos, d: Domain, vl=None) -> int:
    """Return code of a linker (for error-path tests)."""
    p = _pos(pos)
    n = p.size // 3
    out = np.empty(max(n, 1), np.uint32)
    L = lib()
    if fn == "brute":
        return L.or_link_brute(n, p, C.byref(d), vl.ptr if vl else None, out)
    f = L.or_link_ring if fn == "ring" else L.or_link_uf
    return f(n, p, C.byref(d), vl.ptr if vl else None, out, None)


def labels_from_ring(next0):
    nxt = np.ascontiguousarray(next0, dtype=np.uint32)
    lab = np.empty_like(nxt)
    lib().or_labels_from_ring(len(nxt), nxt, lab)
    return lab


def ring_from_labels(label0):
    lab = np.ascontiguousarray(label0, dtype=np.uint32)
    nxt = np.empty_like(lab)
    lib().or_ring_from_labels(len(lab), lab, nxt)
    return nxt


def retrieve(next0, pos, vel, d: Domain):
    """Retrieval + Checkhalo + Output_Halos sums over a ring: structured array
    of every group passing the size cut, in ascending head order."""
    nxt = np.ascontiguousarray(next0, dtype=np.uint32)
    p, v = _pos(pos), _pos(vel)
    n = len(nxt)
    L = lib()
    cnt = L.or_retrieve(n, nxt, p, v, C.byref(d), None, 0)
    if cnt < 0:
        raise MemoryError("oracle retrieve")
    out = np.zeros(cnt, HALO_DTYPE)
    L.or_retrieve(n, nxt, p, v, C.byref(d), out.ctypes.data_as(C.c_void_p), cnt)
    return out


def props_compensated(label0, pos, vel, heads):
    lab = np.ascontiguousarray(label0, dtype=np.uint32)
    h = np.ascontiguousarray(heads, dtype=np.uint32)
    sp = np.zeros((len(h), 3), np.float64)
    sv = np.zeros((len(h), 3), np.float64)
    lib().or_props_comp(len(lab), lab, _pos(pos), _pos(vel), len(h), h, sp.reshape(-1), sv.reshape(-1))
    return sp, sv


def moments(next0, pos, vel, heads):
    """Second moments of the halos headed by `heads` over the ring next0
    (src_v2/FOF.c:811-824): (raw, centred), each (nh, 12) -- Pos
    {xx,yy,zz,xy,xz,yz} then Vel; raw in ring order, centred compensated."""
    nxt = np.ascontiguousarray(next0, dtype=np.uint32)
    h = np.ascontiguousarray(heads, dtype=np.uint32)
    raw = np.zeros((len(h), 12), np.float64)
    cen = np.zeros((len(h), 12), np.float64)
    lib().or_moments(len(nxt), nxt, _pos(pos), _pos(vel), len(h), h, raw.reshape(-1), cen.reshape(-1))
    return raw, cen


def moments_reference_out(n, sum_pos, sum_vel, raw):
    """The reference's written INERTIA and DISPERSION floats for one halo
    (src_v2/FOF.c:848-860)."""
    ine = np.zeros(6, np.float32)
    dis = np.zeros(6, np.float32)
    lib().or_moments_reference_out(int(n), np.ascontiguousarray(sum_pos, np.float64),
                                   np.ascontiguousarray(sum_vel, np.float64),
                                   np.ascontiguousarray(raw, np.float64), ine, dis)
    return ine, dis


def minid_labels(label0, ids):
    """Canonical labels: every particle's group minimum particle ID (the
    relabelling under which membership is compared, independent of order)."""
    lab = np.asarray(label0, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.uint64)
    m = np.full(len(lab), np.iinfo(np.uint64).max, np.uint64)
    np.minimum.at(m, lab, ids)
    return m[lab]


def canonical(label0):
    """Labels are already canonical (group minimum index); kept for symmetry."""
    return np.asarray(label0, dtype=np.uint32)
