"""ctypes binding of include/fof_io.h: the reference's particle file formats
(ASCII, Gadget-1, PICOLA lightcone) and halo catalogue layouts, implemented
in the library's host code."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

# Every function fof_io.h declares.
IO_SYMBOLS = ("fof_read_ascii", "fof_read_gadget1", "fof_read_picola", "fof_write_halos",
              "fof_write_info", "fof_subsample", "fof_write_subsample")

OUT_UNFORMATTED, OUT_PARTICLES, OUT_INERTIA, OUT_DISPERSION, OUT_PARTICLE_ID = 0x1, 0x2, 0x4, 0x8, 0x10


class GadgetHeader(C.Structure):
    _fields_ = [
        ("npart", C.c_uint32 * 6),
        ("mass", C.c_double * 6),
        ("time", C.c_double),
        ("redshift", C.c_double),
        ("flag_sfr", C.c_int32),
        ("flag_feedback", C.c_int32),
        ("npart_total", C.c_uint32 * 6),
        ("flag_cooling", C.c_int32),
        ("num_files", C.c_int32),
        ("box_size", C.c_double),
        ("omega0", C.c_double),
        ("omega_lambda", C.c_double),
        ("hubble_param", C.c_double),
        ("flag_stellarage", C.c_int32),
        ("flag_metals", C.c_int32),
        ("hashtabsize", C.c_int32),
        ("fill", C.c_char * 84),
    ]


assert C.sizeof(GadgetHeader) == 256


class Subsample(C.Structure):
    """fof_subsample_v1"""
    _fields_ = [
        ("samp_in", C.c_double),
        ("samp_out", C.c_double),
        ("rmin", C.c_double * 3),
        ("rmax", C.c_double * 3),
        ("nphalomin", C.c_double),
        ("closed_upper", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


UNIFORM_FN = C.CFUNCTYPE(C.c_double, C.c_void_p)

_bound = False


def _lib():
    global _bound
    lib = L.lib()
    if not _bound:
        vp, u64 = C.c_void_p, C.c_uint64
        for f in (lib.fof_read_ascii, lib.fof_read_picola):
            f.argtypes = [C.c_char_p, C.c_int, vp, u64, C.POINTER(C.c_uint64)]
            f.restype = C.c_int
        lib.fof_read_gadget1.argtypes = [C.c_char_p, C.c_int, vp, u64, C.POINTER(C.c_uint64),
                                         C.POINTER(GadgetHeader)]
        lib.fof_read_gadget1.restype = C.c_int
        lib.fof_write_halos.argtypes = [C.c_char_p, C.c_uint32, vp, vp, u64, C.POINTER(L.Particles), vp]
        lib.fof_write_halos.restype = C.c_int
        lib.fof_write_info.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                       C.c_int, C.c_int, vp, vp, vp]
        lib.fof_write_info.restype = C.c_int
        lib.fof_subsample.argtypes = [C.POINTER(L.Particles), vp, C.POINTER(Subsample), UNIFORM_FN, vp,
                                      vp, u64, C.POINTER(C.c_uint64), vp, u64, C.POINTER(C.c_uint64)]
        lib.fof_subsample.restype = C.c_int
        lib.fof_write_subsample.argtypes = [C.c_char_p, C.c_uint32, C.POINTER(L.Particles), vp, u64]
        lib.fof_write_subsample.restype = C.c_int
        _bound = True
    return lib


def _read(fn, path, with_id, *extra):
    n = C.c_uint64(0)
    L.check(fn(str(path).encode(), int(with_id), None, 0, C.byref(n), *extra), "read (count)")
    out = np.empty(n.value, L.PART_DTYPE)
    L.check(fn(str(path).encode(), int(with_id), out.ctypes.data, len(out), C.byref(n), *extra), "read")
    return out


def read_ascii(path, with_id=True) -> np.ndarray:
    """ASCII particles (src_v2/read_data.c:286-312) as reference records."""
    return _read(_lib().fof_read_ascii, path, with_id)


def read_gadget1(path, with_id=True):
    """Gadget-1 snapshot (src_v2/read_data.c:321-365): (records, header)."""
    h = GadgetHeader()
    recs = _read(_lib().fof_read_gadget1, path, with_id, C.byref(h))
    return recs, h


def read_picola(path, with_id=True) -> np.ndarray:
    """PICOLA unformatted lightcone (src_v2/read_data.c:207-277)."""
    return _read(_lib().fof_read_picola, path, with_id)


def read_rc(kind, path, with_id=True, cap=None):
    """Return code of a reader (for error tests); cap None: a counting call."""
    lib = _lib()
    n = C.c_uint64(0)
    out = None if cap is None else np.empty(max(cap, 1), L.PART_DTYPE)
    ptr = None if out is None else out.ctypes.data
    c = 0 if cap is None else cap
    if kind == "gadget1":
        return lib.fof_read_gadget1(str(path).encode(), int(with_id), ptr, c, C.byref(n), None)
    fn = lib.fof_read_ascii if kind == "ascii" else lib.fof_read_picola
    return fn(str(path).encode(), int(with_id), ptr, c, C.byref(n))


def write_halos(path, flags, halos, moments=None, P=None, next_out=None) -> int:
    """One rank's halo file in the reference's layouts (src_v2/FOF.c:546-913)."""
    halos = np.ascontiguousarray(halos, dtype=L.HALO_DTYPE)
    mom = None if moments is None else np.ascontiguousarray(moments, dtype=L.MOMENTS_DTYPE)
    parts = None if P is None else L.particles(P)
    nxt = None if next_out is None else np.ascontiguousarray(next_out, dtype=np.uint32)
    return _lib().fof_write_halos(str(path).encode(), flags, halos.ctypes.data,
                                  None if mom is None else mom.ctypes.data, len(halos),
                                  C.byref(parts) if parts is not None else None,
                                  None if nxt is None else nxt.ctypes.data)


def write_info(path, lmin, lmax, nx, ny, nz, nhalos, ninhalo=None, nouthalo=None) -> int:
    """The .info file (src_v2/FOF.c:917-957)."""
    lo = (C.c_double * 3)(*[float(x) for x in lmin])
    hi = (C.c_double * 3)(*[float(x) for x in lmax])
    nh = np.ascontiguousarray(nhalos, dtype=np.int32)
    ni = None if ninhalo is None else np.ascontiguousarray(ninhalo, dtype=np.uint32)
    no = None if nouthalo is None else np.ascontiguousarray(nouthalo, dtype=np.uint32)
    return _lib().fof_write_info(str(path).encode(), lo, hi, nx, ny, nz, nh.ctypes.data,
                                 None if ni is None else ni.ctypes.data,
                                 None if no is None else no.ctypes.data)


def subsample(P, next_out, samp_in, samp_out, rmin, rmax, nphalomin, uniform, closed_upper=0,
              in_cap=None, out_cap=None):
    """Subsampling (src_v2/FOF.c:172-284) of the host records P over the
    1-based ring next_out; uniform() is called once per draw.  Returns
    (rc, inhalo, outhalo) with 1-based indices in visit order."""
    parts = L.particles(P)
    nxt = np.ascontiguousarray(next_out, dtype=np.uint32)
    n = len(nxt)
    s = Subsample(samp_in, samp_out, (C.c_double * 3)(*rmin), (C.c_double * 3)(*rmax), nphalomin,
                  closed_upper, 0)
    ic = n if in_cap is None else in_cap
    oc = n if out_cap is None else out_cap
    ih = np.zeros(max(ic, 1), np.uint32)
    oh = np.zeros(max(oc, 1), np.uint32)
    ni, no = C.c_uint64(), C.c_uint64()
    fn = UNIFORM_FN(lambda _state: float(uniform()))
    rc = _lib().fof_subsample(C.byref(parts), nxt.ctypes.data, C.byref(s), fn, None, ih.ctypes.data,
                              ic, C.byref(ni), oh.ctypes.data, oc, C.byref(no))
    return rc, ih[:ni.value].copy(), oh[:no.value].copy()


def write_subsample(path, flags, P, idx) -> int:
    """A subsample file (src_v2/FOF.c:432-544) of the 1-based indices idx."""
    parts = L.particles(P)
    ix = np.ascontiguousarray(idx, dtype=np.uint32)
    return _lib().fof_write_subsample(str(path).encode(), flags, C.byref(parts),
                                      ix.ctypes.data if len(ix) else None, len(ix))
