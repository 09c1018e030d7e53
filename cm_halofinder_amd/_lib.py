"""ctypes binding of the C ABI declared in include/fof_gpu.h.

Buffers may be numpy arrays (host memory) or torch tensors on the GPU
(device memory; torch is only used as an allocator here).  The compute path
is the HIP library; there is no CPU fallback: if ``libcmfof.so`` is missing
the import of this module's ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcmfof.so")

ABI_VERSION = 1
MEM_HOST, MEM_DEVICE = 0, 1
OPT_TIMING, OPT_OUT_DEVICE, OPT_NO_RING, OPT_NO_HALOS = 0x1, 0x2, 0x4, 0x8
OPT_GRAPH, OPT_SPARSE, OPT_AUTOTUNE = 0x10, 0x20, 0x40
OPT_WRITE_PAIRS, OPT_WRITE_DIRECT = 0x80, 0x100
NOALLOC_BYTES = 256 << 20  # FOF_NOALLOC_BYTES

E_ARG, E_DOMAIN, E_RANGE, E_NOMEM, E_HIP, E_NODEV = -1, -2, -3, -4, -5, -6

# Every extern "C" function the header declares.
ABI_SYMBOLS = (
    "fof_ctx_create",
    "fof_ctx_destroy",
    "fof_link_gpu",
    "fof_link_gpu_batch",
    "fof_link_gpu_ex",
    "fof_link_gpu_noalloc",
    "fof_plan",
    "fof_strerror",
    "fof_device_info",
    "fof_gen_host",
    "fof_gen_device",
    "fof_ghosts_host",
    "fof_ghosts_device",
)


class Particles(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("location", C.c_uint32),
        ("base", C.c_void_p),
        ("n", C.c_uint64),
        ("stride", C.c_uint32),
        ("off_id", C.c_int32),
        ("off_pos", C.c_uint32),
        ("off_vel", C.c_uint32),
    ]


class VarLink(C.Structure):
    _fields_ = [
        ("origin", C.c_double * 3),
        ("r2max", C.c_double),
        ("ntab", C.c_uint32),
        ("location", C.c_uint32),
        ("l2tab", C.c_void_p),
    ]


class Domain(C.Structure):
    _fields_ = [
        ("rmin", C.c_double * 3),
        ("rmax", C.c_double * 3),
        ("rmin_buff", C.c_double * 3),
        ("rmax_buff", C.c_double * 3),
        ("lcell", C.c_double * 3),
        ("linklength", C.c_double),
        ("nphalomin", C.c_double),
        ("varlink", C.POINTER(VarLink)),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("flags", C.c_uint32),
        ("band", C.c_double),
        ("dense_threshold", C.c_uint32),
        ("pipelines", C.c_uint32),
        ("big_grid", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Halo(C.Structure):
    _fields_ = [
        ("head", C.c_uint32),
        ("n", C.c_uint32),
        ("owned", C.c_uint32),
        ("edge_warn", C.c_uint32),
        ("sum_pos", C.c_double * 3),
        ("sum_vel", C.c_double * 3),
    ]


HALO_DTYPE = np.dtype(
    [
        ("head", "<u4"),
        ("n", "<u4"),
        ("owned", "<u4"),
        ("edge_warn", "<u4"),
        ("sum_pos", "<f8", 3),
        ("sum_vel", "<f8", 3),
    ]
)
assert HALO_DTYPE.itemsize == C.sizeof(Halo) == 64

NPHASE = 17
PHASES = ("keys", "scan", "scatter", "uf_init", "subsort", "link_sparse", "link_dense",
          "flatten", "label", "groups", "group_sort", "ring", "halo_list", "halo_sums", "output",
          "link_small", "total")


class Stats(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("nvol", C.c_uint32),
        ("n", C.c_uint64),
        ("ncells", C.c_uint64),
        ("ndense", C.c_uint64),
        ("pair_tests", C.c_uint64),
        ("rechecks", C.c_uint64),
        ("ngroups", C.c_uint64),
        ("nmembers", C.c_uint64),
        ("nhalos", C.c_uint64),
        ("nhalos_owned", C.c_uint64),
        ("sparse", C.c_uint32),
        ("dense_threshold", C.c_uint32),
        ("phase_ms", C.c_double * NPHASE),
        ("link_kernel_ms", C.c_double),
        ("pair_tests_kernel", C.c_uint64 * 3),
        ("pairs_joined", C.c_uint64),
        ("pipelines", C.c_uint32),
        ("graph", C.c_uint32),
        ("big_grid", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_
             if k not in ("phase_ms", "pair_tests_kernel")}
        d["phase_ms"] = dict(zip(PHASES, self.phase_ms))
        d["pair_tests_kernel"] = dict(zip(LINK_KERNELS, self.pair_tests_kernel))
        return d


# pair_tests_kernel order
LINK_KERNELS = ("k_link_sparse", "k_link_dense", "k_link_small")


class Moments(C.Structure):
    _fields_ = [
        ("pos2", C.c_double * 6),
        ("vel2", C.c_double * 6),
        ("cpos2", C.c_double * 6),
        ("cvel2", C.c_double * 6),
    ]


MOMENTS_DTYPE = np.dtype([("pos2", "<f8", 6), ("vel2", "<f8", 6), ("cpos2", "<f8", 6),
                          ("cvel2", "<f8", 6)])
assert MOMENTS_DTYPE.itemsize == C.sizeof(Moments) == 192


class Outputs(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("location", C.c_uint32),
        ("next", C.c_void_p),
        ("label", C.c_void_p),
        ("minid", C.c_void_p),
        ("halos", C.c_void_p),
        ("moments", C.c_void_p),
        ("halo_cap", C.c_uint64),
        ("halo_offsets", C.c_void_p),
        ("nhalos", C.c_uint64),
        ("stats", C.c_void_p),
    ]


class Plan(C.Structure):
    _fields_ = [
        ("version", C.c_uint32),
        ("sparse", C.c_uint32),
        ("n", C.c_uint64),
        ("cells", C.c_uint64),
        ("grid_cells", C.c_double),
        ("workspace_bytes", C.c_uint64),
        ("dense_threshold", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Volume(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("n", C.c_uint64), ("domain", Domain)]


class Gen(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("n", C.c_uint64),
        ("box", C.c_double),
        ("nclumps", C.c_uint64),
        ("clump_frac", C.c_double),
        ("mmin", C.c_double),
        ("mmax", C.c_double),
        ("alpha", C.c_double),
        ("a0", C.c_double),
        ("sigma_v", C.c_double),
        ("umax", C.c_double),
        ("pair_frac", C.c_double),
        ("pair_rel", C.c_double),
        ("pair_l", C.c_double),
    ]


class Ghost(C.Structure):
    _fields_ = [
        ("box", C.c_double),
        ("shell", C.c_double),
        ("gx", C.c_uint32),
        ("gy", C.c_uint32),
        ("gz", C.c_uint32),
        ("pside", C.c_uint32),
        ("linklength", C.c_double),
        ("nphalomin", C.c_double),
    ]


# Particle record of the reference with PARTICLE_ID + MEMORY_MODE
# (src_v2/vars.h:36-47): u64 ID, float Pos[3], float Vel[3].
PART_DTYPE = np.dtype([("ID", "<u8"), ("Pos", "<f4", 3), ("Vel", "<f4", 3)])
assert PART_DTYPE.itemsize == 32

_lib = None


class FofError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = lib().fof_strerror(code).decode()
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


def lib():
    """Load the HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cm_halofinder_amd.build` "
                "(there is no CPU fallback)"
            )
        # One HIP runtime per process: torch's libtorch_hip NEEDs the
        # unversioned "libamdhip64.so" and ships its own copy, while this
        # library NEEDs "libamdhip64.so.7".  Loading torch first makes this
        # library bind to torch's already-loaded runtime (same soname);
        # loading it first would make torch load a second runtime, which then
        # sees no devices.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
        L.fof_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.fof_ctx_create.restype = C.c_int
        L.fof_ctx_destroy.argtypes = [vp]
        L.fof_ctx_destroy.restype = None
        L.fof_link_gpu.argtypes = [
            vp, C.POINTER(Particles), C.POINTER(Domain), C.POINTER(Opts),
            vp, vp, vp, u64, C.POINTER(C.c_uint64), C.POINTER(Stats),
        ]
        L.fof_link_gpu.restype = C.c_int
        L.fof_link_gpu_batch.argtypes = [
            vp, C.POINTER(Particles), C.POINTER(Volume), u32, C.POINTER(Opts),
            vp, vp, vp, u64, vp, C.POINTER(Stats),
        ]
        L.fof_link_gpu_batch.restype = C.c_int
        L.fof_link_gpu_ex.argtypes = [
            vp, C.POINTER(Particles), C.POINTER(Volume), u32, C.POINTER(Opts), C.POINTER(Outputs),
        ]
        L.fof_link_gpu_ex.restype = C.c_int
        L.fof_link_gpu_noalloc.argtypes = [vp, C.POINTER(Particles), C.POINTER(Domain),
                                           C.POINTER(Opts), C.POINTER(Outputs)]
        L.fof_link_gpu_noalloc.restype = C.c_int
        L.fof_plan.argtypes = [C.POINTER(Particles), C.POINTER(Volume), u32, C.POINTER(Opts),
                               C.POINTER(Plan)]
        L.fof_plan.restype = C.c_int
        L.fof_strerror.argtypes = [C.c_int]
        L.fof_strerror.restype = C.c_char_p
        L.fof_device_info.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
        L.fof_device_info.restype = C.c_int
        L.fof_gen_host.argtypes = [C.POINTER(Gen), vp]
        L.fof_gen_host.restype = C.c_int
        L.fof_gen_device.argtypes = [C.POINTER(Gen), vp]
        L.fof_gen_device.restype = C.c_int
        for f in (L.fof_ghosts_host, L.fof_ghosts_device):
            f.argtypes = [C.POINTER(Ghost), vp, u64, vp, u64, vp, C.POINTER(Volume)]
            f.restype = C.c_int
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        raise FofError(rc, what)


# ---------------------------------------------------------------------------
# buffers
# ---------------------------------------------------------------------------
def _is_torch(x):
    return type(x).__module__.startswith("torch")


def ptr(x):
    """(address, location) of a numpy array or torch tensor (None -> NULL)."""
    if x is None:
        return None, MEM_HOST
    if _is_torch(x):
        return x.data_ptr(), (MEM_DEVICE if x.is_cuda else MEM_HOST)
    if isinstance(x, np.ndarray):
        assert x.flags["C_CONTIGUOUS"]
        return x.ctypes.data, MEM_HOST
    raise TypeError(type(x))


def particles(buf, n=None, stride=32, off_id=0, off_pos=8, off_vel=20) -> Particles:
    """Describe an AoS particle array (default: the 32-byte reference record)."""
    p = Particles()
    p.version = ABI_VERSION
    addr, loc = ptr(buf)
    p.base = addr
    p.location = loc
    if n is None:
        nbytes = buf.nbytes if isinstance(buf, np.ndarray) else buf.numel() * buf.element_size()
        n = nbytes // stride
    p.n = n
    p.stride, p.off_id, p.off_pos, p.off_vel = stride, off_id, off_pos, off_vel
    return p


def domain(rmin, rmax, rmin_buff, rmax_buff, lcell, linklength, nphalomin, varlink=None) -> Domain:
    d = Domain()
    for k in range(3):
        d.rmin[k], d.rmax[k] = float(rmin[k]), float(rmax[k])
        d.rmin_buff[k], d.rmax_buff[k] = float(rmin_buff[k]), float(rmax_buff[k])
        d.lcell[k] = float(lcell[k])
    d.linklength = float(linklength)
    d.nphalomin = float(nphalomin)
    if varlink is not None:
        d.varlink = C.pointer(varlink.s)
        d._vl_keep = varlink  # keep the table alive with the struct
    return d


class VarLinkTable:
    """Variable linking length table (uniform samples of l^2 on r^2)."""

    def __init__(self, origin, r2max, l2tab):
        self.tab = np.ascontiguousarray(l2tab, dtype=np.float64)
        self.s = VarLink()
        for k in range(3):
            self.s.origin[k] = float(origin[k])
        self.s.r2max = float(r2max)
        self.s.ntab = len(self.tab)
        self.s.location = MEM_HOST
        self.s.l2tab = self.tab.ctypes.data


def opts(flags=0, band=0.0, dense_threshold=0, pipelines=0, big_grid=0) -> Opts:
    o = Opts()
    o.version = ABI_VERSION
    o.flags = flags
    o.band = band
    o.dense_threshold = dense_threshold
    o.pipelines = pipelines
    o.big_grid = big_grid
    return o


def plan(parts: Particles, vols, o: Opts | None = None) -> Plan:
    """fof_plan: the addressing, grid sizes and workspace of a call (host only)."""
    arr = (Volume * len(vols))(*vols)
    p = Plan()
    check(lib().fof_plan(C.byref(parts), arr, len(vols), C.byref(o) if o else None, C.byref(p)),
          "fof_plan")
    return p


def device_info(dev=0) -> str:
    buf = C.create_string_buffer(512)
    lib().fof_device_info(dev, buf, 512)
    return buf.value.decode()


class Context:
    """Owns a fof_ctx (device workspace reused across calls)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().fof_ctx_create(device, C.byref(h)), "fof_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().fof_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def link(self, parts: Particles, dom: Domain, o: Opts | None = None, next_out=None,
             label_out=None, halos=None, halo_cap=0, stats: Stats | None = None):
        nh = C.c_uint64(0)
        rc = lib().fof_link_gpu(
            self.h, C.byref(parts), C.byref(dom), C.byref(o) if o else None,
            ptr(next_out)[0], ptr(label_out)[0], ptr(halos)[0], halo_cap,
            C.byref(nh), C.byref(stats) if stats else None,
        )
        return rc, nh.value

    def link_ex(self, parts: Particles, vols, o: Opts | None = None, next_out=None,
                label_out=None, minid_out=None, halos=None, moments=None, halo_cap=0,
                halo_offsets=None, stats: Stats | None = None, location=None):
        """fof_link_gpu_ex: returns (rc, nhalos).  `location` defaults to where
        the first given output array lives."""
        arr = (Volume * len(vols))(*vols)
        out = Outputs()
        out.version = ABI_VERSION
        if location is None:
            location = MEM_HOST
            for x in (next_out, label_out, minid_out, halos, moments):
                if x is not None:
                    location = ptr(x)[1]
                    break
        out.location = location
        out.next, out.label, out.minid = ptr(next_out)[0], ptr(label_out)[0], ptr(minid_out)[0]
        out.halos, out.moments = ptr(halos)[0], ptr(moments)[0]
        out.halo_cap = halo_cap
        out.halo_offsets = ptr(halo_offsets)[0]
        out.stats = C.cast(C.pointer(stats), C.c_void_p) if stats is not None else None
        rc = lib().fof_link_gpu_ex(self.h, C.byref(parts), arr, len(vols),
                                   C.byref(o) if o else None, C.byref(out))
        return rc, out.nhalos

    def link_noalloc(self, parts: Particles, dom: Domain, o: Opts | None = None, next_out=None,
                     label_out=None, minid_out=None, halos=None, moments=None, halo_cap=0):
        """fof_link_gpu_noalloc (device-resident arrays only): returns (rc, nhalos)."""
        out = Outputs()
        out.version = ABI_VERSION
        out.location = MEM_DEVICE
        out.next, out.label, out.minid = ptr(next_out)[0], ptr(label_out)[0], ptr(minid_out)[0]
        out.halos, out.moments = ptr(halos)[0], ptr(moments)[0]
        out.halo_cap = halo_cap
        rc = lib().fof_link_gpu_noalloc(self.h, C.byref(parts), C.byref(dom),
                                        C.byref(o) if o else None, C.byref(out))
        return rc, out.nhalos

    def link_batch(self, parts: Particles, vols, o: Opts | None = None, next_out=None,
                   label_out=None, halos=None, halo_cap=0, halo_offsets=None,
                   stats: Stats | None = None):
        arr = (Volume * len(vols))(*vols)
        rc = lib().fof_link_gpu_batch(
            self.h, C.byref(parts), arr, len(vols), C.byref(o) if o else None,
            ptr(next_out)[0], ptr(label_out)[0], ptr(halos)[0], halo_cap,
            ptr(halo_offsets)[0], C.byref(stats) if stats else None,
        )
        return rc


# ---------------------------------------------------------------------------
# workload tools
# ---------------------------------------------------------------------------
def gen_params(seed, n, box, nclumps, clump_frac=0.5, mmin=20.0, mmax=5000.0, alpha=1.9,
               a0=0.05, sigma_v=300.0, umax=0.99, pair_frac=0.0, pair_rel=1e-6, pair_l=0.0) -> Gen:
    g = Gen()
    g.seed, g.n, g.box, g.nclumps, g.clump_frac = seed, n, box, nclumps, clump_frac
    g.mmin, g.mmax, g.alpha, g.a0, g.sigma_v, g.umax = mmin, mmax, alpha, a0, sigma_v, umax
    g.pair_frac, g.pair_rel, g.pair_l = pair_frac, pair_rel, pair_l
    return g


def gen_host(g: Gen) -> np.ndarray:
    out = np.empty(g.n, PART_DTYPE)
    check(lib().fof_gen_host(C.byref(g), out.ctypes.data), "fof_gen_host")
    return out


def gen_device(g: Gen, out_tensor):
    check(lib().fof_gen_device(C.byref(g), out_tensor.data_ptr()), "fof_gen_device")
    return out_tensor


def ghost_params(box, shell, gx, gy, gz, pside, linklength, nphalomin) -> Ghost:
    g = Ghost()
    g.box, g.shell, g.gx, g.gy, g.gz = box, shell, gx, gy, gz
    g.pside, g.linklength, g.nphalomin = pside, linklength, nphalomin
    return g


def ghosts_host(g: Ghost, recs: np.ndarray):
    nvol = g.gx * g.gy * g.gz
    counts = np.zeros(nvol, np.uint64)
    vols = (Volume * nvol)()
    check(lib().fof_ghosts_host(C.byref(g), recs.ctypes.data, len(recs), None, 0,
                                counts.ctypes.data, vols), "fof_ghosts_host")
    out = np.empty(int(counts.sum()), PART_DTYPE)
    check(lib().fof_ghosts_host(C.byref(g), recs.ctypes.data, len(recs), out.ctypes.data,
                                len(out), counts.ctypes.data, vols), "fof_ghosts_host")
    return out, counts, list(vols)


def ghosts_device(g: Ghost, recs_dev, n, out_dev=None, out_cap=0):
    """Device ghost builder; with out_dev None only counts/volumes are made."""
    nvol = g.gx * g.gy * g.gz
    counts = np.zeros(nvol, np.uint64)
    vols = (Volume * nvol)()
    check(lib().fof_ghosts_device(C.byref(g), recs_dev.data_ptr(), n,
                                  out_dev.data_ptr() if out_dev is not None else None,
                                  out_cap, counts.ctypes.data, vols), "fof_ghosts_device")
    return counts, list(vols)
