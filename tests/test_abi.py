"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, the ctypes mirror matches the header's struct layouts
(compiled here with gcc), and argument errors are reported before any device
work."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from cm_halofinder_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fof_gpu.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(fof_\w+)\s*\(", txt, flags=re.M)))


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert set(names) == set(L.ABI_SYMBOLS), names
    lib = L.lib()
    for nm in names:
        assert hasattr(lib, nm), nm
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    for nm in names:
        assert re.search(rf"\bT {nm}\b", out), nm


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "fof_gpu.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
int main(void) {
  S(fof_particles_v1) O(fof_particles_v1, base) O(fof_particles_v1, n) O(fof_particles_v1, off_vel)
  S(fof_varlink_v1) O(fof_varlink_v1, l2tab)
  S(fof_domain_v1) O(fof_domain_v1, linklength) O(fof_domain_v1, varlink)
  S(fof_opts_v1) O(fof_opts_v1, band) O(fof_opts_v1, dense_threshold)
  S(fof_halo_v1) O(fof_halo_v1, sum_pos) O(fof_halo_v1, sum_vel)
  S(fof_stats_v1) O(fof_stats_v1, phase_ms) O(fof_stats_v1, link_kernel_ms)
  S(fof_volume_v1) O(fof_volume_v1, domain)
  S(fof_gen_v1) O(fof_gen_v1, pair_l)
  S(fof_ghost_v1) O(fof_ghost_v1, nphalomin)
  return 0;
}
"""


def test_struct_layouts_match_header():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "p.c"), os.path.join(d, "p")
        open(src, "w").write(PROBE)
        subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    py = {
        "fof_particles_v1": L.Particles, "fof_varlink_v1": L.VarLink, "fof_domain_v1": L.Domain,
        "fof_opts_v1": L.Opts, "fof_halo_v1": L.Halo, "fof_stats_v1": L.Stats, "fof_volume_v1": L.Volume,
        "fof_gen_v1": L.Gen, "fof_ghost_v1": L.Ghost,
    }
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert getattr(py[t], f).offset == int(val), key
        else:
            assert C.sizeof(py[key]) == int(val), key
    assert L.HALO_DTYPE.itemsize == int(got["fof_halo_v1"])


def test_strerror_and_argument_errors():
    lib = L.lib()
    for code in (0, -1, -2, -3, -4, -5, -6):
        assert lib.fof_strerror(code)
    P = np.zeros(4, L.PART_DTYPE)
    d = L.domain([0] * 3, [1] * 3, [0] * 3, [1] * 3, [1] * 3, 0.1, 2)
    p = L.particles(P)
    p.version = 99
    assert lib.fof_link_gpu(None, C.byref(p), C.byref(d), None, None, None, None, 0, None, None) == L.E_ARG
    p = L.particles(P)
    p.stride = 30
    assert lib.fof_link_gpu(None, C.byref(p), C.byref(d), None, None, None, None, 0, None, None) == L.E_ARG
    p = L.particles(P)
    p.n = 1 << 31
    assert lib.fof_link_gpu(None, C.byref(p), C.byref(d), None, None, None, None, 0, None, None) == L.E_RANGE
    p = L.particles(P)
    o = L.opts(band=0.5)
    assert lib.fof_link_gpu(None, C.byref(p), C.byref(d), C.byref(o), None, None, None, 0, None, None) == L.E_ARG
    # batch volumes must tile the particle array
    v = (L.Volume * 2)()
    v[0].offset, v[0].n, v[1].offset, v[1].n = 0, 2, 3, 1
    assert lib.fof_link_gpu_batch(None, C.byref(p), v, 2, None, None, None, None, 0, None, None) == L.E_ARG


def test_no_device_reported_as_nodev():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert L.lib().fof_ctx_create(0, C.byref(h)) == L.E_NODEV
    P = np.zeros(4, L.PART_DTYPE)
    d = L.domain([0] * 3, [1] * 3, [0] * 3, [1] * 3, [1] * 3, 0.1, 2)
    assert L.lib().fof_link_gpu(None, C.byref(L.particles(P)), C.byref(d), None, None, None, None, 0,
                                None, None) == L.E_NODEV
