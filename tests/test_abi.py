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
  O(fof_stats_v1, pair_tests_kernel) O(fof_stats_v1, pairs_joined) O(fof_stats_v1, graph)
  S(fof_moments_v1) O(fof_moments_v1, vel2) O(fof_moments_v1, cvel2)
  S(fof_outputs_v1) O(fof_outputs_v1, minid) O(fof_outputs_v1, halo_cap) O(fof_outputs_v1, nhalos)
  O(fof_outputs_v1, stats)
  S(fof_plan_v1) O(fof_plan_v1, grid_cells) O(fof_plan_v1, workspace_bytes)
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
        "fof_gen_v1": L.Gen, "fof_ghost_v1": L.Ghost, "fof_moments_v1": L.Moments,
        "fof_outputs_v1": L.Outputs, "fof_plan_v1": L.Plan,
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


def test_domain_error_on_host_before_device_work():
    # host-resident input outside [rmin_buff, rmax_buff] (or NaN) fails on
    # the host, with or without a GPU
    P = np.zeros(100, L.PART_DTYPE)
    P["Pos"] = 0.5
    d = L.domain([0] * 3, [1] * 3, [0] * 3, [1] * 3, [0.25] * 3, 0.1, 2)
    for bad in (1.5, -0.01, np.nan):
        Q = P.copy()
        Q["Pos"][57, 1] = bad
        assert L.lib().fof_link_gpu(None, C.byref(L.particles(Q)), C.byref(d), None, None, None,
                                    None, 0, None, None) == L.E_DOMAIN
    Q = P.copy()
    Q["Pos"][3] = 1.0  # rmax_buff itself is inside (the last cell)
    rc = L.lib().fof_link_gpu(None, C.byref(L.particles(Q)), C.byref(d), None, None, None, None, 0,
                              None, None)
    assert rc != L.E_DOMAIN


def fake_particles(n):
    p = L.Particles()
    p.version, p.location, p.base, p.n = L.ABI_VERSION, L.MEM_DEVICE, 4096, n
    p.stride, p.off_id, p.off_pos, p.off_vel = 32, 0, 8, 20
    return p


def volume(n, side, lcell, l=0.2, offset=0):
    v = L.Volume()
    v.offset, v.n = offset, n
    v.domain = L.domain([0] * 3, [side] * 3, [0] * 3, [side] * 3, [lcell] * 3, l, 20)
    return v


def test_plan_addressing_choice():
    # dense cell table while it stays below max(3N + 4M, ...) entries and 2^31
    n = 1 << 20
    p = L.plan(fake_particles(n), [volume(n, 100.0, 1.0)])
    assert p.sparse == 0 and p.cells == 101 ** 3 and p.n == n and p.workspace_bytes > 64 * n
    # far more cells than particles: occupied-cell addressing
    p = L.plan(fake_particles(n), [volume(n, 1000.0, 1.0)])
    assert p.sparse == 1 and p.cells == 0
    # 2^31 cells or more fall back to occupied-cell addressing even when the
    # particles would allow a table that large (u32 cell indices)
    n = 800_000_000
    side = 1300.0  # 1301^3 = 2.2e9 cells < 3n + 4M
    p = L.plan(fake_particles(n), [volume(n, side, 1.0)])
    assert p.grid_cells == 1301.0 ** 3 and p.grid_cells < 3 * n + 4 * 2 ** 20
    assert p.sparse == 1
    p = L.plan(fake_particles(n), [volume(n, 1200.0, 1.0)])  # 1.73e9 cells
    assert p.sparse == 0 and p.cells == 1201 ** 3
    # the option forces it
    p = L.plan(fake_particles(1000), [volume(1000, 10.0, 1.0)], L.opts(L.OPT_SPARSE))
    assert p.sparse == 1


def test_plan_batch_workspace_grows_with_volumes():
    vols = [volume(1000, 10.0, 1.0, offset=1000 * k) for k in range(4)]
    p4 = L.plan(fake_particles(4000), vols)
    p1 = L.plan(fake_particles(1000), vols[:1])
    assert p4.cells == 4 * p1.cells and p4.workspace_bytes > p1.workspace_bytes


def test_plan_internal_grid():
    # a constant l far inside the reference cells: each such axis is cut into
    # ceil(lcell / 6.9 l) cells (the same linked pairs, sub-cells stay on);
    # l at or above lcell (1 - 1e-6), or a variable l, keeps the reference grid
    n = 1 << 16
    assert L.plan(fake_particles(n), [volume(n, 100.0, 2.0)]).grid_cells == 101.0 ** 3  # 2 a side
    assert L.plan(fake_particles(n), [volume(n, 100.0, 3.0)]).grid_cells == 101.0 ** 3  # 3 a side
    assert L.plan(fake_particles(n), [volume(n, 100.0, 1.0)]).grid_cells == 101.0 ** 3  # as given
    assert L.plan(fake_particles(n), [volume(n, 100.0, 0.2)]).grid_cells == 501.0 ** 3  # l = lcell
    v = volume(n, 100.0, 3.0)
    vl = L.VarLinkTable([0.0, 0.0, 0.0], 3e4, np.full(4, 0.04))
    v.domain = L.domain([0] * 3, [100.0] * 3, [0] * 3, [100.0] * 3, [3.0] * 3, 0.2, 20, vl)
    assert L.plan(fake_particles(n), [v]).grid_cells == 35.0 ** 3  # ceil(100/3) + 1
