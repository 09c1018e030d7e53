"""Subsampling (src_v2/FOF.c:172-284) and its file (Subsample, :432-544)
without a GPU: the oracle restatement against hand-built rings with scripted
draws (known answers worked out from the reference's loop by hand), then the
library's host implementation against the oracle on random rings and draws."""
import ctypes
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import formats as F

BOX_LO, BOX_HI = (0.0, 0.0, 0.0), (10.0, 10.0, 10.0)

# Ten particles in four groups (1-based, ring order as given):
#   1 -> 4 -> 8 -> 1           (3 members)
#   2 -> 2                      (alone)
#   3 -> 6 -> 10 -> 5 -> 3      (4 members, not in ascending order)
#   7 -> 9 -> 7                 (2 members)
NEXT_OUT = np.array([4, 2, 6, 8, 3, 10, 9, 1, 7, 5], np.uint32)  # next_out[i-1] = next[i]


def records(moves=()):
    r = np.zeros(10, L.PART_DTYPE)
    r["ID"] = 1000 + np.arange(10)
    r["Pos"][:, 0] = 1.0 + 0.5 * np.arange(10)
    r["Pos"][:, 1:] = 5.0
    r["Vel"] = np.arange(30, dtype=np.float32).reshape(10, 3) * 10.0
    for i1, axis, v in moves:  # (1-based particle, axis, coordinate)
        r["Pos"][i1 - 1, axis] = v
    return r


def oracle_run(r, samp_in, samp_out, nphalomin, draws, closed_upper=0):
    pos = np.ascontiguousarray(r["Pos"], np.float32)
    return O.subsample(NEXT_OUT - 1, pos, samp_in, samp_out, BOX_LO, BOX_HI, nphalomin, draws,
                       closed_upper)


def test_oracle_every_member_at_rate_one():
    # heads 1, 2, 3, 7 in ascending order; each walk starts at the head's
    # successor and ends on the head; one draw per particle visited
    ih, oh, used = oracle_run(records(), 1.0, 1.0, 3.0, np.zeros(20))
    assert ih.tolist() == [4, 8, 1, 6, 10, 5, 3]
    assert oh.tolist() == [2, 9, 7]
    assert used == 10


def test_oracle_alternate_draws_and_rate_off():
    # rate 0.5 in halos, none outside: draws 0.1, 0.9, 0.1, ... keep every
    # other particle visited; groups outside halos take no draws
    draws = np.tile([0.1, 0.9], 10)
    ih, oh, used = oracle_run(records(), 0.5, 0.0, 3.0, draws)
    assert ih.tolist() == [4, 1, 10, 3]
    assert oh.tolist() == []
    assert used == 7


def test_oracle_outside_member_ends_its_group():
    # particle 10 outside the box: group 3's walk stops there (6 kept, the
    # draw of 10 taken), the later group is walked in full
    ih, oh, used = oracle_run(records([(10, 0, 12.0)]), 1.0, 1.0, 3.0, np.zeros(20))
    assert ih.tolist() == [4, 8, 1, 6]
    assert oh.tolist() == [2, 9, 7]
    assert used == 8
    # below rmin too
    ih, oh, used = oracle_run(records([(8, 2, -0.5)]), 1.0, 0.0, 3.0, np.zeros(20))
    assert ih.tolist() == [4, 6, 10, 5, 3] and used == 2 + 4


def test_oracle_upper_face():
    # particle 9 exactly on rmax[0]: outside with the half-open box
    # (PERIODIC), inside on the last rank along x without PERIODIC
    r = records([(9, 0, 10.0)])
    ih, oh, used = oracle_run(r, 0.0, 1.0, 3.0, np.zeros(20))
    assert oh.tolist() == [2] and used == 2
    ih, oh, used = oracle_run(r, 0.0, 1.0, 3.0, np.zeros(20), closed_upper=1)
    assert oh.tolist() == [2, 9, 7] and used == 3
    ih, oh, used = oracle_run(r, 0.0, 1.0, 3.0, np.zeros(20), closed_upper=6)  # y and z only
    assert oh.tolist() == [2] and used == 2


def test_oracle_size_cut_is_strict_less():
    # nphalo < nphalomin (src_v2/FOF.c:202): 3 members are a halo at 3.0,
    # not at 3.5
    ih, oh, used = oracle_run(records(), 1.0, 0.0, 3.0, np.zeros(20))
    assert ih.tolist() == [4, 8, 1, 6, 10, 5, 3]
    ih, oh, used = oracle_run(records(), 1.0, 0.0, 3.5, np.zeros(20))
    assert ih.tolist() == [6, 10, 5, 3] and used == 4


class Draws:
    def __init__(self, u):
        self.u, self.k = np.asarray(u, np.float64), 0

    def __call__(self):
        v = self.u[self.k]
        self.k += 1
        return v


def lib_run(r, samp_in, samp_out, nphalomin, draws, closed_upper=0, next_out=NEXT_OUT, **kw):
    d = Draws(draws)
    rc, ih, oh = F.subsample(r, next_out, samp_in, samp_out, BOX_LO, BOX_HI, nphalomin, d,
                             closed_upper, **kw)
    return rc, ih, oh, d.k


@pytest.mark.parametrize("moves,samp_in,samp_out,nphalomin,closed", [
    ((), 1.0, 1.0, 3.0, 0),
    ((), 0.5, 0.0, 3.0, 0),
    (((10, 0, 12.0),), 1.0, 1.0, 3.0, 0),
    (((9, 0, 10.0),), 0.0, 1.0, 3.0, 1),
    (((9, 0, 10.0),), 0.0, 1.0, 3.0, 0),
    ((), 1.0, 0.0, 3.5, 0),
])
def test_library_hand_cases(moves, samp_in, samp_out, nphalomin, closed):
    r = records(moves)
    draws = np.tile([0.1, 0.9], 10) if samp_in == 0.5 else np.zeros(20)
    rc, ih, oh, used = lib_run(r, samp_in, samp_out, nphalomin, draws, closed)
    oih, ooh, oused = oracle_run(r, samp_in, samp_out, nphalomin, draws, closed)
    assert rc == 0
    assert ih.tolist() == oih.tolist() and oh.tolist() == ooh.tolist() and used == oused


def random_case(seed, n=3000):
    rng = np.random.default_rng(seed)
    # groups of 1-40 members with random ring orders
    labels = np.empty(n, np.int64)
    i, g = 0, 0
    while i < n:
        s = int(min(n - i, rng.integers(1, 41) if rng.random() < 0.5 else 1))
        labels[i:i + s] = g
        i, g = i + s, g + 1
    perm = rng.permutation(n)
    lab = labels[perm]
    next_out = np.zeros(n, np.uint32)
    for gg in np.unique(lab):
        mem = np.flatnonzero(lab == gg)
        order = rng.permutation(mem)
        next_out[order] = np.roll(order, -1) + 1
    r = np.zeros(n, L.PART_DTYPE)
    r["ID"] = rng.integers(0, 2**63, n, dtype=np.uint64)
    r["Pos"] = rng.uniform(-0.5, 10.5, (n, 3)).astype(np.float32)
    r["Pos"][rng.random(n) < 0.02, 0] = 10.0  # on the upper face
    r["Vel"] = rng.normal(0, 300, (n, 3)).astype(np.float32)
    return r, next_out, rng.random(n)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_library_matches_oracle_on_random_rings(seed):
    r, next_out, draws = random_case(seed)
    pos = np.ascontiguousarray(r["Pos"], np.float32)
    for samp_in, samp_out, nphalomin, closed in [(0.3, 0.05, 20.0, 0), (0.9, 0.5, 5.0, 5),
                                                 (1.0, 0.0, 1.0, 7), (0.0, 1.0, 2.0, 2)]:
        d = Draws(draws)
        rc, ih, oh = F.subsample(r, next_out, samp_in, samp_out, BOX_LO, BOX_HI, nphalomin, d, closed)
        oih, ooh, oused = O.subsample(next_out - 1, pos, samp_in, samp_out, BOX_LO, BOX_HI, nphalomin,
                                      draws, closed)
        assert rc == 0
        assert np.array_equal(ih, oih) and np.array_equal(oh, ooh) and d.k == oused


def test_library_errors():
    r = records()
    # a list past its capacity
    rc, ih, oh, used = lib_run(r, 1.0, 1.0, 3.0, np.zeros(20), in_cap=5)
    assert rc == L.E_RANGE and len(ih) == 5
    # not a ring: two particles pointing at one
    bad = NEXT_OUT.copy()
    bad[1] = 4
    rc, ih, oh, used = lib_run(r, 1.0, 1.0, 3.0, np.zeros(20), next_out=bad)
    assert rc == L.E_ARG
    bad = NEXT_OUT.copy()
    bad[0] = 11  # out of range
    rc, ih, oh, used = lib_run(r, 1.0, 1.0, 3.0, np.zeros(20), next_out=bad)
    assert rc == L.E_ARG


def test_write_subsample_layouts(tmp_path):
    r = records()
    idx = np.array([4, 8, 1, 6], np.uint32)
    sel = r[idx - 1]
    # ASCII, with and without the ID (src_v2/FOF.c:527-533)
    for flags in (0, F.OUT_PARTICLE_ID):
        p = tmp_path / f"a{flags}"
        assert F.write_subsample(p, flags, r, idx) == 0
        lines = open(p).read().splitlines()
        assert len(lines) == 4
        for k, ln in enumerate(lines):
            v = ln.split()
            if flags:
                assert int(v[0]) == int(sel["ID"][k]) and len(v) == 7
                v = v[1:]
            want = np.concatenate([sel["Pos"][k], sel["Vel"][k]])
            assert np.allclose([float(x) for x in v], want, atol=5e-7)
            assert ln.startswith(" " * (15 - len(ln.split()[0])))  # %15 fields
    # unformatted: count, positions, velocities, IDs as Fortran records
    for flags in (F.OUT_UNFORMATTED, F.OUT_UNFORMATTED | F.OUT_PARTICLE_ID):
        p = tmp_path / f"u{flags}"
        assert F.write_subsample(p, flags, r, idx) == 0
        b = open(p, "rb").read()
        m0, cnt, m1 = struct.unpack_from("<iIi", b, 0)
        assert (m0, cnt, m1) == (4, 4, 4)
        o = 12
        for field in ("Pos", "Vel"):
            (m,) = struct.unpack_from("<i", b, o)
            assert m == 48
            arr = np.frombuffer(b, np.float32, 12, o + 4).reshape(4, 3)
            assert np.array_equal(arr, sel[field])
            assert struct.unpack_from("<i", b, o + 4 + 48)[0] == 48
            o += 56
        if flags & F.OUT_PARTICLE_ID:
            assert struct.unpack_from("<i", b, o)[0] == 32
            assert np.array_equal(np.frombuffer(b, np.uint64, 4, o + 4), sel["ID"])
            o += 40
        assert o == len(b)
    # nothing to write: no file (nsubsamp > 0 guards it, src_v2/FOF.c:457)
    p = tmp_path / "empty"
    assert F.write_subsample(p, 0, r, np.zeros(0, np.uint32)) == 0
    assert not os.path.exists(p)


def test_struct_layout_matches_header(tmp_path):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src, exe = tmp_path / "p.c", tmp_path / "p"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fof_io.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(fof_subsample_v1),\n'
        "         offsetof(fof_subsample_v1, samp_out), offsetof(fof_subsample_v1, rmax),\n"
        "         offsetof(fof_subsample_v1, nphalomin), offsetof(fof_subsample_v1, closed_upper),\n"
        "         offsetof(fof_subsample_v1, reserved));\n"
        "  return 0;\n}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    S = F.Subsample
    assert got == [ctypes.sizeof(S), S.samp_out.offset, S.rmax.offset, S.nphalomin.offset,
                   S.closed_upper.offset, S.reserved.offset]


def test_records_without_id_and_writer_errors(tmp_path):
    # the 24-byte record (no PARTICLE_ID): Pos at 0, Vel at 12
    r = records([(10, 0, 12.0)])
    rec24 = np.zeros(10, np.dtype([("Pos", "<f4", 3), ("Vel", "<f4", 3)]))
    rec24["Pos"], rec24["Vel"] = r["Pos"], r["Vel"]
    P24 = L.particles(rec24, stride=24, off_id=-1, off_pos=0, off_vel=12)
    d = Draws(np.zeros(20))
    s = F.Subsample(1.0, 1.0, (ctypes.c_double * 3)(*BOX_LO), (ctypes.c_double * 3)(*BOX_HI), 3.0, 0, 0)
    ih = np.zeros(10, np.uint32)
    oh = np.zeros(10, np.uint32)
    ni, no = ctypes.c_uint64(), ctypes.c_uint64()
    fn = F.UNIFORM_FN(lambda _s: float(d()))
    lib = F._lib()
    rc = lib.fof_subsample(ctypes.byref(P24), NEXT_OUT.ctypes.data, ctypes.byref(s), fn, None,
                           ih.ctypes.data, 10, ctypes.byref(ni), oh.ctypes.data, 10, ctypes.byref(no))
    assert rc == 0
    assert ih[:ni.value].tolist() == [4, 8, 1, 6] and oh[:no.value].tolist() == [2, 9, 7]
    # IDs asked of records without one; an index outside 1..n
    assert lib.fof_write_subsample(str(tmp_path / "x").encode(), F.OUT_PARTICLE_ID, ctypes.byref(P24),
                                   ih.ctypes.data, 2) == L.E_ARG
    bad = np.array([1, 11], np.uint32)
    assert F.write_subsample(tmp_path / "y", 0, r, bad) == L.E_ARG
    # the 24-byte records written without IDs
    assert lib.fof_write_subsample(str(tmp_path / "z").encode(), 0, ctypes.byref(P24), ih.ctypes.data,
                                   ni.value) == 0
    vals = np.loadtxt(tmp_path / "z")
    assert np.allclose(vals[:, :3], rec24["Pos"][ih[:ni.value] - 1], atol=5e-7)
