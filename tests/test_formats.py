"""File formats without a GPU: the readers of the reference's three particle
formats against fixtures written by tests/golden/make_fixtures.py (its own
struct-level writer), and the catalogue writers against an independent
parse of the reference's layouts (src_v2/FOF.c:546-957) and, where the
reference checkout is present, against its own reader of the binary halo
file (files/prepare_cmhalos.c, built into oracle/_ref)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import formats as F
from cm_halofinder_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
REF_READER = os.path.join(ROOT, "oracle", "_ref", "prepare_cmhalos")


def expected():
    return np.load(os.path.join(GOLD, "expected.npy"), allow_pickle=False)


def test_symbols_exported():
    import re

    txt = open(os.path.join(ROOT, "include", "fof_io.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"^\s*int\s+(fof_\w+)\s*\(", txt, flags=re.M))
    assert names == set(F.IO_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    for nm in names:
        assert re.search(rf"\bT {nm}\b", out), nm


def test_read_ascii():
    exp = expected()
    got = F.read_ascii(os.path.join(GOLD, "particles.ascii"))
    assert got.tobytes() == exp.tobytes()
    noid = F.read_ascii(os.path.join(GOLD, "particles.ascii"), with_id=False)
    # without IDs the first column is read as a coordinate: only the count holds
    assert len(noid) == len(exp) and np.array_equal(noid["ID"], np.arange(len(exp)))


def test_read_gadget1():
    exp = expected()
    got, hdr = F.read_gadget1(os.path.join(GOLD, "particles.gadget1"))
    assert got.tobytes() == exp.tobytes()
    assert hdr.npart[1] == len(exp) and hdr.npart_total[1] == len(exp)
    assert hdr.box_size == 15.625 and hdr.num_files == 1 and hdr.mass[1] == 1.0


def test_read_picola():
    exp = expected()
    got = F.read_picola(os.path.join(GOLD, "particles.picola"))
    assert got.tobytes() == exp.tobytes()  # IDs below 2^32 survive the u32 record


def test_reader_errors(tmp_path):
    g = open(os.path.join(GOLD, "particles.gadget1"), "rb").read()
    bad = tmp_path / "trunc.gadget1"
    bad.write_bytes(g[: len(g) - 100])
    assert F.read_rc("gadget1", bad, cap=1000) == L.E_ARG
    mark = bytearray(g)
    struct.pack_into("<I", mark, 256 + 8, 12)  # wrong size marker of the position block
    (tmp_path / "mark.gadget1").write_bytes(bytes(mark))
    assert F.read_rc("gadget1", tmp_path / "mark.gadget1", cap=1000) == L.E_ARG
    assert F.read_rc("gadget1", os.path.join(GOLD, "particles.gadget1"), cap=999) == L.E_RANGE
    assert F.read_rc("picola", os.path.join(GOLD, "particles.picola"), cap=999) == L.E_RANGE
    assert F.read_rc("ascii", os.path.join(GOLD, "particles.ascii"), cap=999) == L.E_RANGE
    (tmp_path / "x.ascii").write_text("1 2 3\n")
    assert F.read_rc("ascii", tmp_path / "x.ascii", cap=10) == L.E_ARG
    assert F.read_rc("ascii", tmp_path / "missing", cap=10) == L.E_ARG


# ---- catalogue writers -------------------------------------------------------
def halo_case(nphalomin=8, scale=16, seed=72):
    """A halo table, ring and second moments for a small periodic box, from
    the oracle (the engine's GPU output is checked against the same
    quantities in the GPU tests)."""
    bx = W.plain(scale, seed=seed, nphalomin=nphalomin)
    _, P, vols = W.host_domain(bx)
    d = vols[0].domain
    od = O.make_domain(d.rmin, d.rmax, d.rmin_buff, d.rmax_buff, d.lcell, d.linklength, d.nphalomin)
    lab, _ = O.link_uf(P["Pos"], od)
    ring = O.ring_from_labels(lab)
    oh = O.retrieve(ring, P["Pos"], P["Vel"], od)
    raw, cen = O.moments(ring, P["Pos"], P["Vel"], oh["head"])
    H = np.zeros(len(oh), L.HALO_DTYPE)
    for f in ("head", "n", "owned", "edge_warn", "sum_pos", "sum_vel"):
        H[f] = oh[f]
    M = np.zeros(len(oh), L.MOMENTS_DTYPE)
    M["pos2"], M["vel2"] = raw[:, :6], raw[:, 6:]
    M["cpos2"], M["cvel2"] = cen[:, :6], cen[:, 6:]
    return P, ring, H, M, oh, raw


def values(H, oh, raw, k, inertia, disp):
    v = list(oh["out"][k])  # (float)(sum * (1/n)), src_v2/FOF.c:837-842
    ine, dis = O.moments_reference_out(H["n"][k], H["sum_pos"][k], H["sum_vel"][k], raw[k])
    if inertia:
        v += list(ine)
    if disp:
        v += list(dis)
    return np.array(v, np.float32)


@pytest.mark.parametrize("inertia,disp", [(False, False), (True, False), (False, True), (True, True)])
def test_write_halos_unformatted(tmp_path, inertia, disp):
    P, ring, H, M, oh, raw = halo_case()
    flags = F.OUT_UNFORMATTED | (F.OUT_INERTIA if inertia else 0) | (F.OUT_DISPERSION if disp else 0)
    path = tmp_path / "halos.0"
    assert F.write_halos(path, flags, H, M) == 0
    b = path.read_bytes()
    own = np.flatnonzero(H["owned"])
    hs = 7 + 6 * inertia + 6 * disp
    # [4][nhalos][4] [nhalos*hs*4][blocks][same]
    assert struct.unpack_from("<3i", b, 0) == (4, len(own), 4)
    ln = struct.unpack_from("<i", b, 12)[0]
    assert ln == len(own) * hs * 4 and struct.unpack_from("<i", b, 16 + ln)[0] == ln
    assert len(b) == 16 + ln + 4
    blk = np.frombuffer(b, np.float32, len(own) * hs, 16).reshape(-1, hs)
    for row, k in zip(blk, own):
        assert row[0] == np.float32(H["n"][k])
        assert np.array_equal(row[1:], values(H, oh, raw, k, inertia, disp))


def test_write_halos_ascii_and_members(tmp_path):
    P, ring, H, M, oh, raw = halo_case()
    own = np.flatnonzero(H["owned"])
    nxt1 = (ring + 1).astype(np.uint32)  # the engine's 1-based ring
    path = tmp_path / "h.ascii"
    assert F.write_halos(path, F.OUT_INERTIA | F.OUT_DISPERSION, H, M) == 0
    lines = path.read_text().splitlines()
    assert len(lines) == len(own)
    for line, k in zip(lines, own):
        v = values(H, oh, raw, k, True, True)
        exp = "%15d" % H["n"][k] + "".join(" %15.6f" % float(x) for x in v)
        assert line == exp
    # members after each halo, in ring order from the head, with IDs
    path = tmp_path / "hp.ascii"
    assert F.write_halos(path, F.OUT_PARTICLES | F.OUT_PARTICLE_ID, H, None, P, nxt1) == 0
    lines = path.read_text().splitlines()
    assert lines[0] == "%15d" % len(own)
    i = 1
    for k in own:
        assert lines[i].split()[0] == str(H["n"][k])
        i += 1
        j = int(H["head"][k])
        for _ in range(H["n"][k]):
            r = P[j]
            exp = "%15d" % r["ID"] + "".join(" %15.6f" % float(x) for x in list(r["Pos"]) + list(r["Vel"]))
            assert lines[i] == exp
            i += 1
            j = int(ring[j])
        assert j == H["head"][k]
    assert i == len(lines)


def test_write_halos_unformatted_members(tmp_path):
    P, ring, H, M, oh, raw = halo_case()
    own = np.flatnonzero(H["owned"])
    path = tmp_path / "hp.bin"
    assert F.write_halos(path, F.OUT_UNFORMATTED | F.OUT_PARTICLES | F.OUT_PARTICLE_ID, H, None, P,
                         ring + 1) == 0
    b = path.read_bytes()
    assert struct.unpack_from("<3i", b, 0) == (4, len(own), 4)
    o = 12
    for k in own:
        assert struct.unpack_from("<i", b, o)[0] == 28
        row = np.frombuffer(b, np.float32, 7, o + 4)
        assert row[0] == np.float32(H["n"][k]) and np.array_equal(row[1:], oh["out"][k])
        o += 36
        j = int(H["head"][k])
        for _ in range(H["n"][k]):
            assert struct.unpack_from("<i", b, o)[0] == 32
            assert struct.unpack_from("<Q", b, o + 4)[0] == P["ID"][j]
            vals = np.frombuffer(b, np.float32, 6, o + 12)
            assert np.array_equal(vals[:3], P["Pos"][j]) and np.array_equal(vals[3:], P["Vel"][j])
            assert struct.unpack_from("<i", b, o + 36)[0] == 32
            o += 40
            j = int(ring[j])
    assert o == len(b)


@pytest.mark.skipif(not os.path.exists(REF_READER), reason="reference reader not built (no checkout)")
def test_reference_reader_reads_halo_file(tmp_path):
    # files/prepare_cmhalos.c reads the INERTIA layout (13 floats a halo) of
    # files <base>.0 .. <base>.(n-2) (its loop stops one short of the count it
    # is given) and prints "%13d\t" then "%13.6lf\t" per value
    P, ring, H, M, oh, raw = halo_case()
    base = str(tmp_path / "cat")
    assert F.write_halos(base + ".0", F.OUT_UNFORMATTED | F.OUT_INERTIA, H, M) == 0
    r = subprocess.run([REF_READER, base, "2"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    lines = open(base + "_reform_c.0").read().splitlines()
    own = np.flatnonzero(H["owned"])
    assert int(lines[0]) == len(own)
    for line, k in zip(lines[1:], own):
        v = values(H, oh, raw, k, True, False)
        exp = "%13d\t" % int(np.float32(H["n"][k])) + "".join("%13.6lf\t" % float(x) for x in v)
        assert line == exp
    assert len(lines) == len(own) + 1


def test_write_info(tmp_path):
    path = tmp_path / "cat.info"
    nh = np.array([5, 0, 7, 1, 2, 3, 4, 9], np.int32)
    assert F.write_info(path, [0, -10, 5], [100, 90, 45], 2, 2, 2, nh, np.arange(8), 2 * np.arange(8)) == 0
    lines = path.read_text().splitlines()
    assert lines[0].startswith("#    FILENUM      XMIN")
    assert len(lines) == 9
    # rank i = (tx, ty, tz) x-fastest; region x0 = tx (L/Nx) + Lmin (src_v2/FOF.c:939-947)
    assert lines[1 + 5] == "%12d %12.6f %12.6f %12.6f %12.6f %12.6f %12.6f %12d %12u %12u" % (
        5, 50.0, -10.0, 25.0, 100.0, 40.0, 45.0, 3, 5, 10)
