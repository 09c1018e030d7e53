"""A plain C caller, written as INTEGRATION.md patches src_v2/FOF.c, compiled
against include/ and linked with libcmfof.so: it reads a Gadget-1 fixture,
calls fof_link_gpu in place of src_v2/FOF.c:44-161 and walks next[1..n]
with the reference's retrieval loop and Checkhalo, then subsamples the ring
through fof_subsample with its own generator.  Its halos and lists must be
the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from cm_halofinder_amd import _lib as L
from cm_halofinder_amd import formats as F

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_caller_walks_ring_like_reference(tmp_path):
    src = os.path.join(ROOT, "tests", "c_caller", "fof_caller.c")
    exe = str(tmp_path / "fof_caller")
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run(["gcc", "-O2", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                    "-L", libdir, "-lcmfof", f"-Wl,-rpath,{libdir}"], check=True)
    gad = os.path.join(ROOT, "tests", "golden", "particles.gadget1")
    box, pside, l, nmin = 15.625, 10, 0.3125, 5
    samp_in, samp_out, seed = 0.3, 0.05, 12345
    r = subprocess.run([exe, gad, str(box), str(pside), str(l), str(nmin), str(samp_in), str(samp_out),
                        str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [line.split() for line in r.stdout.strip().splitlines()]
    got = [g for g in lines if g[0] != "S"]
    sub = {g[1]: [int(x) for x in g[3:]] for g in lines if g[0] == "S"}
    P, _ = F.read_gadget1(gad)
    od = O.make_domain([0] * 3, [box] * 3, [0] * 3, [box] * 3, [box / pside] * 3, l, nmin)
    lab, _ = O.link_uf(P["Pos"], od)
    oh = O.retrieve(O.ring_from_labels(lab), P["Pos"], P["Vel"], od)
    oh = oh[oh["owned"] == 1]
    assert len(got) == len(oh) and len(oh) > 0
    for g, h in zip(got, oh):
        assert int(g[0]) == h["head"] + 1 and int(g[1]) == h["n"]
        assert np.array_equal(np.array([float(x) for x in g[2:]]), h["sum_pos"] / h["n"])
    # the caller's generator, replayed: x <- a x + c (mod 2^64), u = (x >> 11) / 2^53
    x, draws = seed, []
    for _ in range(len(P)):
        x = (x * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        draws.append((x >> 11) / 9007199254740992.0)
    oih, ooh, _ = O.subsample(O.ring_from_labels(lab), np.ascontiguousarray(P["Pos"]), samp_in, samp_out,
                              [0] * 3, [box] * 3, nmin, np.array(draws))
    assert sub["in"] == oih.tolist() and sub["out"] == ooh.tolist()
    assert len(sub["in"]) > 0 and len(sub["out"]) > 0
