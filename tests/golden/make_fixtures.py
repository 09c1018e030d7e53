"""Write the particle-format fixtures under tests/golden/.

A small clumpy periodic box (10^3 particles in 15.625 Mpc/h, two Plummer
clumps, the generator of cm_halofinder_amd) written in the three input
formats of the reference (src_v2/read_data.c), by this script's own
struct-level writer -- independent of the library's readers:

* particles.ascii    "%llu %.9g x6" per line (:291-294; %.9g round-trips a float)
* particles.gadget1  Fortran records: 256-byte header (npart[1] = n, box size),
                     positions, velocities, u64 IDs (:321-365)
* particles.picola   two records per chunk: u32 count, then {u32 ID, Pos, Vel}
                     per particle, in chunks of 300 (:248-271)
* expected.npy       the records themselves (ID u64, Pos f4[3], Vel f4[3])

Run from the repository root: python tests/golden/make_fixtures.py
"""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cm_halofinder_amd import _lib as L  # noqa: E402
from cm_halofinder_amd import workloads as W  # noqa: E402


def rec(f, payload: bytes):
    f.write(struct.pack("<I", len(payload)))
    f.write(payload)
    f.write(struct.pack("<I", len(payload)))


def main():
    bx = W.plain(10, seed=71)
    P = L.gen_host(bx.gen())
    P["ID"] = P["ID"] * 7919 + 12345  # IDs unlike the indices
    np.save(os.path.join(HERE, "expected.npy"), P, allow_pickle=False)
    with open(os.path.join(HERE, "particles.ascii"), "w") as f:
        for r in P:
            vals = " ".join("%.9g" % float(x) for x in list(r["Pos"]) + list(r["Vel"]))
            f.write("%d %s\n" % (int(r["ID"]), vals))
    n = len(P)
    hdr = bytearray(256)
    struct.pack_into("<6I", hdr, 0, 0, n, 0, 0, 0, 0)
    struct.pack_into("<d", hdr, 24 + 8, 1.0)  # mass[1]
    struct.pack_into("<6I", hdr, 96, 0, n, 0, 0, 0, 0)  # npartTotal
    struct.pack_into("<i", hdr, 124, 1)  # num_files
    struct.pack_into("<d", hdr, 128, bx.box)  # BoxSize
    with open(os.path.join(HERE, "particles.gadget1"), "wb") as f:
        rec(f, bytes(hdr))
        rec(f, P["Pos"].astype("<f4").tobytes())
        rec(f, P["Vel"].astype("<f4").tobytes())
        rec(f, P["ID"].astype("<u8").tobytes())
    with open(os.path.join(HERE, "particles.picola"), "wb") as f:
        for a in range(0, n, 300):
            chunk = P[a:a + 300]
            rec(f, struct.pack("<I", len(chunk)))
            body = b"".join(struct.pack("<I", int(r["ID"]) & 0xFFFFFFFF) +
                            r["Pos"].astype("<f4").tobytes() + r["Vel"].astype("<f4").tobytes()
                            for r in chunk)
            rec(f, body)
    print(f"wrote {n} particles in three formats to {HERE}")


if __name__ == "__main__":
    main()
