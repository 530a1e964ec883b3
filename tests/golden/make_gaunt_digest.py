#!/usr/bin/env python3
"""Write tests/golden/gaunt_L20.json: the SHA-256 digest (and entry count) of
the reference's coupling tables, read as data from the reference's generated
Fortran include file src/include/gaunt__body.f90 (its 43,289 "T(i,j,k) = x"
literals, each a REAL(4) literal widened to REAL(8); src/gaunt.f90:14-17).

Each literal is parsed to float32 with the C library's strtof (correctly
rounded, as a Fortran REAL(4) literal is), and the digest is taken over the
same canonical "table i j k bits" lines (0-based indices, sorted) that
specfab_amd/tables/make_gaunt.py hashes.  A handful of sampled entries are
stored as well so a mismatch can be localised.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_gaunt_digest.py
"""
import ctypes
import hashlib
import json
import os
import re
import struct

REF = "/root/reference/src/include/gaunt__body.f90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gaunt_L20.json")
TABLES = {"GC": 0, "GCm": 1, "GC_m1": 2, "GC_p1": 3}

libc = ctypes.CDLL(None)
libc.strtof.restype = ctypes.c_float
libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def f32_bits(text):
    v = libc.strtof(text.encode(), None)
    return struct.unpack("<I", struct.pack("<f", v))[0]


def main():
    pat = re.compile(r"^\s*(GC|GCm|GC_m1|GC_p1)\((\d+),(\d+),(\d+)\)\s*=\s*(\S+)\s*$")
    entries = []
    with open(REF) as f:
        for line in f:
            m = pat.match(line)
            if not m:
                continue
            t = TABLES[m.group(1)]
            i, j, k = (int(m.group(n)) - 1 for n in (2, 3, 4))
            entries.append((t, i, j, k, f32_bits(m.group(5))))
    entries.sort()
    h = hashlib.sha256()
    for e in entries:
        h.update(("%d %d %d %d %08x\n" % e).encode())
    sample = [list(e) for e in entries[:: max(1, len(entries) // 64)]]
    json.dump({"source": "src/include/gaunt__body.f90", "entries": len(entries),
               "sha256": h.hexdigest(), "sample": sample}, open(OUT, "w"), indent=1)
    print("wrote", OUT, len(entries), h.hexdigest())


if __name__ == "__main__":
    main()
