"""CPU restatement of MagLua's checkpoint byte stream, for the tests only (test
infrastructure: nothing in the package imports it).

  murmur3_x86_32     MurmurHash3_x86_32 as published by Austin Appleby (public domain),
                     the hash behind MagLua's type ids: hash32(name) uses seed 1000
                     (common/luabaseobject/factory.cpp:123-131) and MagLua's tail
                     rotation of 16 (MurmurHash3.cpp:143; the published hash has 15).
  darray_bytes       encodeT<dArray> (common/luabaseobject/luabaseobject.h:101-106) then
                     Array<double>::encode (cpu/array/array_core_cpu.h:200-224).
  spinsystem_bytes   SpinSystem::encode (cpu/core/spinsystem.cpp:669-760), from the
                     object's ENCODE_PREAMBLE (luabaseobject.h:63-72) on.

Pinned by: the published MurmurHash3 test vectors and the reference's own
MurmurHash3.cpp built into oracle/_ref (tests/test_checkpoint.py); the byte layout by a
hand-written literal stream of a 2x1x1 system in the same tests.
"""
import struct

import numpy as np

LUA_TUSERDATA = 7
MAGIC_NEW = 111


def murmur3_x86_32(data, seed, tail_rot=15):
    """tail_rot=15 is the published hash; MagLua's copy rotates the tail word by 16
    (MurmurHash3.cpp:143), which is what its type ids use."""
    data = bytes(data)
    c1, c2, m = 0xCC9E2D51, 0x1B873593, 0xFFFFFFFF

    def rotl(x, r):
        return ((x << r) | (x >> (32 - r))) & m

    h = seed & m
    nb = len(data) // 4
    for i in range(nb):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = rotl((k * c1) & m, 15) * c2 & m
        h = (rotl(h ^ k, 13) * 5 + 0xE6546B64) & m
    tail = data[4 * nb:]
    if tail:
        k = int.from_bytes(tail, "little")
        h ^= rotl((k * c1) & m, tail_rot) * c2 & m
    h ^= len(data)
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & m
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & m
    h ^= h >> 16
    return h


def hash32(name):
    """Signed, as MagLua stores it in an int."""
    h = murmur3_x86_32(name.encode(), 1000, tail_rot=16)
    return h - (1 << 32) if h >= (1 << 31) else h


def darray_bytes(index, dims, v):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    out = struct.pack("<ii", LUA_TUSERDATA, hash32("Array.Double"))
    out += struct.pack("<Bi", MAGIC_NEW, index) + b"\0" + struct.pack("<3i", *dims)
    if np.all(v == v[0]):
        return out + struct.pack("<id", 1, v[0])
    return out + struct.pack("<i", 0) + v.astype("<f8").tobytes()


def spinsystem_bytes(dims, alpha, dt, gamma, time, x, y, z, ms, site_alpha=None, site_gamma=None,
                     first_index=0, extra="partial"):
    """extra: "partial" writes the count 0 (no site has extra data); "all" writes -1 and
    a nil for every site, the form MagLua picks when more than half the sites carry data."""
    n = int(np.prod(dims))
    idx = first_index
    out = struct.pack("<Bi", MAGIC_NEW, idx) + b"\0" + struct.pack("<3i", *dims)
    out += struct.pack("<4d", alpha, dt, gamma, time)
    for a in (x, y, z, ms):
        idx += 1
        out += darray_bytes(idx, dims, a)
    for a in (site_alpha, site_gamma):
        out += struct.pack("<i", 0 if a is None else 1)
        if a is not None:
            idx += 1
            out += darray_bytes(idx, dims, a)
    if extra == "all":
        out += struct.pack("<i", -1) + struct.pack("<i", 0) * n
    else:
        out += struct.pack("<i", 0)
    return out, idx + 1 - first_index
