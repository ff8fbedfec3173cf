"""CPU checks of the checkpoint byte stream: the type-id hash against published vectors
and the reference's own MurmurHash3 build, the oracle's SpinSystem stream against a
hand-assembled one, and the Checkpointer's values and transforms (no GPU calls)."""
import os
import random
import struct
import zlib

import pytest

import maglua_amd as M
from maglua_amd import checkpoint as CP
from oracle import checkpoint_oracle as O
from tests.golden import checkpoint_streams as G

REF_MURMUR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                          "libref_murmur.so")

# MurmurHash3_x86_32 published verification values (key, seed, hash)
PUBLISHED = [
    (b"", 0, 0x00000000), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\0\0\0\0", 0, 0x2362F9DE), (b"aaaa", 0x9747B28C, 0x5A97808A), (b"abc", 0, 0xB3DD93FA),
    (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
]


def test_murmur_oracle_published_vectors():
    for key, seed, h in PUBLISHED:
        assert O.murmur3_x86_32(key, seed) == h
    # MagLua's tail rotation changes only keys whose length is not a multiple of 4
    for key, seed, h in PUBLISHED:
        same = O.murmur3_x86_32(key, seed, tail_rot=16) == h
        assert same == (len(key) % 4 == 0)


def _ref_lib():
    import ctypes as C
    if not os.path.exists(REF_MURMUR):
        pytest.skip("oracle/_ref not built (the reference's sources are absent)")
    L = C.CDLL(REF_MURMUR)

    def ref(b, seed):
        o = C.c_uint32()
        L.ref_murmur3_x86_32(b, len(b), C.c_uint32(seed), C.byref(o))
        return o.value
    return ref


def test_murmur_oracle_matches_reference_build():
    ref = _ref_lib()
    rng = random.Random(0)
    for _ in range(500):
        b = bytes(rng.getrandbits(8) for _ in range(rng.randint(0, 40)))
        s = rng.getrandbits(32)
        assert ref(b, s) == O.murmur3_x86_32(b, s, tail_rot=16)


def test_type_ids_of_the_library():
    """msd_hash32 (host code, no device) against the oracle and the reference build."""
    names = ["SpinSystem", "Array.Double", "Array.Float", "Exchange", "Anisotropy", "LLG.Cartesian", "Dipole",
             "Thermal", "Checkpointer", "x", "ab", "abc"]
    for nm in names:
        assert CP.hash32(nm) == O.hash32(nm)
    assert CP.hash32("SpinSystem") == G.SPINSYSTEM_ID
    assert CP.hash32("Array.Double") == G.ARRAY_DOUBLE_ID
    if os.path.exists(REF_MURMUR):
        ref = _ref_lib()
        for nm in names:
            r = ref(nm.encode(), 1000)
            assert CP.hash32(nm) == (r - (1 << 32) if r >= 1 << 31 else r)


def test_oracle_spinsystem_stream_golden():
    b, k = O.spinsystem_bytes((2, 1, 1), 0.5, 0.01, 1.0, 0.25, [1.0, 0.0], [0.0, 0.0], [0.0, 1.0], [1.0, 1.0])
    assert b == G.SPINSYSTEM_2x1x1 and k == 5
    # the explicit form: -1 then one nil per site
    b2, _ = O.spinsystem_bytes((2, 1, 1), 0.5, 0.01, 1.0, 0.25, [1.0, 0.0], [0.0, 0.0], [0.0, 1.0], [1.0, 1.0],
                               extra="all")
    assert b2 == G.SPINSYSTEM_2x1x1[:-4] + struct.pack("<3i", -1, 0, 0)


def test_checkpointer_values_and_pack_golden():
    cp = M.Checkpointer()
    for v in (1.5, "ab", True, None, [2.0]):
        cp.add(v)
    assert cp._buf == G.VALUES_RAW
    assert cp.transformation() == "None"
    cp.transform("pack")                      # names match case-insensitively
    assert cp._buf == G.VALUES_PACKED
    assert cp.decodeHeader() == ("LUA", "Pack", "None", len(G.VALUES_PACKED), len(G.VALUES_RAW))
    cp.untransform()
    assert cp.get() == (1.5, "ab", True, None, {1: 2.0})


@pytest.mark.parametrize("ops", [("Pack",), ("Pack", "Compress"), ("Pack", "Base64"), ("Pack", "Compress", "Base64")])
def test_transforms_round_trip(ops):
    vals = (3.25, "checkpoint", False, {"a": 1.0, "b": [1.0, "x"]}, -7.0)
    cp = M.Checkpointer()
    cp.addTable(vals)
    cp.transform(*ops)
    sig, new, old, ns, os_ = cp.decodeHeader()
    assert new == {"Pack": "Pack", "Compress": "Compress", "Base64": "Base64"}[ops[-1]]
    if ops[-1] == "Compress":
        packed = M.Checkpointer()
        packed.addTable(vals)
        packed.transform("Pack")
        assert sig == "ZZZ" and old == "Pack" and os_ == len(packed._buf)
        assert zlib.decompress(cp._buf[20:ns]) == packed._buf
    if ops[-1] == "Base64":
        s = cp.toString()
        assert s.startswith("QjY") and ns == len(s)
        cp = M.Checkpointer()
        cp.fromString(s)
    while cp._transformation() > 0:
        cp.untransform()
    got = cp.get()
    assert got == (3.25, "checkpoint", False, {"a": 1.0, "b": {1: 1.0, 2: "x"}}, -7.0)


def test_base64_quirks():
    # base64.cpp appends one '=' when no padding is needed and stops reading at '='
    assert CP.b64encode(b"abc") == b"YWJj="
    assert CP.b64encode(b"ab") == b"YWI="
    assert CP.b64decode(b"YWJj=") == b"abc"
    assert CP.b64decode(b"YWI=trailing") == b"ab"
    # the header of a Base64 buffer decodes from its first 40 characters
    cp = M.Checkpointer()
    cp.add(1.0)
    cp.transform("Pack", "Base64")
    raw = CP.b64decode(cp.toString().encode())
    assert raw[:4] == b"B64\0" and struct.unpack_from("<4i", raw, 4)[:2] == (2, 0)


def test_repeated_table_is_a_reference():
    t = [1.0]
    cp = M.Checkpointer()
    cp.add(t)
    cp.add(t)
    tail = cp._buf[-9:]
    assert tail == struct.pack("<iBi", 5, 222, 0)
    a, b = cp.get()
    assert a is b and a == {1: 1.0}


def test_file_and_string_round_trip(tmp_path):
    fn = str(tmp_path / "cp.dat")
    M.checkpointSave(fn, 1.0, "x", {"k": 2.0})
    with open(fn, "rb") as f:
        assert f.read(4) == b"LUA\0"
    assert M.checkpointLoad(fn) == (1.0, "x", {"k": 2.0})
    s = M.checkpointToString(4.0, None, True)
    assert s.startswith("QjY")
    assert M.checkpointFromString(s) == (4.0, None, True)


def test_checkpointer_errors():
    cp = M.Checkpointer()
    cp.add(1.0)
    with pytest.raises(M.MagLuaError, match="Packed state"):
        cp.transform("Compress")
    with pytest.raises(M.MagLuaError, match="valid operation"):
        cp.transform("Zip")
    cp.transform("Pack")
    with pytest.raises(M.MagLuaError, match="internal encoding"):
        cp.add(2.0)
    with pytest.raises(M.MagLuaError, match="converted to a string"):
        cp.toString()
    with pytest.raises(M.MagLuaError, match="not recognized"):
        M.Checkpointer().fromString("abcdefgh")
    with pytest.raises(M.MagLuaError, match="cannot export"):
        M.Checkpointer().add(object())
    bad = M.Checkpointer()
    bad.add(1.0)
    bad.transform("Pack", "Compress")
    bad._buf = bad._buf[:24] + b"\0" * 8
    with pytest.raises(M.MagLuaError, match="compressed datastream"):
        bad.untransform()
