"""Hand-assembled checkpoint byte streams, field by field from the reference's encoders
(data for the checkpoint tests: each line is one encoder call of the reference).

SPINSYSTEM_2x1x1: SpinSystem::encode (cpu/core/spinsystem.cpp:669-760) of a 2x1x1 system
with alpha 0.5, dt 0.01, gamma 1, time 0.25, spins (1,0,0) and (0,0,1), ms 1, no site
alpha/gamma, as the first object of a buffer.
VALUES_RAW: _exportLuaVariable (common/luabaseobject/luamigrate.cpp:57-286) of the values
1.5, "ab", true, nil, {2.0}; VALUES_PACKED: the same after the "Pack" transform
(common/checkpoint/checkpointer.cpp:455-479)."""
import struct


def i(v):
    return struct.pack("<i", v)


def d(v):
    return struct.pack("<d", v)


NEW = b"\x6f"               # ENCODE_MAGIC_NEW, 111 (luabaseobject.h:64)
SPINSYSTEM_ID = -687101837  # hash32("SpinSystem")
ARRAY_DOUBLE_ID = -1895565407  # hash32("Array.Double")

SPINSYSTEM_2x1x1 = b"".join([
    NEW, i(0),                               # ENCODE_PREAMBLE: new object, index 0
    b"\x00",                                 # version
    i(2), i(1), i(1),                        # nx ny nz
    d(0.5), d(0.01), d(1.0), d(0.25),        # alpha dt gamma time
    # x: encodeT<dArray> (LUA_TUSERDATA, type) then Array::encode: preamble, version,
    # dims, flag 0 and the raw values
    i(7), i(ARRAY_DOUBLE_ID), NEW, i(1), b"\x00", i(2), i(1), i(1), i(0), d(1.0), d(0.0),
    # y: every site equal -> flag 1 and the one value
    i(7), i(ARRAY_DOUBLE_ID), NEW, i(2), b"\x00", i(2), i(1), i(1), i(1), d(0.0),
    # z
    i(7), i(ARRAY_DOUBLE_ID), NEW, i(3), b"\x00", i(2), i(1), i(1), i(0), d(0.0), d(1.0),
    # ms
    i(7), i(ARRAY_DOUBLE_ID), NEW, i(4), b"\x00", i(2), i(1), i(1), i(1), d(1.0),
    i(0),                                    # site_alpha_exists
    i(0),                                    # site_gamma_exists
    i(0),                                    # extra data: partial form, 0 entries
])

VALUES_RAW = b"".join([
    i(3), d(1.5),                            # LUA_TNUMBER
    i(4), i(3), b"ab\x00",                   # LUA_TSTRING: strlen + 1, bytes, NUL
    i(1), i(1),                              # LUA_TBOOLEAN
    i(0),                                    # LUA_TNIL
    i(5), NEW, i(0), i(1),                   # LUA_TTABLE: new, 0, one pair
    i(3), d(1.0), i(3), d(2.0),              #   [1] = 2
    i(0),                                    #   metatable: nil
])

VALUES_PACKED = b"LUA\x00" + i(0) + i(-1) + i(20 + 4 + len(VALUES_RAW)) + i(len(VALUES_RAW)) + i(5) + VALUES_RAW

# Exchange::encode (cpu/exchange/spinoperationexchange.cpp:52-72) of a 2x1x1 Exchange with
# normalisation on, pbc (1,1,1) and one pathway 0 -> 1 of strength 0.5, as object 0
EXCHANGE_2x1x1 = b"".join([
    NEW, i(0),                               # ENCODE_PREAMBLE
    b"\x00", i(2), i(1), i(1), d(1.0),       # SpinOperation::encode: version, dims, scale
    b"\x01",                                 # version 1
    i(1),                                    # normalizeMoments
    i(1), i(1), i(1),                        # pbc
    i(1),                                    # pathway count
    i(0), i(1), d(0.5),                      # from, to, strength
])

# Dipole::encode (cpu/dipole/spinoperationdipole.cpp:35-41) of a 2x2x1 Dipole with
# strength 0.25, the default truncation and the identity unit cell, as object 3
DIPOLE_2x2x1 = b"".join([
    NEW, i(3),
    b"\x00", i(2), i(2), i(1), d(1.0),       # SpinOperation::encode
    b"\x00",                                 # LongRange version
    i(1000), i(1000), i(1000), i(1000),      # gmax
    d(0.25),                                 # g
    d(1.0), d(0.0), d(0.0), d(0.0), d(1.0), d(0.0), d(0.0), d(0.0), d(1.0),   # ABC
    b"\x00",                                 # Dipole version
])

# Anisotropy::encode (cpu/anisotropy/spinoperationanisotropy.cpp:191-208) of a 2x1x1
# Anisotropy with one op at site 1, axis z, K 0.5, as object 2
ANISOTROPY_2x1x1 = b"".join([
    NEW, i(2),
    b"\x00", i(2), i(1), i(1), d(1.0),       # SpinOperation::encode
    b"\x00",                                 # version
    i(1),                                    # op count
    i(1), d(0.0), d(0.0), d(1.0), d(0.5),    # site, axis, K
])

# Thermal::encode (cpu/thermal/spinoperationthermal.cpp:51-61) of a 2x1x1 Thermal with
# scale 2, temperature 0.75 and site 2 scaled by 0.5, as object 4: the scale array is
# written by its own encode (scale->encode: preamble as object 5, no Lua type tag)
THERMAL_2x1x1 = b"".join([
    NEW, i(4),
    b"\x00", i(2), i(1), i(1), d(2.0),       # SpinOperation::encode: version, dims, scale
    b"\x00",                                 # version
    d(0.75),                                 # temperature
    NEW, i(5), b"\x00", i(2), i(1), i(1), i(0), d(1.0), d(0.5),   # Array::encode of the scale
])

# LLG::encode (cpu/llg/llg.cpp:50-60) of an LLG with thermalOnlyFirstTerm off,
# renormalisation disabled, precession scale 0.5 and damping scale 2, as object 1
LLG_FLAGS = b"".join([
    NEW, i(1),                               # ENCODE_PREAMBLE (no SpinOperation part)
    b"\x01",                                 # version
    i(0),                                    # thermalOnlyFirstTerm
    i(1),                                    # disableRenormalization
    d(0.5), d(2.0),                          # precess_scale, damp_scale
])

# AppliedField::encode (cpu/appliedfield/spinoperationappliedfield.cpp:33-42) of a 2x1x1
# field (0.1, -0.2, 0.3) with scale 1.5, as object 0
APPLIEDFIELD_2x1x1 = b"".join([
    NEW, i(0),
    b"\x00", i(2), i(1), i(1), d(1.5),       # SpinOperation::encode
    b"\x00",                                 # version
    d(0.1), d(-0.2), d(0.3),                 # B
])

# AppliedField_Heterogeneous::encode (spinoperationappliedfield_heterogeneous.cpp:39-48)
# of a 2x1x1 field with site 1 at (0.5, 0, 0) and site 2 at (0.5, 0.25, -1), as object 2:
# the three component arrays written by their own encode, objects 3, 4 and 5
APPLIEDFIELD_HET_2x1x1 = b"".join([
    NEW, i(2),
    b"\x00", i(2), i(1), i(1), d(1.0),       # SpinOperation::encode
    b"\x00",                                 # version
    NEW, i(3), b"\x00", i(2), i(1), i(1), i(1), d(0.5),              # hx: every site 0.5
    NEW, i(4), b"\x00", i(2), i(1), i(1), i(0), d(0.0), d(0.25),     # hy
    NEW, i(5), b"\x00", i(2), i(1), i(1), i(0), d(0.0), d(-1.0),     # hz
])

# AnisotropyUniaxial::encode (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp:181-199)
# of a 2x1x1 operator with ops (site 1, z, K1 0.5, K2 0) and (site 0, x, 0.25, 0.125)
# in that order, scale 2, as object 1
UNIAXIAL_2x1x1 = b"".join([
    NEW, i(1),
    b"\x00", i(2), i(1), i(1), d(2.0),       # SpinOperation::encode
    b"\x00",                                 # version
    i(2),                                    # op count
    i(1), d(0.0), d(0.0), d(1.0), d(0.5), d(0.0),
    i(0), d(1.0), d(0.0), d(0.0), d(0.25), d(0.125),
])
