"""MagLua's Checkpointer and the checkpointSave / checkpointLoad / checkpointToString /
checkpointFromString functions (common/checkpoint/checkpointer.cpp,
common/checkpoint/checkpointer_luafuncs.lua:143-175), over the SpinSystem byte codec of
libmsd.so (msd_system_encode / msd_system_decode).

A checkpoint holds n values written one after another as luamigrate.cpp:57-286
(_exportLuaVariable) writes them: an int32 Lua type tag, then
  nil       nothing
  boolean   int32 0/1
  number    double
  string    int32 length including the NUL, the bytes, the NUL
  table     magic char 111 and int32 0 (or 222 and the index of a table written earlier),
            int32 pair count, the key/value pairs, then the metatable (nil here)
  userdata  int32 type id (MurmurHash3 of the type name, seed 1000) and the object's own
            encoding, which opens with its magic char and index in the object list.
All integers and doubles are little-endian.  Transforms wrap the buffer behind a 20-byte
header (3-char signature and a NUL, then int32 new transform, old transform, new size,
old size): "Pack" (signature LUA, with n in front of the data), "Compress" (ZZZ, zlib
level 9 of the packed buffer) and "Base64" (the header and data as base64 text, which
opens with QjY).

Python values map to Lua values as: None - nil, bool - boolean, int/float - number,
str - string (UTF-8), list/tuple/dict - table (lists take keys 1..n), SpinSystem,
Exchange, Anisotropy, Anisotropy.Uniaxial, Dipole, Thermal, LLG.Cartesian, LLG.Quaternion,
AppliedField and AppliedField.Heterogeneous - userdata.  Decoded tables come back as dicts, numbers as floats (integral table keys as
ints).  Table pairs are written in list order or dict insertion order; Lua writes them in
its own table-traversal order, so streams holding tables with several keys match
MagLua's bytes only where the two orders agree.
"""
import base64
import ctypes as C
import numbers
import struct
import zlib

from ._lib import MagLuaError, MSD_EINVAL, MSD_EUNSUPPORTED, call
from .maglua import LLG, AppliedField, Anisotropy, Dipole, Exchange, SpinSystem, Thermal

# userdata this container writes and reads, by MagLua type name
_USERDATA = (("SpinSystem", SpinSystem), ("Exchange", Exchange), ("Anisotropy", Anisotropy), ("Dipole", Dipole),
             ("Thermal", Thermal), ("LLG.Cartesian", LLG.Cartesian), ("LLG.Quaternion", LLG.Quaternion),
             ("AppliedField", AppliedField), ("AppliedField.Heterogeneous", AppliedField.Heterogeneous),
             ("Anisotropy.Uniaxial", Anisotropy.Uniaxial))

LUA_TNIL, LUA_TBOOLEAN, LUA_TNUMBER, LUA_TSTRING, LUA_TTABLE, LUA_TUSERDATA = 0, 1, 3, 4, 5, 7
MAGIC_NEW, MAGIC_OLD = 111, 222
OP_UNKNOWN, OP_NONE, OP_PACK, OP_ZLIB, OP_B64 = -2, -1, 0, 1, 2
HEADER_SIZE = 20
_OP_NAMES = {0: "None", 1: "Pack", 2: "Compress", 3: "Base64", -1: "Unknown"}
_SIGNATURES = {OP_PACK: b"LUA\0", OP_ZLIB: b"ZZZ\0", OP_B64: b"B64\0"}


def hash32(name):
    """MagLua's type id of a type name (factory.cpp:123-131)."""
    v = C.c_int32()
    call("msd_hash32", name.encode(), C.byref(v))
    return v.value


def _err(msg, code=MSD_EINVAL):
    return MagLuaError(code, msg)


def b64encode(data):
    """base64.cpp's encoder: the standard alphabet with '=' padding, plus one '=' when no
    padding is needed (base64.cpp:60-87)."""
    s = base64.b64encode(bytes(data))
    return s + b"=" if len(data) % 3 == 0 else s


def b64decode(text):
    """base64.cpp's decoder: reads up to the first '=' or non-alphabet character."""
    alpha = b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    t = bytes(text)
    end = 0
    while end < len(t) and t[end] in alpha:
        end += 1
    t = t[:end]
    if len(t) % 4 == 1:
        t = t[:-1]
    return base64.b64decode(t + b"=" * (-len(t) % 4))


def _header(op, old, new_size, old_size):
    return _SIGNATURES[op] + struct.pack("<4i", op, old, new_size, old_size)


def state_of(data):
    """Checkpointer::currentStateS (checkpointer.cpp:558-578): the transform a buffer
    carries, from its first three characters."""
    if data is None:
        return OP_UNKNOWN
    sig = bytes(data[:3])
    return {b"LUA": OP_PACK, b"ZZZ": OP_ZLIB, b"QjY": OP_B64}.get(sig, OP_NONE)


def decode_header(data):
    """decodeHeader (checkpointer.cpp:386-430): (signature, new op, old op, new size, old size)."""
    op = state_of(data)
    if op == OP_B64:
        h = b64decode(bytes(data[:2 * HEADER_SIZE]))[:HEADER_SIZE]
    elif op in (OP_PACK, OP_ZLIB):
        h = bytes(data[:HEADER_SIZE])
    else:
        return "", op, op, 0, 0
    if len(h) < HEADER_SIZE:
        raise _err("Checkpointer: header truncated")
    ne, oe, ns, os_ = struct.unpack_from("<4i", h, 4)
    return h[:3].decode("latin-1"), ne, oe, ns, os_


class _Reader:
    def __init__(self, data):
        self.d, self.at = bytes(data), 0

    def take(self, k):
        if k < 0 or self.at + k > len(self.d):
            raise _err("Checkpointer: data stream ends early")
        v = self.d[self.at:self.at + k]
        self.at += k
        return v

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def ch(self):
        return self.take(1)[0]


class Checkpointer:
    """Checkpointer (common/checkpoint/checkpointer.cpp): values are added while the buffer
    is untransformed, then transformed for a file or a string and untransformed back."""

    def __init__(self):
        self.clear()

    def clear(self):
        self._buf = b""        # current data: raw values, or a transformed buffer
        self._pos = 0          # b->pos: length of the raw values
        self._n = 0
        self._objs = []        # the stream's object list (userdata and their arrays)
        self._tabs = []        # tables written so far, for repeats

    # state
    def _state(self):
        return OP_NONE if not self._buf else state_of(self._buf)

    def _internal_size(self):
        """Checkpointer::internalStateSize (checkpointer.cpp:601-621)."""
        if self._state() == OP_NONE:
            return self._pos
        return decode_header(self._buf)[3]

    def _transformation(self):
        return self._state() + 1

    def transformation(self):
        return _OP_NAMES.get(self._transformation(), "ERROR")

    def decodeHeader(self):
        sig, ne, oe, ns, os_ = decode_header(self._buf) if self._buf else ("", OP_NONE, OP_NONE, 0, 0)
        return sig, _OP_NAMES.get(ne + 1), _OP_NAMES.get(oe + 1), ns, os_

    # values
    def _export(self, v, out):
        if v is None:
            out += struct.pack("<i", LUA_TNIL)
        elif isinstance(v, bool):
            out += struct.pack("<ii", LUA_TBOOLEAN, int(v))
        elif isinstance(v, numbers.Real):
            out += struct.pack("<id", LUA_TNUMBER, float(v))
        elif isinstance(v, (str, bytes)):
            b = v.encode("utf-8", "surrogateescape") if isinstance(v, str) else bytes(v)
            b = b.split(b"\0", 1)[0]          # strlen, as lua_tostring's C string reads it
            out += struct.pack("<ii", LUA_TSTRING, len(b) + 1) + b + b"\0"
        elif isinstance(v, (list, tuple, dict)):
            out += struct.pack("<i", LUA_TTABLE)
            for k, t in enumerate(self._tabs):
                if t is v:
                    out += struct.pack("<Bi", MAGIC_OLD, k)
                    return
            self._tabs.append(v)
            pairs = list(v.items()) if isinstance(v, dict) else [(i + 1, x) for i, x in enumerate(v)]
            out += struct.pack("<Bii", MAGIC_NEW, 0, len(pairs))
            for k, x in pairs:
                self._export(k, out)
                self._export(x, out)
            out += struct.pack("<i", LUA_TNIL)   # no metatable
        elif type(v) in tuple(c for _, c in _USERDATA):
            name = next(nm for nm, cls in _USERDATA if type(v) is cls)
            out += struct.pack("<ii", LUA_TUSERDATA, hash32(name))
            for k, o in enumerate(self._objs):
                if o is v:
                    out += struct.pack("<Bi", MAGIC_OLD, k)
                    return
            data, k = v.encode(len(self._objs))
            self._objs += [v] + [None] * (k - 1)
            out += data
        else:
            raise _err(f"Checkpointer: cannot export a {type(v).__name__}", MSD_EUNSUPPORTED)

    def _import(self, r, objs, tabs):
        t = r.i32()
        if t == LUA_TNIL:
            return None
        if t == LUA_TBOOLEAN:
            return bool(r.i32())
        if t == LUA_TNUMBER:
            return r.f64()
        if t == LUA_TSTRING:
            s = r.take(r.i32())
            return s.split(b"\0", 1)[0].decode("utf-8", "surrogateescape")
        if t == LUA_TTABLE:
            magic, pos = r.ch(), r.i32()
            if magic == MAGIC_OLD:
                if not 0 <= pos < len(tabs):
                    raise _err("Checkpointer: malformed data stream (table index)")
                return tabs[pos]
            if magic != MAGIC_NEW:
                raise _err("Checkpointer: malformed data stream (table magic)")
            tab = {}
            tabs.append(tab)
            for _ in range(r.i32()):
                k = self._import(r, objs, tabs)
                x = self._import(r, objs, tabs)
                if isinstance(k, float) and k.is_integer():
                    k = int(k)
                tab[k] = x
            self._import(r, objs, tabs)          # metatable: not kept
            return tab
        if t == LUA_TUSERDATA:
            type_id = r.i32()
            cls = next((c for nm, c in _USERDATA if hash32(nm) == type_id), None)
            if cls is None:
                raise _err(f"Checkpointer: userdata of type id {type_id} is not supported", MSD_EUNSUPPORTED)
            if r.d[r.at] == MAGIC_OLD:
                r.take(1)
                pos = r.i32()
                if not 0 <= pos < len(objs) or objs[pos] is None:
                    raise _err("Checkpointer: malformed data stream (object index)")
                return objs[pos]
            o, used, k = cls.decode(r.d[r.at:], len(objs))
            r.at += used
            objs += [o] + [None] * (k - 1)
            return o
        raise _err(f"Checkpointer: Lua type {t} is not supported", MSD_EUNSUPPORTED)

    def add(self, v):
        if self._state() != OP_NONE:
            raise _err("Cannot add data to Checkpointer with internal encoding")
        out = bytearray(self._buf)
        self._export(v, out)
        self._buf = bytes(out)
        self._pos = len(self._buf)
        self._n += 1

    def addTable(self, t):
        """Adds t[1..maxn(t)] (checkpointer_luafuncs.lua addTable)."""
        if isinstance(t, dict):
            m = max([k for k in t if isinstance(k, int) and not isinstance(k, bool)] or [0])
            vals = [t.get(i) for i in range(1, m + 1)]
        else:
            vals = list(t)
        for v in vals:
            self.add(v)

    def get(self):
        """All values, in order (Lua returns them as multiple results)."""
        if self._state() != OP_NONE:
            raise _err("Cannot get data from Checkpointer with internal encoding")
        r, objs, tabs = _Reader(self._buf[:self._pos]), [], []
        return tuple(self._import(r, objs, tabs) for _ in range(self._n))

    def getTable(self):
        return list(self.get())

    # transforms (Checkpointer::operate_data / deoperate_data, checkpointer.cpp:284-556)
    def transform(self, *names):
        for name in names:
            name = "" if name is None else str(name)
            k = next((i for i in (1, 2, 3) if _OP_NAMES[i].lower() == name.lower()), 0)
            if k == 0:
                raise _err(f"Failed to match `{name}' to a valid operation")
            self._operate(k - 1)

    def _operate(self, op):
        old = self._state()
        old_size = self._internal_size()
        data = self._buf[:old_size]
        if op == OP_PACK:
            extra = struct.pack("<i", self._n) if old == OP_NONE else b""
            self._buf = _header(OP_PACK, old, HEADER_SIZE + len(extra) + old_size, old_size) + extra + data
        elif op == OP_ZLIB:
            if old != OP_PACK:
                raise _err("Must be in the Packed state to compress")
            z = zlib.compress(data, 9)
            self._buf = _header(OP_ZLIB, OP_PACK, HEADER_SIZE + len(z), old_size) + z
        elif op == OP_B64:
            L = HEADER_SIZE + old_size
            new_size = 4 * ((L + 2) // 3) + (1 if L % 3 == 0 else 0)
            self._buf = b64encode(_header(OP_B64, old, new_size, old_size) + data)

    def untransform(self):
        op = self._state()
        if op == OP_NONE:
            return
        sig, ne, oe, ns, os_ = decode_header(self._buf)
        if oe == OP_NONE and ne == OP_PACK:
            self._n = struct.unpack_from("<i", self._buf, HEADER_SIZE)[0]
            self._buf = self._buf[HEADER_SIZE + 4:HEADER_SIZE + 4 + os_]
            self._pos = os_
            self._objs, self._tabs = [], []
            return
        if ne == OP_PACK:
            self._buf = self._buf[HEADER_SIZE:HEADER_SIZE + os_]
        elif ne == OP_ZLIB:
            try:
                d = zlib.decompressobj().decompress(self._buf[HEADER_SIZE:])
            except zlib.error:
                d = b""
            if len(d) != os_:
                raise _err("Error in compressed datastream.")
            self._buf = d
        elif ne == OP_B64:
            self._buf = b64decode(self._buf)[HEADER_SIZE:HEADER_SIZE + os_]

    # strings and files
    def toString(self):
        if self._state() != OP_B64:
            raise _err("Internal state cannot be converted to a string")
        return self._buf.decode("ascii")

    def fromString(self, s):
        self.clear()
        b = s.encode("ascii") if isinstance(s, str) else bytes(s)
        if len(b) <= 6:
            raise _err("Data stream too short to identify")
        if not b.startswith(b"QjY"):
            raise _err("Data stream not recognized")
        self._buf = b

    def saveToFile(self, fn):
        if self._state() < 0:
            raise _err("Internal state must be in encode state")
        with open(fn, "wb") as f:
            f.write(self._buf[:decode_header(self._buf)[3]])

    def loadFromFile(self, fn):
        self.clear()
        with open(fn, "rb") as f:
            self._buf = f.read()

    def checksum(self):
        """Adler-32 of the first b->pos bytes of the buffer (checkpointer.cpp:589-599)."""
        return zlib.adler32(self._buf[:self._pos])

    def copy(self):
        c = Checkpointer()
        c._buf, c._pos, c._n = self._buf, self._pos, self._n
        c._objs, c._tabs = list(self._objs), list(self._tabs)
        return c


def checkpointSave(fn, *values):
    """checkpointer_luafuncs.lua:143-148: values packed (not compressed) into a file."""
    cp = Checkpointer()
    cp.addTable(values)
    cp.transform("Pack")
    cp.saveToFile(fn)


def checkpointLoad(fn):
    """checkpointer_luafuncs.lua:150-159: every transform undone; the values as a tuple."""
    cp = Checkpointer()
    cp.loadFromFile(fn)
    while cp._transformation() > 0:
        cp.untransform()
    return cp.get()


def checkpointToString(*values):
    cp = Checkpointer()
    cp.addTable(values)
    cp.transform("Pack", "Base64")
    return cp.toString()


def checkpointFromString(txt):
    cp = Checkpointer()
    cp.fromString(txt)
    while cp._transformation() > 0:
        cp.untransform()
    return cp.get()
