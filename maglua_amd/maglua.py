"""MagLua's hot-path object API over libmsd.so.

Names, argument meaning and error behaviour follow the Lua-visible API of the
reference modules so that scripts port line by line:

    ss  = SpinSystem.new(32, 32, 1)          -- SpinSystem.new(nx, ny, nz)
    ex  = Exchange.new(ss)
    ex:add({i,j}, {i+1,j}, J)                -- 1-based sites, periodic wrap
    llg:apply(ss, 0.5, ssP, ssC)             -- LLG:apply(from, scale, dmdt, to)

Sites are 1-based as in Lua (``{x,y,z}`` tuples, missing trailing coordinates
are 1).  Failures raise MagLuaError carrying msd_last_error(), where MagLua
raises luaL_error.  All state lives on the GPU; there is no CPU path.
"""
import ctypes as C
import os
import struct

import numpy as np

from ._lib import MagLuaError, call, lib

_vp = C.c_void_p


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _site3(p, dims):
    p = tuple(int(v) for v in p) if not np.isscalar(p) else (int(p),)
    p = p + (1,) * (3 - len(p))
    return p[0] - 1, p[1] - 1, p[2] - 1


def _index(p, dims):
    x, y, z = _site3(p, dims)
    nx, ny, nz = dims
    if not (0 <= x < nx and 0 <= y < ny and 0 <= z < nz):
        raise MagLuaError(-1, f"site ({x + 1}, {y + 1}, {z + 1}) is not part of the {nx}x{ny}x{nz} lattice")
    return x + nx * (y + ny * z)


def _dims_of(a):
    """Constructor arguments: (nx[,ny[,nz]]), ({nx,ny,nz}) or (ss) -- spinoperation.cpp:43-64."""
    if len(a) == 1 and isinstance(a[0], SpinSystem):
        return a[0].nx(), a[0].ny(), a[0].nz()
    if len(a) >= 1 and isinstance(a[0], (tuple, list)):
        a = tuple(a[0])
    d = tuple(int(v) for v in a) + (1, 1, 1)
    return d[0], d[1], d[2]


class _Handle:
    _destroy = None

    def __init__(self, h):
        self._h = h

    @property
    def h(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and self._destroy:
            try:
                getattr(lib(), self._destroy)(h)
            except Exception:
                pass
            self._h = None


def _encode_handle(fn, h, first_index):
    """An object's checkpoint bytes through msd_*_encode: (bytes, objects taken)."""
    n, k = C.c_int64(), C.c_int()
    call(fn, h, int(first_index), None, 0, C.byref(n), C.byref(k))
    buf = C.create_string_buffer(max(n.value, 1))
    call(fn, h, int(first_index), buf, n.value, C.byref(n), C.byref(k))
    return buf.raw[:n.value], k.value


def _decode_handle(cls, fn, data, first_index):
    """A new cls wrapping the handle msd_*_decode makes: (object, bytes read, objects taken)."""
    data = bytes(data)
    p, used, k = _vp(), C.c_int64(), C.c_int()
    call(fn, data, len(data), int(first_index), C.byref(p), C.byref(used), C.byref(k))
    o = cls.__new__(cls)
    _Handle.__init__(o, p.value)
    o._dims = struct.unpack_from("<3i", data, 6)
    return o, used.value, k.value


# ---------------------------------------------------------------- SpinSystem
class SpinSystem(_Handle):
    """SpinSystem (cpu/core/spinsystem.h:33-140): FP32 SoA spins + named field slots on the GPU."""

    _destroy = "msd_system_destroy"

    def __init__(self, *a, _h=None):
        if _h is None:
            nx, ny, nz = _dims_of(a)
            p = _vp()
            call("msd_system_create", nx, ny, nz, C.byref(p))
            _h = p.value
        super().__init__(_h)
        nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
        call("msd_system_dims", self._h, C.byref(nx), C.byref(ny), C.byref(nz))
        self._dims = (nx.value, ny.value, nz.value)

    @classmethod
    def new(cls, *a):
        return cls(*a)

    def nx(self):
        return self._dims[0]

    def ny(self):
        return self._dims[1]

    def nz(self):
        return self._dims[2]

    def nxyz(self):
        return self._dims[0] * self._dims[1] * self._dims[2]

    def copy(self):
        p = _vp()
        call("msd_system_copy", self._h, C.byref(p))
        return SpinSystem(_h=p.value)

    def copySpinsTo(self, other):
        call("msd_system_copy_spins_to", self._h, other.h)

    # checkpoint byte stream (SpinSystem::encode/decode, cpu/core/spinsystem.cpp:669-846)
    def encode(self, first_index=0):
        """The bytes SpinSystem::encode appends to a checkpoint buffer holding first_index
        objects already; returns (bytes, number of objects the system and its arrays take)."""
        cap = 242 + 48 * self.nxyz()
        buf = C.create_string_buffer(cap)
        n, k = C.c_int64(), C.c_int()
        call("msd_system_encode", self._h, int(first_index), buf, cap, C.byref(n), C.byref(k))
        return buf.raw[:n.value], k.value

    @classmethod
    def decode(cls, data, first_index=0):
        """Inverse of encode: returns (SpinSystem, bytes read, number of objects)."""
        data = bytes(data)
        p, used, k = _vp(), C.c_int64(), C.c_int()
        call("msd_system_decode", data, len(data), int(first_index), C.byref(p), C.byref(used), C.byref(k))
        return cls(_h=p.value), used.value, k.value

    # spins
    def setSpin(self, site, s, m=None):
        call("msd_system_set_spin", self._h, _index(site, self._dims), float(s[0]), float(s[1]), float(s[2]),
             -1.0 if m is None else abs(float(m)))

    def spin(self, *site):
        if len(site) == 1 and not np.isscalar(site[0]):
            site = site[0]
        out = (C.c_double * 4)()
        call("msd_system_get_spin", self._h, _index(site, self._dims), out)
        return out[0], out[1], out[2], out[3]

    def setSpins(self, x, y, z, m=None):
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (x, y, z)]
        mm = None if m is None else np.ascontiguousarray(m, dtype=np.float64).reshape(-1)
        for v in a + ([mm] if mm is not None else []):
            if v.size != self.nxyz():
                raise MagLuaError(-1, f"setSpins: expected {self.nxyz()} values, got {v.size}")
        call("msd_system_set_spins", self._h, _d(a[0]), _d(a[1]), _d(a[2]), None if mm is None else _d(mm))

    def spins(self):
        """(x, y, z, ms) as FP64 arrays shaped (nz, ny, nx)."""
        n = self.nxyz()
        o = [np.zeros(n) for _ in range(4)]
        call("msd_system_get_spins", self._h, *(_d(v) for v in o))
        shp = (self._dims[2], self._dims[1], self._dims[0])
        return tuple(v.reshape(shp) for v in o)

    def spinArrayX(self):
        return self.spins()[0]

    def spinArrayY(self):
        return self.spins()[1]

    def spinArrayZ(self):
        return self.spins()[2]

    def spinArrayM(self):
        return self.spins()[3]

    def randomize(self, seed):
        """Unit spins uniform on the sphere from the seeded counter-based stream."""
        call("msd_system_random_spins", self._h, int(seed))

    # parameters
    def _set(self, k, v):
        call("msd_system_set_param", self._h, k, float(v))

    def _get(self, k):
        v = C.c_double()
        call("msd_system_get_param", self._h, k, C.byref(v))
        return v.value

    def setAlpha(self, a):
        if np.isscalar(a):
            self._set(0, a)
        else:
            self.setSiteAlpha(a)

    def alpha(self):
        return self._get(0)

    def setGamma(self, g):
        if np.isscalar(g):
            self._set(1, g)
        else:
            self.setSiteGamma(g)

    def gamma(self):
        return self._get(1)

    def setTimeStep(self, dt):
        self._set(2, dt)

    def timeStep(self):
        return self._get(2)

    def setTime(self, t):
        self._set(3, t)

    def time(self):
        return self._get(3)

    def setSiteAlpha(self, a):
        if a is None:
            call("msd_system_set_site_alpha", self._h, None)
        else:
            v = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            call("msd_system_set_site_alpha", self._h, _d(v))

    def setSiteGamma(self, g):
        if g is None:
            call("msd_system_set_site_gamma", self._h, None)
        else:
            v = np.ascontiguousarray(g, dtype=np.float64).reshape(-1)
            call("msd_system_set_site_gamma", self._h, _d(v))

    # fields
    def resetFields(self):
        call("msd_system_reset_fields", self._h)

    zeroFields = resetFields

    def sumFields(self):
        call("msd_system_sum_fields", self._h)

    def field(self, slot="Total"):
        n = self.nxyz()
        o = [np.zeros(n) for _ in range(3)]
        call("msd_system_get_field", self._h, slot.encode(), *(_d(v) for v in o))
        shp = (self._dims[2], self._dims[1], self._dims[0])
        return tuple(v.reshape(shp) for v in o)

    def fieldArrayX(self, slot="Total"):
        return self.field(slot)[0]

    def fieldArrayY(self, slot="Total"):
        return self.field(slot)[1]

    def fieldArrayZ(self, slot="Total"):
        return self.field(slot)[2]

    def slotUsed(self, slot):
        u = C.c_int()
        call("msd_system_slot_used", self._h, slot.encode(), C.byref(u))
        return bool(u.value)

    # observables
    def netMoment(self):
        """Sums (not means), as spinsystem_luafuncs.lua:17-54 returns them."""
        out = (C.c_double * 8)()
        call("msd_system_net_moment", self._h, out)
        return tuple(out)

    def fieldDot(self, slot):
        v = C.c_double()
        call("msd_system_field_dot", self._h, slot.encode(), C.byref(v))
        return v.value

    def diff(self, other):
        out = (C.c_double * 4)()
        call("msd_system_diff", self._h, other.h, out)
        return tuple(out)


# ------------------------------------------------------------------ operators
class Exchange(_Handle):
    """Exchange (cpu/exchange/spinoperationexchange.cpp): path list, slot 'Exchange'."""

    _destroy = "msd_exchange_destroy"

    def __init__(self, *a):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_exchange_create", *self._dims, C.byref(p))
        super().__init__(p.value)

    new = classmethod(lambda cls, *a: cls(*a))

    # checkpoint byte stream (Exchange::encode/decode, spinoperationexchange.cpp:52-113)
    def encode(self, first_index=0):
        return _encode_handle("msd_exchange_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        return _decode_handle(cls, "msd_exchange_decode", data, first_index)

    def setPeriodicXYZ(self, px, py, pz):
        call("msd_exchange_set_pbc", self._h, int(bool(px)), int(bool(py)), int(bool(pz)))

    def normalizeMoments(self, yes=True):
        call("msd_exchange_set_normalize", self._h, int(bool(yes)))

    def setScale(self, s):
        call("msd_exchange_set_scale", self._h, float(s))

    def add(self, a, b, J=1.0):
        """ex:add({a}, {b}, J): couples from site a onto site b (1-based, periodic wrap)."""
        pa = (C.c_int * 3)(*_site3(a, self._dims))
        pb = (C.c_int * 3)(*_site3(b, self._dims))
        added = C.c_int()
        call("msd_exchange_add", self._h, pa, pb, float(J), C.byref(added))
        return bool(added.value)

    addPath = add

    def addPaths(self, frm, to, J):
        """Bulk addPath over 0-based site index arrays (a script's add loop)."""
        f = np.ascontiguousarray(frm, dtype=np.int32).reshape(-1)
        t = np.ascontiguousarray(to, dtype=np.int32).reshape(-1)
        j = np.ascontiguousarray(np.broadcast_to(np.asarray(J, dtype=np.float64), f.shape))
        if f.size != t.size:
            raise MagLuaError(-1, "addPaths: from/to lengths differ")
        call("msd_exchange_add_paths", self._h, f.ctypes.data_as(C.POINTER(C.c_int)),
             t.ctypes.data_as(C.POINTER(C.c_int)), _d(j), int(f.size))

    def numberOfPaths(self):
        n = C.c_int()
        call("msd_exchange_num_paths", self._h, C.byref(n))
        return n.value

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_exchange_apply", self._h, s.h)


class Anisotropy(_Handle):
    """Anisotropy (cpu/anisotropy/spinoperationanisotropy.cpp): per-site axis and K, slot 'Anisotropy'."""

    _destroy = "msd_anisotropy_destroy"

    def __init__(self, *a):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_anisotropy_create", *self._dims, C.byref(p))
        super().__init__(p.value)

    new = classmethod(lambda cls, *a: cls(*a))

    # checkpoint byte stream (Anisotropy::encode/decode, spinoperationanisotropy.cpp:191-240)
    def encode(self, first_index=0):
        return _encode_handle("msd_anisotropy_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        return _decode_handle(cls, "msd_anisotropy_decode", data, first_index)

    def setScale(self, s):
        call("msd_anisotropy_set_scale", self._h, float(s))

    def add(self, site, axis, K):
        ax = (C.c_double * 3)(*[float(v) for v in axis])
        added = C.c_int()
        call("msd_anisotropy_add", self._h, _index(site, self._dims), ax, float(K), C.byref(added))
        return bool(added.value)

    def addMany(self, sites, axes, K):
        """Bulk add over 0-based site indices; axes (count,3), K (count,)."""
        s = np.ascontiguousarray(sites, dtype=np.int32).reshape(-1)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(axes, dtype=np.float64), (s.size, 3)))
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(K, dtype=np.float64), s.shape))
        call("msd_anisotropy_add_many", self._h, s.ctypes.data_as(C.POINTER(C.c_int)), _d(a), _d(k), int(s.size))

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_anisotropy_apply", self._h, s.h)


class _AnisotropyUniaxial(_Handle):
    """Anisotropy.Uniaxial (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp): ops of
    (site, unit axis, K1, K2), slot 'UniaxialAnisotropy',
    h = n (2 K1 d + 4 K2 d^3) scale with d = sx nx + sy ny + sz nx as the reference
    computes it (its line 287 pairs sz with nx)."""

    _destroy = "msd_uniaxial_destroy"

    def __init__(self, *a):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_uniaxial_create", *self._dims, C.byref(p))
        super().__init__(p.value)

    new = classmethod(lambda cls, *a: cls(*a))

    def setScale(self, s):
        call("msd_uniaxial_set_scale", self._h, float(s))

    # checkpoint byte stream (AnisotropyUniaxial::encode/decode, :181-234)
    def encode(self, first_index=0):
        return _encode_handle("msd_uniaxial_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        return _decode_handle(cls, "msd_uniaxial_decode", data, first_index)

    def add(self, site, axis, K1=None, K2=None):
        """add({x,y,z}, {ax,ay,az}, K1[, K2]) (l_add, :387-431).  As in the reference's
        binding, a fourth number sets K2 to K1's value (it reads K1's stack slot twice)."""
        if K1 is None:
            raise MagLuaError(-1, "anisotropy needs strength")
        k2 = float(K1) if K2 is not None else 0.0
        ax = (C.c_double * 3)(*[float(v) for v in axis])
        added = C.c_int()
        call("msd_uniaxial_add", self._h, _index(site, self._dims), ax, float(K1), k2, C.byref(added))
        return bool(added.value)

    def numberOfAxes(self):
        n = C.c_int()
        call("msd_uniaxial_num_axes", self._h, C.byref(n))
        return n.value

    def _op(self, i):
        site, K1, K2 = C.c_int(), C.c_double(), C.c_double()
        ax = (C.c_double * 3)()
        call("msd_uniaxial_axis_at", self._h, int(i), C.byref(site), ax, C.byref(K1), C.byref(K2))
        return site.value, [ax[0], ax[1], ax[2]], K1.value, K2.value

    def axis(self, i):
        """i-th op, 1-based: (axis, K1, K2); l_axisat pushes the site too but returns
        only the top three values."""
        _, ax, k1, k2 = self._op(int(i) - 1)
        return ax, k1, k2

    def get(self, site):
        """The first op at the site (getAnisotropy, :126-139): (axis, K1, K2), or
        ({0,0,1}, 0, 0) when the site has none."""
        idx = _index(site, self._dims)
        for i in range(self.numberOfAxes()):
            s, ax, k1, k2 = self._op(i)
            if s == idx:
                return ax, k1, k2
        return [0.0, 0.0, 1.0], 0.0, 0.0

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_uniaxial_apply", self._h, s.h)


Anisotropy.Uniaxial = _AnisotropyUniaxial


def _decode_applied(cls, data, first_index, het):
    data = bytes(data)
    p, used, k = _vp(), C.c_int64(), C.c_int()
    call("msd_appliedfield_decode", data, len(data), int(first_index), int(het), C.byref(p), C.byref(used), C.byref(k))
    o = cls.__new__(cls)
    _Handle.__init__(o, p.value)
    o._dims = struct.unpack_from("<3i", data, 6)
    o._scale = struct.unpack_from("<d", data, 18)[0]
    return o, used.value, k.value


class _AppliedFieldHeterogeneous(_Handle):
    """AppliedField.Heterogeneous (cpu/appliedfield/spinoperationappliedfield_heterogeneous.cpp):
    a field per site, added (times scale) to slot 'Zeeman_Heterogeneous' by apply."""

    _destroy = "msd_appliedfield_destroy"

    def __init__(self, *a):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_appliedfield_create", *self._dims, 1, C.byref(p))
        super().__init__(p.value)
        self._scale = 1.0

    new = classmethod(lambda cls, *a: cls(*a))

    def set(self, p, h):
        """set({x,y,z} 1-based site, {hx,hy,hz}) (appliedfield_heterogeneous_luafuncs.lua:7-18)."""
        if not isinstance(p, (tuple, list)):
            raise MagLuaError(-1, "first argument needs to be a table")
        if not isinstance(h, (tuple, list)):
            raise MagLuaError(-1, "second argument needs to be a table")
        call("msd_appliedfield_set_site", self._h, _index(p, self._dims), float(h[0]), float(h[1]), float(h[2]))

    def get(self, p):
        if not isinstance(p, (tuple, list)):
            raise MagLuaError(-1, "first argument needs to be a table")
        v = (C.c_double * 3)()
        call("msd_appliedfield_get_site", self._h, _index(p, self._dims), v)
        return [v[0], v[1], v[2]]

    def setScale(self, s):
        self._scale = float(s)
        call("msd_appliedfield_set_scale", self._h, self._scale)

    def scale(self):
        return self._scale

    def slotName(self):
        return "Zeeman_Heterogeneous"

    # checkpoint byte stream (AppliedField_Heterogeneous::encode/decode, :39-72)
    def encode(self, first_index=0):
        return _encode_handle("msd_appliedfield_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        return _decode_applied(cls, data, first_index, 1)

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_appliedfield_apply", self._h, s.h)


class AppliedField(_Handle):
    """AppliedField (cpu/appliedfield/spinoperationappliedfield.cpp and its Lua methods,
    appliedfield_luafuncs.lua): a uniform field; apply sets slot 'Zeeman' to B*scale."""

    _destroy = "msd_appliedfield_destroy"
    Heterogeneous = _AppliedFieldHeterogeneous

    def __init__(self, *a):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_appliedfield_create", *self._dims, 0, C.byref(p))
        super().__init__(p.value)
        self._scale = 1.0

    new = classmethod(lambda cls, *a: cls(*a))

    # checkpoint byte stream (AppliedField::encode/decode, spinoperationappliedfield.cpp:33-58)
    def encode(self, first_index=0):
        return _encode_handle("msd_appliedfield_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        return _decode_applied(cls, data, first_index, 0)

    def _B(self):
        v = (C.c_double * 3)()
        call("msd_appliedfield_get", self._h, v)
        return [v[0], v[1], v[2]]

    def _setB(self, b):
        call("msd_appliedfield_set", self._h, float(b[0]), float(b[1]), float(b[2]))

    def x(self):
        return self._B()[0]

    def y(self):
        return self._B()[1]

    def z(self):
        return self._B()[2]

    def setX(self, v):
        b = self._B(); b[0] = v; self._setB(b)

    def setY(self, v):
        b = self._B(); b[1] = v; self._setB(b)

    def setZ(self, v):
        b = self._B(); b[2] = v; self._setB(b)

    def set(self, x=None, y=None, z=None):
        """set({x,y,z}) or set(x,y,z); missing components are 0 (appliedfield_luafuncs.lua:7-20)."""
        if isinstance(x, (tuple, list)):
            t = list(x) + [None] * 3
            x, y, z = (t[0] or 0), (t[1] or 0), (t[2] or 0)
        self._setB([x or 0, y or 0, z or 0])

    def add(self, x, y=None, z=None):
        """add({x,y,z}) or add(x,y,z) (appliedfield_luafuncs.lua:23-32)."""
        if isinstance(x, (tuple, list)):
            t = list(x) + [None] * 3
            x, y, z = (t[0] or 0), (t[1] or 0), (t[2] or 0)
        b = self._B()
        self._setB([b[0] + x, b[1] + y, b[2] + z])

    def get(self):
        return tuple(self._B())

    def toTable(self):
        return self._B()

    def setScale(self, s):
        self._scale = float(s)
        call("msd_appliedfield_set_scale", self._h, self._scale)

    def scale(self):
        return self._scale

    def slotName(self):
        return "Zeeman"

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_appliedfield_apply", self._h, s.h)


def _urandom_seed():
    """A nonzero 64-bit seed from the OS entropy pool, as an unseeded MagLua RNG takes one
    from /dev/urandom (RNG::seed, common/random/random.cpp:29-62)."""
    while True:
        v = int.from_bytes(os.urandom(8), "little")
        if v:
            return v


class Random:
    """Seed holder standing in for Random.Isaac/MersenneTwister: the build's noise is
    counter-based (Philox4x32-10), so an RNG object only carries the seed.  No seed, or
    seed 0, draws one from /dev/urandom as RNG::luaInit does (common/random/random.cpp:29-46)."""

    def __init__(self, seed=None):
        self.seed = int(seed) if seed else _urandom_seed()

    @classmethod
    def new(cls, seed=None):
        return cls(seed)


class Thermal(_Handle):
    """Thermal (cpu/thermal/spinoperationthermal.cpp): h = sqrt(2 a T s/(ms dt g)) N(0,1), slot 'Thermal'."""

    _destroy = "msd_thermal_destroy"

    def __init__(self, *a, seed=None):
        """Thermal.new(n [, rng]): the noise is keyed by the RNG's seed, else by `seed`, else
        (no seed, or 0) by a seed from /dev/urandom, so two unseeded Thermals never share
        a stream.  setSeed(seed, draw) gives a reproducible one."""
        rng = [v for v in a if isinstance(v, Random)]
        a = tuple(v for v in a if not isinstance(v, Random))
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_thermal_create", *self._dims, C.byref(p))
        super().__init__(p.value)
        self._T = 0.0
        key = rng[0].seed if rng else (int(seed) if seed else _urandom_seed())
        call("msd_thermal_set_seed", self._h, key, 0)

    new = classmethod(lambda cls, *a, **k: cls(*a, **k))

    # checkpoint byte stream (Thermal::encode/decode, spinoperationthermal.cpp:51-80); the
    # noise stream is not in it: a decoded Thermal has a fresh seed (setSeed to choose one)
    def encode(self, first_index=0):
        return _encode_handle("msd_thermal_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        o, used, k = _decode_handle(cls, "msd_thermal_decode", data, first_index)
        o._T = struct.unpack_from("<d", bytes(data), 27)[0]      # after preamble, dims, scale, version
        return o, used, k

    def set(self, T):
        self._T = float(T)
        call("msd_thermal_set_temperature", self._h, self._T)

    setTemperature = set

    def temperature(self):
        return self._T

    def setScale(self, s):
        call("msd_thermal_set_scale", self._h, float(s))

    def scaleSite(self, site, v):
        call("msd_thermal_scale_site", self._h, _index(site, self._dims), float(v))

    def setSeed(self, seed, draw=0):
        call("msd_thermal_set_seed", self._h, int(seed), int(draw))

    def draw(self):
        d = C.c_uint64()
        call("msd_thermal_get_draw", self._h, C.byref(d))
        return d.value

    def seed(self):
        d = C.c_uint64()
        call("msd_thermal_get_seed", self._h, C.byref(d))
        return d.value

    def apply(self, ss, rng=None):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_thermal_apply", self._h, s.h)


class Dipole(_Handle):
    """Dipole : LongRange (cpu/dipole, cpu/longrange): layered FFT convolution, slot 'Dipole'."""

    _destroy = "msd_dipole_destroy"
    NAMES = ("XX", "XY", "XZ", "YY", "YZ", "ZZ")

    def __init__(self, *a):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_dipole_create", *self._dims, C.byref(p))
        super().__init__(p.value)
        self._g = 1.0
        self._gmax = (1000, 1000, 1000, 1000)

    new = classmethod(lambda cls, *a: cls(*a))

    # checkpoint byte stream (Dipole::encode/decode, spinoperationdipole.cpp:35-59)
    def encode(self, first_index=0):
        return _encode_handle("msd_dipole_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        o, used, k = _decode_handle(cls, "msd_dipole_decode", data, first_index)
        *gmax, g = struct.unpack_from("<4id", bytes(data), 27)
        o._gmax, o._g = tuple(gmax), g
        return o, used, k

    def setStrength(self, g):
        self._g = float(g)
        call("msd_dipole_set_strength", self._h, self._g)

    def strength(self):
        return self._g

    def schedule(self):
        """The transform schedule in use: {"rows_per_tile", "threads", "split"} (split = rows
        per tile when the row kernels take the first radix stage of the y transform)."""
        ty, tpb, sp = C.c_int(), C.c_int(), C.c_int()
        call("msd_dipole_schedule", self._h, C.byref(ty), C.byref(tpb), C.byref(sp))
        return {"rows_per_tile": ty.value, "threads": tpb.value, "split": sp.value}

    def setScale(self, s):
        call("msd_dipole_set_scale", self._h, float(s))

    def setTruncation(self, *a):
        """l_settrunc (spinoperationlongrange.cpp:505-573): a number sets all four
        limits; a table {x,y,z} (or three numbers) sets x,y,z and radial = max.
        math.huge (float("inf")) is the reference's extrapolated sum, which ends as the
        square of images |X|, |Y| <= 4999 (dipolesupport.cpp:756-840)."""
        if len(a) == 1 and np.isscalar(a[0]):
            g = (-1,) * 4 if a[0] == float("inf") else (int(a[0]),) * 4
        else:
            v = tuple(a[0]) if len(a) == 1 else tuple(a)
            lim = [max(1, int(t)) for t in (tuple(v) + (1, 1, 1))[:3]]
            g = (max(lim), lim[0], lim[1], lim[2])
        self._gmax = g
        call("msd_dipole_set_truncation", self._h, (C.c_int * 4)(*g))

    def truncation(self):
        return tuple(float("inf") if v == -1 else v for v in self._gmax)

    def setUnitCell(self, A, B, Cc):
        m = np.ascontiguousarray(np.array([A, B, Cc], dtype=np.float64).reshape(-1))
        call("msd_dipole_set_unit_cell", self._h, _d(m))

    def _which(self, name):
        try:
            return self.NAMES.index(name.upper())
        except ValueError:
            raise MagLuaError(-1, "1st argument must be matrix name: XX, XY, XZ, YY, YZ or ZZ")

    def getMatrix(self, name, offset):
        o = tuple(int(v) for v in offset) + (0, 0, 0)
        v = C.c_double()
        call("msd_dipole_get_matrix", self._h, self._which(name), o[0], o[1], o[2], C.byref(v))
        return v.value

    def setMatrix(self, name, offset, value):
        o = tuple(int(v) for v in offset) + (0, 0, 0)
        call("msd_dipole_set_matrix", self._h, self._which(name), o[0], o[1], o[2], float(value))

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_dipole_apply", self._h, s.h)


LongRange = Dipole


class _TensorArray:
    """One tensor plane of a LongRange2D operator (the reference hands out the operator's
    own dArray): get/set/addAt take 1-based (x, y) as the reference's Array does, and
    write into the operator's copy.  As in the reference, a change reaches the field only
    after setCompileRequired(true) (or a makeData that sets it)."""

    def __init__(self, owner, ab, dest, src):
        self._o, self._k = owner, (ab, dest, src)

    def _plane(self):
        return self._o._T[self._k]

    def nx(self):
        return self._o._dims[0]

    def ny(self):
        return self._o._dims[1]

    def get(self, x, y):
        return float(self._plane()[int(y) - 1, int(x) - 1])

    def set(self, x, y, v):
        self._plane()[int(y) - 1, int(x) - 1] = float(v)
        self._o._host_dirty = True

    def addAt(self, x, y, v):
        self._plane()[int(y) - 1, int(x) - 1] += float(v)
        self._o._host_dirty = True

    def zero(self):
        self._plane()[:] = 0.0
        self._o._host_dirty = True

    def toArray(self):
        """numpy copy [ny][nx]"""
        return self._plane().copy()


class LongRange2D(_Handle):
    """LongRange2D (cpu/longrange2d/spinoperationlongrange2d.cpp): a layered in-plane FFT
    convolution with one tensor per ordered layer pair T_ab[dest][src](x, y), per-layer
    source-moment and destination-field scales and the operator scale, slot 'LongRange2D'.
    The field of layer d reads the pair stored under [s][d] for every source layer s, as
    the reference's apply does (:463-526).  Tensors are made by the function given to
    setMakeDataFunction when new data is required (Dipole2D installs its own), or set
    through tensorArray / setTensorArray; they are transformed when a compile is required.
    Both flags start true, as in the reference; a plain LongRange2D with no make-data
    function raises at apply until setNewDataRequired(false)."""

    _destroy = "msd_dipole_destroy"
    NAMES = ("XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ")
    _slot_default = "LongRange2D"

    def __init__(self, *a):
        self._dims = _dims_of(a)
        nx, ny, nz = self._dims
        p = _vp()
        call("msd_lr2d_create", nx, ny, nz, self._slot_default.encode(), C.byref(p))
        super().__init__(p.value)
        self._T = np.zeros((9, nz, nz, ny, nx))
        self._sms = np.ones(nz)
        self._dfs = np.ones(nz)
        self._scale = 1.0
        self._compile = True
        self._newdata = True
        self._host_dirty = True
        self._makefn = None
        self._internal = None

    new = classmethod(lambda cls, *a: cls(*a))

    def nx(self):
        return self._dims[0]

    def ny(self):
        return self._dims[1]

    def nz(self):
        return self._dims[2]

    def slotName(self):
        return self._slot_default

    # ---- scales (l_ssms / l_gsms / l_sdfs / l_gdfs, :533-580)
    def _layer(self, layer):
        k = int(layer) - 1
        if k < 0 or k >= self._dims[2]:
            raise MagLuaError(-1, "Invalid Layer")
        return k

    def setSourceMomentScale(self, layer, v):
        self._sms[self._layer(layer)] = float(v)
        call("msd_lr2d_set_layer_scales", self._h, _d(self._sms), None)

    def sourceMomentScale(self, layer):
        return float(self._sms[self._layer(layer)])

    def setDestinationFieldScale(self, layer, v):
        self._dfs[self._layer(layer)] = float(v)
        call("msd_lr2d_set_layer_scales", self._h, None, _d(self._dfs))

    def destinationFieldScale(self, layer):
        return float(self._dfs[self._layer(layer)])

    def setScale(self, s):
        self._scale = float(s)
        call("msd_dipole_set_scale", self._h, self._scale)

    def scale(self):
        return self._scale

    def setStrength(self, *a):
        raise MagLuaError(-3, "setStrength is depreciated. Please use setSourceMomentScale and setDestinationFieldScale.")

    def strength(self, *a):
        raise MagLuaError(-3, "strength is depreciated. Please use sourceMomentScale and sestinationFieldScale.")

    # ---- tensors (l_getTensorArray / l_setTensorArray, :585-602)
    def _ab(self, name):
        try:
            return self.NAMES.index(str(name).upper()[:2])
        except ValueError:
            raise MagLuaError(-1, "tensor name must be one of XX XY XZ YX YY YZ ZX ZY ZZ")

    def tensorArray(self, dest, src, ab):
        return _TensorArray(self, self._ab(ab), self._layer(dest), self._layer(src))

    def setTensorArray(self, dest, src, ab, a):
        """Copies a ([ny][nx] array, nx*ny values x fastest, or a tensorArray) into the plane."""
        v = a.toArray() if isinstance(a, _TensorArray) else np.asarray(a, dtype=np.float64)
        self._T[self._ab(ab), self._layer(dest), self._layer(src)] = v.reshape(self._dims[1], self._dims[0])
        self._host_dirty = True

    def getMatrix(self, dest, src, x, y, ab):
        """t.getMatrix (longrange2d_luafuncs.lua:6-9): tensorArray(dest, src, ab):get(x, y)."""
        return self.tensorArray(dest, src, ab).get(x, y)

    # ---- state flags and the make-data hook (:604-679)
    def setCompileRequired(self, b=True):
        self._compile = bool(b)

    def setNewDataRequired(self, b=True):
        self._newdata = bool(b)

    def setInternalData(self, v):
        self._internal = v

    def internalData(self):
        return self._internal

    def setMakeDataFunction(self, f):
        if not callable(f):
            raise MagLuaError(-1, "function required as argument")
        self._makefn = f

    def makeDataFunction(self):
        return self._makefn

    def makeData(self):
        """LongRange2D::makeNewData (:377-394)."""
        if not self._newdata:
            return
        self._newdata = False
        if self._makefn is None:
            raise MagLuaError(-1, "make data function not set")
        self._makefn(self)
        self._compile = True

    def _pull(self):
        """Copy the library's tensors (after a builder ran there) into this object's copy."""
        nz = self._dims[2]
        buf = np.empty(self._dims[0] * self._dims[1])
        for ab in range(9):
            for dd in range(nz):
                for ss in range(nz):
                    call("msd_lr2d_get_tensor", self._h, dd, ss, ab, _d(buf))
                    self._T[ab, dd, ss] = buf.reshape(self._dims[1], self._dims[0])
        self._host_dirty = False

    def _sync(self):
        """What apply does first: new data, then the compile (the library transforms the
        tensors it holds at the next field evaluation)."""
        if self._newdata:
            self.makeData()
        if self._compile:
            if self._host_dirty:
                nz = self._dims[2]
                for ab in range(9):
                    for dd in range(nz):
                        for ss in range(nz):
                            call("msd_lr2d_set_tensor", self._h, dd, ss, ab,
                                 _d(np.ascontiguousarray(self._T[ab, dd, ss]).reshape(-1)))
                self._host_dirty = False
            self._compile = False

    def apply(self, ss):
        self._sync()
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_dipole_apply", self._h, s.h)


class Dipole2D(LongRange2D):
    """Dipole2D (cpu/dipole2d): LongRange2D whose tensors are the truncated dipole box sums of
    makeData (dip2d_luafuncs.lua:559-668), slot 'Dipole2D'.  Internal data: per-layer unit
    cells (identity) and offsets (0), truncation 25 in every direction.  physical=False
    reproduces the reference's tensors, which evaluate source and destination at the same
    in-plane cell, so only the layers' vertical coupling enters; physical=True places the
    destination at the origin (the truncated in-plane dipole sum).  As in the reference,
    changing the cell, offsets or truncation after the tensors are made takes
    setNewDataRequired() to rebuild them."""

    _slot_default = "Dipole2D"

    def __init__(self, *a, physical=False):
        super().__init__(*a)
        self._physical = bool(physical)
        self.setMakeDataFunction(Dipole2D._make)
        self.setInternalData(None)

    def setPhysical(self, yes=True):
        self._physical = bool(yes)

    def _id(self):
        """initializeInternalData (:6-26)"""
        if self._internal is None:
            nz = self._dims[2]
            self._internal = {"ABC": [[[1, 0, 0], [0, 1, 0], [0, 0, 1]] for _ in range(nz)],
                              "offset": [[0, 0, 0] for _ in range(nz)],
                              "truncation": 25, "truncationX": 25, "truncationY": 25, "truncationZ": 25}
        return self._internal

    def _check_layer(self, layer):
        if not isinstance(layer, (int, float, np.integer, np.floating)):
            raise MagLuaError(-1, "Layer expected as first argument")
        if layer < 1 or layer > self._dims[2]:
            raise MagLuaError(-1, "Layer must be in the range [1:%d]" % self._dims[2])
        return int(layer) - 1

    def setUnitCell(self, layer, A, B=None, Cv=None):
        k = self._check_layer(layer)
        if B is None:
            return self.setUnitCell(layer, A[0], A[1], A[2])
        try:
            cell = [[float(v) for v in A], [float(v) for v in B], [float(v) for v in Cv]]
        except TypeError:
            raise MagLuaError(-1, "Tables expected for the A, B, C components (2nd through 4th arguments)")
        self._id()["ABC"][k] = cell

    def unitCell(self, layer):
        k = self._check_layer(layer)
        A, B, Cv = self._id()["ABC"][k]
        return list(A), list(B), list(Cv)

    def setOffset(self, layer, v):
        k = self._check_layer(layer)
        try:
            self._id()["offset"][k] = [float(x) for x in v]
        except TypeError:
            raise MagLuaError(-1, "Table expected for 2nd argument")

    def offset(self, layer=None):
        t = self._id()
        if layer is None:
            return [list(o) for o in t["offset"]]
        return list(t["offset"][self._check_layer(layer)])

    def setTruncation(self, trunc=None, truncX=None, truncY=None, truncZ=None):
        t = self._id()
        for k, v in (("truncation", trunc), ("truncationX", truncX), ("truncationY", truncY), ("truncationZ", truncZ)):
            if v is not None:
                t[k] = int(v)

    def truncation(self):
        t = self._id()
        return t["truncation"], t["truncationX"], t["truncationY"], t["truncationZ"]

    def _make(self):
        t = self._id()
        nz = self._dims[2]
        ABC = np.ascontiguousarray(np.asarray(t["ABC"], dtype=np.float64).reshape(nz * 9))
        off = np.ascontiguousarray(np.asarray(t["offset"], dtype=np.float64).reshape(nz * 3))
        tr = (C.c_int * 4)(t["truncation"], t["truncationX"], t["truncationY"], t["truncationZ"])
        call("msd_lr2d_build_dipole2d", self._h, _d(ABC), _d(off), tr, int(self._physical))
        self._pull()
        self.setCompileRequired(True)


class Magnetostatics2D(LongRange2D):
    """Magnetostatics2D (cpu/magnetostatics2d): LongRange2D whose tensors are the
    prism-prism demagnetising tensors of makeData (mag2d_luafuncs.lua:598-750), slot
    'Magnetostatics2D'.  Internal data (:7-23): per-layer unit cells (identity) and prism
    (grain) sizes (1, 1, 1), truncation 25 radially and along x, y and z.  -N(r; source
    grain, destination grain) is summed over the images of each cell and stored under the
    labels [source][destination].  As in the reference, the scales are the user's: 1/V of
    the source moments and 4 pi on the destination fields in CGS, through
    setSourceMomentScale / setDestinationFieldScale.  Changing the cells, grains or
    truncation after the tensors are made takes setNewDataRequired() to rebuild them."""

    _slot_default = "Magnetostatics2D"

    def __init__(self, *a):
        super().__init__(*a)
        self.setMakeDataFunction(Magnetostatics2D._make)
        self.setInternalData(None)

    def _id(self):
        """initializeInternalData (:7-23)"""
        if self._internal is None:
            nz = self._dims[2]
            self._internal = {"ABC": [[[1, 0, 0], [0, 1, 0], [0, 0, 1]] for _ in range(nz)],
                              "grainSize": [[1.0, 1.0, 1.0] for _ in range(nz)],
                              "truncation": 25, "truncationX": 25, "truncationY": 25, "truncationZ": 25}
        return self._internal

    def _check_layer(self, layer, what="first argument"):
        if not isinstance(layer, (int, float, np.integer, np.floating)):
            raise MagLuaError(-1, "Layer expected as " + what)
        if layer < 1 or layer > self._dims[2]:
            raise MagLuaError(-1, "Layer must be in the range [1:%d]" % self._dims[2])
        return int(layer) - 1

    def setUnitCell(self, layer, A, B=None, Cv=None):
        """setUnitCell (:64-78)"""
        k = self._check_layer(layer)
        try:
            cell = [[float(v) for v in A], [float(v) for v in B], [float(v) for v in Cv]]
        except TypeError:
            raise MagLuaError(-1, "Tables expected for the A, B, C components (2nd through 4th arguments)")
        self._id()["ABC"][k] = cell

    def unitCell(self, layer):
        k = self._check_layer(layer)
        A, B, Cv = self._id()["ABC"][k]
        return list(A), list(B), list(Cv)

    def setGrainSize(self, layer, X, Y, Z):
        """setGrainSize (:93-107)"""
        k = self._check_layer(layer)
        if not all(isinstance(v, (int, float, np.integer, np.floating)) for v in (X, Y, Z)):
            raise MagLuaError(-1, "Numbers expected for the X, Y, Z components (2nd through 4th arguments)")
        self._id()["grainSize"][k] = [float(X), float(Y), float(Z)]

    def grainSize(self, layer):
        """grainSize (:109-122): x, y, z and the prism volume"""
        k = self._check_layer(layer, "argument")
        x, y, z = self._id()["grainSize"][k]
        return x, y, z, x * y * z

    def setTruncation(self, trunc=None, truncX=None, truncY=None, truncZ=None):
        """setTruncation (:124-132): a nil argument keeps that limit"""
        t = self._id()
        for k, v in (("truncation", trunc), ("truncationX", truncX), ("truncationY", truncY), ("truncationZ", truncZ)):
            if v is not None:
                t[k] = int(v)

    def truncation(self):
        t = self._id()
        return t["truncation"], t["truncationX"], t["truncationY"], t["truncationZ"]

    def getMatrix(self, dest, src, ab, x, y):
        """getMatrix (:32-45): makes the data if needed, then tensorArray(dest, src, ab):get(x, y)
        with 0-based offsets."""
        self.makeData()
        return self.tensorArray(dest, src, ab).get(x + 1, y + 1)

    def setMatrix(self, dest, src, ab, x, y, v):
        """setMatrix (:47-62)"""
        self.makeData()
        self.tensorArray(dest, src, ab).set(x + 1, y + 1, v)
        self.setCompileRequired(True)

    def _make(self):
        t = self._id()
        nz = self._dims[2]
        ABC = np.ascontiguousarray(np.asarray(t["ABC"], dtype=np.float64).reshape(nz * 9))
        grain = np.ascontiguousarray(np.asarray(t["grainSize"], dtype=np.float64).reshape(nz * 3))
        tr = (C.c_int * 4)(t["truncation"], t["truncationX"], t["truncationY"], t["truncationZ"])
        call("msd_lr2d_build_magnetostatics2d", self._h, _d(ABC), _d(grain), tr)
        self._pull()
        self.setCompileRequired(True)


class LongRange3D(_Handle):
    """LongRange3D (cpu/longrange3d/spinoperationlongrange3d.cpp): 3D FFT convolution, slot
    'LongRange3D'.  open=False: periodic, dipole lattice sum to radius R (setRadius, default
    1000); open=True: open boundaries, zero padding to the doubled grid."""

    _destroy = "msd_longrange3d_destroy"
    NAMES = {"XX": 0, "XY": 1, "XZ": 2, "YX": 1, "YY": 3, "YZ": 4, "ZX": 2, "ZY": 4, "ZZ": 5}

    def __init__(self, *a, open=False):
        self._dims = _dims_of(a)
        p = _vp()
        call("msd_longrange3d_create", *self._dims, int(bool(open)), C.byref(p))
        super().__init__(p.value)
        self._g = 1.0
        self._open = bool(open)

    new = classmethod(lambda cls, *a, **k: cls(*a, **k))

    def setStrength(self, g):
        self._g = float(g)
        call("msd_longrange3d_set_strength", self._h, self._g)

    def strength(self):
        return self._g

    def setScale(self, s):
        call("msd_longrange3d_set_scale", self._h, float(s))

    def setRadius(self, R):
        call("msd_longrange3d_set_radius", self._h, int(R))

    def _which(self, name):
        try:
            return self.NAMES[name.upper()]
        except KeyError:
            raise MagLuaError(-1, "1st argument must be a tensor name: XX, XY, ... ZZ")

    def getMatrix(self, name, offset):
        o = tuple(int(v) for v in offset) + (0, 0, 0)
        v = C.c_double()
        call("msd_longrange3d_get_matrix", self._h, self._which(name), o[0], o[1], o[2], C.byref(v))
        return v.value

    def setMatrix(self, name, offset, value):
        o = tuple(int(v) for v in offset) + (0, 0, 0)
        call("msd_longrange3d_set_matrix", self._h, self._which(name), o[0], o[1], o[2], float(value))

    def apply(self, ss):
        for s in (ss if isinstance(ss, (list, tuple)) else (ss,)):
            call("msd_longrange3d_apply", self._h, s.h)

    def setPassWidths(self, y=0, z=0):
        """kx columns per workgroup of the y pass and the z pass (0 keeps a width); only
        compiled widths of the compile-time pass forms are accepted."""
        call("msd_longrange3d_set_pass_widths", self._h, int(y), int(z))

    def passWidths(self):
        """{"y", "z"}: the widths in use; {"y_choices", "z_choices"}: the widths each pass
        form offers, default first (empty: the runtime-length pass, width fixed)."""
        wy, wz, ny, nz = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        yc, zc = (C.c_int * 8)(), (C.c_int * 8)()
        call("msd_longrange3d_pass_widths", self._h, C.byref(wy), C.byref(wz), yc, C.byref(ny), zc, C.byref(nz))
        return {"y": wy.value, "z": wz.value, "y_choices": list(yc[:ny.value]), "z_choices": list(zc[:nz.value])}


class _LLGBase(_Handle):
    """LLG.Cartesian (cpu/llg/llgcartesian.cpp:90-278) with LLG::apply argument rules (llg.cpp:100-171)."""

    _destroy = "msd_llg_destroy"
    _create = "msd_llg_create"

    def __init__(self):
        p = _vp()
        call(self._create, C.byref(p))
        super().__init__(p.value)
        self._tf = True
        self._nr = False

    def setThermalOnlyFirstTerm(self, yes=True):
        self._tf = bool(yes)
        call("msd_llg_set_thermal_only_first_term", self._h, int(self._tf))

    def thermalOnlyFirstTerm(self):
        return self._tf

    def setDisableRenormalization(self, yes=True):
        self._nr = bool(yes)
        call("msd_llg_set_disable_renormalization", self._h, int(self._nr))

    def setScales(self, precess=1.0, damp=1.0):
        call("msd_llg_set_scales", self._h, float(precess), float(damp))

    # checkpoint byte stream (LLG::encode/decode, cpu/llg/llg.cpp:50-88)
    def encode(self, first_index=0):
        return _encode_handle("msd_llg_encode", self._h, first_index)

    @classmethod
    def decode(cls, data, first_index=0):
        data = bytes(data)
        p, used, k = _vp(), C.c_int64(), C.c_int()
        call("msd_llg_decode", data, len(data), int(first_index), int(cls._create == "msd_llg_create_quaternion"),
             C.byref(p), C.byref(used), C.byref(k))
        o = cls.__new__(cls)
        _Handle.__init__(o, p.value)
        tf, nr = struct.unpack_from("<ii", data, 6)               # after preamble and version
        o._tf, o._nr = bool(tf), bool(nr)
        return o, used.value, k.value

    def apply(self, *args):
        """apply(from [,scale] [,dmdt] [,to] [,advance]); lists apply element-wise."""
        args = list(args)
        advance = True
        if args and isinstance(args[-1], bool):
            advance = args.pop()
        if not args:
            raise MagLuaError(-1, "apply requires 1, 2 or 3 SpinSystems")
        frm = args.pop(0)
        scale = 1.0
        if args and isinstance(args[0], (int, float)) and not isinstance(args[0], bool):
            scale = float(args.pop(0))
        dmdt = args.pop(0) if args else frm
        to = args.pop(0) if args else frm
        F = frm if isinstance(frm, (list, tuple)) else [frm]
        D = dmdt if isinstance(dmdt, (list, tuple)) else [dmdt]
        T = to if isinstance(to, (list, tuple)) else [to]
        if not (len(F) == len(D) == len(T)):
            raise MagLuaError(-4, "Size mismatch in number of SpinSystems given in each table")
        for f, d, t in zip(F, D, T):
            call("msd_llg_apply", self._h, f.h, scale, d.h, t.h, int(advance))


class _LLGNamespace:
    class Cartesian(_LLGBase):
        @classmethod
        def new(cls):
            return cls()

    class Quaternion(_LLGBase):
        """LLG.Quaternion (cpu/llg/llgquat.cpp:127-373): each apply rotates the spin about
        the damped field by g|H'|dt/(1+a^2); same argument rules and flags as Cartesian."""
        _create = "msd_llg_create_quaternion"

        @classmethod
        def new(cls):
            return cls()

    @staticmethod
    def new(kind="Cartesian"):
        if kind == "Cartesian":
            return _LLGNamespace.Cartesian()
        if kind == "Quaternion":
            return _LLGNamespace.Quaternion()
        raise MagLuaError(-3, f"LLG.{kind} is not provided by this build (LLG.Cartesian and LLG.Quaternion are)")


LLG = _LLGNamespace


# ------------------------------------------------------------------ drivers
def make_pc_step_function(s, calcField, llg, tol, optional_temp=None):
    """common/scripts/PredictorCorrector.lua:54-109, over the per-operator GPU kernels."""
    state = {"P": s.copy(), "C": s.copy()}

    def step(ss=None, skip_temperature=False):
        ss = ss or s
        P, Cc = state["P"], state["C"]
        last = [-1.0]

        def same(a, b):
            d = a.diff(b)[3]
            if d == last[0]:
                raise MagLuaError(-1, "Failing to converge in Predictor Corrector, timestep too large?")
            last[0] = d
            return d < tol

        calcField(ss)
        llg.apply(ss, ss, P)
        q = 0
        while True:
            llg.apply(ss, 0.5, ss, Cc)
            calcField(P)
            llg.apply(Cc, 0.5, P, Cc)
            P, Cc = Cc, P
            q += 1
            if q > 100:
                raise MagLuaError(-1, "Failing to converge in Predictor Corrector, timestep too large?")
            if same(P, Cc):
                break
        state["P"], state["C"] = P, Cc
        P.copySpinsTo(ss)
        ss.setTime(P.time())
        if optional_temp is not None and not skip_temperature:
            ss.resetFields()
            optional_temp.apply(ss)
            ss.sumFields()
            llg.apply(ss, False)
        return True

    return step


def make_adapt_step_function(*a):
    """common/scripts/AdaptiveTimeStep.lua: wrap a step function in step doubling.

    Arguments are sorted by kind as the Lua does (userdata, functions, numbers): the
    SpinSystem, then the LLG operator and an optional thermal operator; the step
    function, then an optional dynamics function; the tolerance.  Each call takes one
    full step and two half steps from the same state; when their max-norm difference
    (SpinSystem.diff) is below the tolerance the two half steps are kept (plus the
    thermal kick) and the next step is longer by 1 + clip(ee^2, 0.1, 0.5) - 0.1,
    ee = (tol - diff)/tol; otherwise the state is kept and the step halves.  Returns
    whether the step was accepted."""
    objs = [v for v in a if v is not None and not callable(v) and not isinstance(v, (int, float))]
    funcs = [v for v in a if callable(v) and not isinstance(v, _Handle)]
    nums = [v for v in a if isinstance(v, (int, float)) and not isinstance(v, bool)]
    ss, step, tol = objs[0], funcs[0], float(nums[0])
    dynamics = funcs[1] if len(funcs) > 1 else (lambda s: None)
    llg = objs[1] if len(objs) > 1 else None
    temp = objs[2] if len(objs) > 2 else None
    ss2, ss3 = ss.copy(), ss.copy()

    def same(A, Bs):
        d = A.diff(Bs)[3]
        return d < tol, (tol - d) / tol

    def adapt_step(s=None, skip_temperature=False):
        s = s or ss
        ss2.setTime(s.time())
        ss3.setTime(s.time())
        ss2.setTimeStep(s.timeStep())
        ss3.setTimeStep(s.timeStep() / 2)
        s.copySpinsTo(ss2)
        s.copySpinsTo(ss3)
        dynamics(ss2)
        step(ss2, True)                  # one full step, no thermal kick
        dynamics(ss3)
        step(ss3, True)                  # two half steps
        dynamics(ss3)
        step(ss3, True)
        ok, ee = same(ss2, ss3)
        if ok:
            ss3.copySpinsTo(s)
            if temp is not None and not skip_temperature:
                dynamics(s)
                s.resetFields()
                temp.apply(s)
                s.sumFields()
                llg.apply(s, False)      # False: the time is not advanced
            s.setTime(ss3.time())
            ee = min(max(ee * ee, 0.1), 0.5)
            s.setTimeStep(s.timeStep() * (1 + ee - 0.1))
        else:
            s.setTimeStep(s.timeStep() * 0.5)
        return ok

    return adapt_step


BUTCHER = {
    "RK1": [[1]],
    "RK2": [[1 / 2], [0, 1]],
    "K3": [[1 / 2], [-1, 2], [1 / 6, 2 / 3, 1 / 6]],
    "RK4": [[1 / 2], [0, 1 / 2], [0, 0, 1], [1 / 6, 1 / 3, 1 / 3, 1 / 6]],
    "RK3": [[1 / 3], [0, 2 / 3], [1 / 4, 0, 3 / 4]],
    "RK4_38": [[1 / 3], [-1 / 3, 1], [1, -1, 1], [1 / 8, 3 / 8, 3 / 8, 1 / 8]],
    "RK6": [[1 / 3], [0, 2 / 3], [1 / 12, 4 / 12, -1 / 12], [-1 / 16, 18 / 16, -3 / 16, -6 / 16],
            [0, 9 / 8, -3 / 8, -6 / 8, 4 / 8], [9 / 44, -36 / 44, 63 / 44, 72 / 44, 0, -64 / 44],
            [11 / 120, 0, 81 / 120, 81 / 120, -32 / 120, -32 / 120, 11 / 120]],
    "BS3": [[1 / 2], [0, 3 / 4], [2 / 9, 1 / 3, 4 / 9], [7 / 24, 1 / 4, 1 / 3, 1 / 8]],
}


def make_rk_step_function(ss, kind, calcField, llg, optional_temp=None, dynamics=None):
    """common/scripts/RungeKutta.lua:57-214: Butcher rows as chained renormalised Euler sub-steps."""
    if isinstance(calcField, (list, tuple)):
        ops = list(calcField)

        def calcField(s):
            s.resetFields()
            for o in ops:
                o.apply(s)
            s.sumFields()
    if kind not in BUTCHER:
        raise MagLuaError(-1, 'RK type must be one of "' + '", "'.join(BUTCHER) + '"')
    bt = BUTCHER[kind]
    nstep = len(bt)
    k = [ss] + [ss.copy() for _ in range(nstep - 1)]
    dyn = dynamics or (lambda s: None)

    def istep(src, f, dst):
        llg.apply(src, f[0], src, dst)
        for i in range(1, nstep):
            if i < len(f) and f[i] != 0:
                llg.apply(dst, f[i], k[i], dst)

    def sfunc(ss_input=None, skip_temperature=False):
        k[0] = ss_input or ss
        dyn(k[0])
        calcField(k[0])
        for i in range(nstep - 1):
            k[i + 1].setTimeStep(k[0].timeStep())
            istep(k[0], bt[i], k[i + 1])
            dyn(k[i + 1])
            calcField(k[i + 1])
        istep(k[0], bt[nstep - 1], k[0])
        if optional_temp is not None and not skip_temperature:
            k[0].resetFields()
            optional_temp.apply(k[0])
            k[0].sumFields()
            llg.apply(k[0], False)
        return True

    return sfunc


class Heun(_Handle):
    """The fused device form of make_pc_step_function with a single corrector pass
    (Heun) over Exchange + Anisotropy + Dipole, with the optional thermal kick."""

    _destroy = "msd_stepper_destroy"

    def __init__(self, ss, ex=None, ani=None, dip=None, temp=None, llg=None, graph=True, compensated=True):
        lr3 = dip if isinstance(dip, LongRange3D) else None
        if lr3 is not None:
            dip = None
        self._keep = (ss, ex, ani, dip, temp, llg, lr3)
        g = lambda o: o.h if o is not None else None
        p = _vp()
        call("msd_stepper_create", ss.h, g(ex), g(ani), g(dip), g(temp), g(llg), C.byref(p))
        super().__init__(p.value)
        call("msd_stepper_set_graph", self._h, int(bool(graph)))
        call("msd_stepper_set_compensated", self._h, int(bool(compensated)))
        if lr3 is not None:
            call("msd_stepper_set_longrange3d", self._h, lr3.h)

    def _sync(self):
        d = self._keep[3]
        if isinstance(d, LongRange2D):
            d._sync()

    def step(self, n=1):
        self._sync()
        call("msd_stepper_heun", self._h, int(n))
        return True

    __call__ = step

    def setUniaxial(self, u):
        """Anisotropy.Uniaxial in the step's field sum (None removes it)."""
        self._uni = u
        call("msd_stepper_set_uniaxial", self._h, u.h if u is not None else None)

    def addAppliedField(self, af):
        """Add an AppliedField or AppliedField.Heterogeneous to the step's field sum (None
        removes them all); later changes to it apply from the next step() call."""
        if af is None:
            self._af = []
            call("msd_stepper_add_applied_field", self._h, None)
            return
        self._af = getattr(self, "_af", []) + [af]
        call("msd_stepper_add_applied_field", self._h, af.h)

    def setSurrogate(self, sur, K=10, tol=1e-3):
        """Learned dipole field: windows of K steps from an exact reference field.  Each window
        is checked when it ends (relative max-norm residual of the surrogate's field against
        the exact one <= tol); a window over tol is restored and run again as exact steps.
        tol=0 makes every window exact.  sur=None returns to exact steps."""
        self._sur = sur
        call("msd_stepper_set_surrogate", self._h, sur.h if sur is not None else None, int(K), float(tol))

    def autotune(self, steps=20):
        """Layered Dipole driver: time every compiled row-tile shape in both y-transform
        schedules (split: the row kernels take a radix-rows_per_tile stage; full columns)
        for `steps` steps and keep the fastest; returns {"rows_per_tile", "threads",
        "split", "ms_per_step"}.  LongRange3D driver: time every pair of compiled y-pass
        and z-pass kx widths; returns {"y_width", "z_width", "ms_per_step"}.  The state is
        restored either way."""
        if self._keep[6] is not None:
            wy, wz, ms = C.c_int(), C.c_int(), C.c_double()
            call("msd_stepper_autotune_passes", self._h, int(steps), C.byref(wy), C.byref(wz), C.byref(ms))
            return {"y_width": wy.value, "z_width": wz.value, "ms_per_step": ms.value}
        ty, tpb, sp, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        call("msd_stepper_autotune", self._h, int(steps), C.byref(ty), C.byref(tpb), C.byref(sp), C.byref(ms))
        return {"rows_per_tile": ty.value, "threads": tpb.value, "split": sp.value, "ms_per_step": ms.value}

    def rk(self, kind, n=1):
        """n steps of make_rk_step_function(ss, kind, ...) with the fused stage kernels
        (Butcher table `kind` of RungeKutta.lua; plain FP32 state)."""
        if kind not in BUTCHER:
            raise MagLuaError(-1, 'RK type must be one of "' + '", "'.join(BUTCHER) + '"')
        self._sync()
        call("msd_stepper_rk", self._h, kind.encode(), int(n))
        return True

    def graphBuilds(self):
        """Step graphs captured so far (values that change between steps reuse the graph)."""
        n = C.c_long()
        call("msd_stepper_graph_builds", self._h, C.byref(n))
        return n.value

    def surrogateStats(self):
        w, e, f = C.c_long(), C.c_long(), C.c_long()
        last, mx = C.c_double(), C.c_double()
        call("msd_stepper_surrogate_stats", self._h, C.byref(w), C.byref(e), C.byref(f), C.byref(last), C.byref(mx))
        return {"windows": w.value, "exact_windows": e.value, "fallbacks": f.value,
                "last_residual": last.value, "max_residual": mx.value}

    def profile(self, n):
        """Kernel -> average ms per launch (HIP events on the library stream).  Layered and
        local forms repeat each launch of one step n times back to back (one step of
        progress); the LongRange3D form runs n more steps with events around every launch."""
        cap = 8
        names = (C.c_char * 32 * cap)()
        avg = (C.c_double * cap)()
        cnt = C.c_int()
        call("msd_stepper_profile", self._h, int(n), cap, names, avg, C.byref(cnt))
        return {bytes(names[i]).split(b"\0")[0].decode(): avg[i] for i in range(cnt.value)}


class Surrogate(_Handle):
    """Learned predictor of the dipole field's change between exact refreshes (the
    layered Dipole or the 3D LongRange3D field):
    dH = s_out * L3(L2(L1(s_in * (m - m_ref)))), three 5x5x3 convolution layers with
    3 -> 16 -> 16 -> 3 channels (BF16 MFMA, FP32 accumulate), periodic in x and y
    (periodic=False: zero-padded there, for lattices with open boundaries),
    zero-padded in z.  Weights of layer l are arrays [3,5,5,cout,cin] indexed
    [dz+1, dy+2, dx+2, out, in].  Field arrays are [3][n] like SpinSystem.spins()."""

    _destroy = "msd_surrogate_destroy"
    CHANNELS = (3, 16, 16, 3)

    def __init__(self, nx, ny=None, nz=None, periodic=True):
        if ny is None:
            nx, ny, nz = _dims_of((nx,))
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz
        self.periodic = bool(periodic)
        p = _vp()
        call("msd_surrogate_create", self.nx, self.ny, self.nz, C.byref(p))
        super().__init__(p.value)
        if not self.periodic:
            call("msd_surrogate_set_open", self._h, 1)

    def _shape(self, l):
        return (3, 5, 5, self.CHANNELS[l + 1], self.CHANNELS[l])

    def setWeights(self, layer, W):
        a = np.ascontiguousarray(np.asarray(W, dtype=np.float64).reshape(self._shape(layer)))
        call("msd_surrogate_set_weights", self._h, int(layer), _d(a))

    def weights(self, layer):
        a = np.zeros(self._shape(layer))
        call("msd_surrogate_get_weights", self._h, int(layer), _d(a))
        return a

    def setScales(self, s_in, s_out, sigma2=1.0):
        call("msd_surrogate_set_scales", self._h, float(s_in), float(s_out), float(sigma2))

    def scales(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        call("msd_surrogate_get_scales", self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def _field(self, v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3, self.n))
        return a

    def setInput(self, dm):
        a = self._field(dm)
        call("msd_surrogate_set_input", self._h, _d(a))

    def setInputSystems(self, m, m_ref):
        call("msd_surrogate_set_input_systems", self._h, m.h, m_ref.h)

    def setTarget(self, t):
        a = self._field(t)
        call("msd_surrogate_set_target", self._h, _d(a))

    def setTargetDipole(self, dip, m, m_ref):
        """Target = the exact field of m - m_ref from dip, a Dipole or a LongRange3D."""
        fn = "msd_surrogate_set_target_longrange3d" if isinstance(dip, LongRange3D) else "msd_surrogate_set_target_dipole"
        call(fn, self._h, dip.h, m.h, m_ref.h)

    def forward(self, download=True):
        """dH of the current input as a [3][n] array (download=False: leave it on the device)."""
        if not download:
            call("msd_surrogate_forward", self._h, None)
            return None
        out = np.zeros((3, self.n))
        call("msd_surrogate_forward", self._h, _d(out))
        return out

    def lossGrads(self):
        loss = C.c_double()
        call("msd_surrogate_loss_grads", self._h, C.byref(loss))
        grads = []
        for l in range(3):
            g = np.zeros(self._shape(l))
            call("msd_surrogate_get_grads", self._h, l, _d(g))
            grads.append(g)
        return loss.value, grads

    def adam(self, lr):
        call("msd_surrogate_adam", self._h, float(lr))

    def resetOptimizer(self):
        call("msd_surrogate_reset_optimizer", self._h)

    def trainStep(self, lr):
        loss = C.c_double()
        call("msd_surrogate_train_step", self._h, float(lr), C.byref(loss))
        return loss.value

    def fieldError(self, dip, m, m_ref):
        """(max|H_ref + dH - H| / max|H|, the same for H_ref alone) with the exact fields of
        dip, a Dipole or a LongRange3D."""
        a, b = C.c_double(), C.c_double()
        fn = "msd_surrogate_field_error_longrange3d" if isinstance(dip, LongRange3D) else "msd_surrogate_field_error"
        call(fn, self._h, dip.h, m.h, m_ref.h, C.byref(a), C.byref(b))
        return a.value, b.value


class Ensemble(_Handle):
    """Independent replicas (uneven sizes, own temperatures and seeds) advanced by one set
    of launches per Heun pass.  members: Heun step drivers of the local, layered (Dipole,
    LongRange2D, Dipole2D) or LongRange3D form."""

    _destroy = "msd_ensemble_destroy"

    def __init__(self, members, graph=True, compensated=True):
        self._members = list(members)
        for m in self._members:
            m._sync()
        arr = (C.c_void_p * len(self._members))(*[m.h for m in self._members])
        p = _vp()
        call("msd_ensemble_create", arr, len(self._members), C.byref(p))
        super().__init__(p.value)
        call("msd_ensemble_set_graph", self._h, int(bool(graph)))
        call("msd_ensemble_set_compensated", self._h, int(bool(compensated)))

    def step(self, n=1):
        for m in self._members:
            m._sync()
        call("msd_ensemble_heun", self._h, int(n))
        return True

    __call__ = step

    def profile(self, n):
        """n more steps with HIP events around every pass group of a step; returns pass
        group -> average ms per launch (k3_ypass_fwd, k3_zconv, k3_ypass_inv,
        k3_rows_pred, k3_rows_corr, films_pred, films_corr)."""
        for m in self._members:
            m._sync()
        cap = 8
        names = (C.c_char * 32 * cap)()
        avg = (C.c_double * cap)()
        cnt = C.c_int()
        call("msd_ensemble_profile", self._h, int(n), cap, names, avg, C.byref(cnt))
        return {bytes(names[i]).split(b"\0")[0].decode(): avg[i] for i in range(cnt.value)}

    def autotune(self, steps=10):
        """LongRange3D pass widths timed in the ensemble's own launches: members grouped by
        lattice, every compiled (y, z) width pair tried on each group in turn; the members'
        states are restored.  Returns {"ms_per_step", "widths": {"NXxNYxNZ[o]": (y, z)}}
        (o: open boundaries)."""
        for m in self._members:
            m._sync()
        ms = C.c_double()
        call("msd_ensemble_autotune_passes", self._h, int(steps), C.byref(ms))
        widths = {}
        for m in self._members:
            lr = m._keep[6]
            if lr is not None:
                pw = lr.passWidths()
                key = "x".join(str(v) for v in lr._dims) + ("o" if lr._open else "")
                widths[key] = (pw["y"], pw["z"])
        return {"ms_per_step": ms.value, "widths": widths}


def synchronize():
    call("msd_synchronize")


def philox_device(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros_like(ctr)
    k = (C.c_uint32 * 2)(*key)
    call("msd_philox_device", ctr.ctypes.data_as(C.POINTER(C.c_uint32)), ctr.shape[0], k,
         out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out
