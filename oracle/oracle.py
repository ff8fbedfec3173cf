"""ctypes view of the FP64 CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.  The classes mirror the MagLua objects
whose algorithms msd_oracle.cpp restates (SpinSystem, Exchange, Anisotropy,
Thermal, Dipole, LLG.Cartesian); coordinates here are 0-based.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(threads=False):
    name = "liboracle_omp.so" if threads else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "ora_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "ora_thermal_normals": (None, [C.c_uint64, C.c_uint64, C.c_int64, _dp]),
        "ora_random_unit_spins": (None, [C.c_uint64, C.c_int64, _dp, _dp, _dp]),
        "ora_fft1d": (C.c_int, [_dp, C.c_int, C.c_int]),
        "ora_dft1d_naive": (None, [_dp, _dp, C.c_int, C.c_int]),
        "ora_fft2d_layers": (C.c_int, [_dp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ora_dipole_tensor": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), _dp,
                                        _dp, _dp, _dp, _dp, _dp, _dp]),
        "ora_sys_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_sys_free": (None, [vp]),
        "ora_sys_copy": (vp, [vp]),
        "ora_sys_copy_spins_to": (None, [vp, vp]),
        "ora_sys_set_spins": (None, [vp, _dp, _dp, _dp]),
        "ora_sys_set_spin": (None, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
        "ora_sys_get_spins": (None, [vp, _dp, _dp, _dp, _dp]),
        "ora_sys_set_params": (None, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
        "ora_sys_get_params": (None, [vp, _dp]),
        "ora_sys_set_site_alpha": (None, [vp, vp]),
        "ora_sys_set_site_gamma": (None, [vp, vp]),
        "ora_sys_slot": (C.c_int, [vp, C.c_char_p]),
        "ora_sys_get_field": (C.c_int, [vp, C.c_char_p, _dp, _dp, _dp]),
        "ora_sys_set_field": (C.c_int, [vp, C.c_char_p, _dp, _dp, _dp]),
        "ora_ex_add_paths": (None, [vp, _ip, _ip, _dp, C.c_int]),
        "ora_ani_add_many": (None, [vp, _ip, _dp, _dp, C.c_int]),
        "ora_sys_reset_fields": (None, [vp]),
        "ora_sys_sum_fields": (None, [vp]),
        "ora_sys_net_moment": (None, [vp, _dp]),
        "ora_sys_field_dot": (C.c_double, [vp, C.c_char_p]),
        "ora_ex_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_ex_free": (None, [vp]),
        "ora_ex_set_pbc": (None, [vp, C.c_int, C.c_int, C.c_int]),
        "ora_ex_set_normalize": (None, [vp, C.c_int]),
        "ora_ex_set_scale": (None, [vp, C.c_double]),
        "ora_ex_add": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
        "ora_ex_add_path": (None, [vp, C.c_int, C.c_int, C.c_double]),
        "ora_ex_apply": (C.c_int, [vp, vp]),
        "ora_ani_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_ani_free": (None, [vp]),
        "ora_ani_set_scale": (None, [vp, C.c_double]),
        "ora_ani_add": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
        "ora_ani_apply": (C.c_int, [vp, vp]),
        "ora_uni_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_uni_free": (None, [vp]),
        "ora_uni_set_scale": (None, [vp, C.c_double]),
        "ora_uni_add": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
        "ora_uni_apply": (C.c_int, [vp, vp]),
        "ora_sim_add_uni": (None, [vp, vp]),
        "ora_af_new": (vp, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "ora_af_free": (None, [vp]),
        "ora_af_set": (None, [vp, C.c_double, C.c_double, C.c_double]),
        "ora_af_set_scale": (None, [vp, C.c_double]),
        "ora_af_set_site": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double]),
        "ora_af_apply": (C.c_int, [vp, vp]),
        "ora_sim_add_af": (None, [vp, vp]),
        "ora_th_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_th_free": (None, [vp]),
        "ora_th_set_temperature": (None, [vp, C.c_double]),
        "ora_th_set_scale": (None, [vp, C.c_double]),
        "ora_th_scale_site": (None, [vp, C.c_int, C.c_double]),
        "ora_th_apply": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_uint64]),
        "ora_lr_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_lr_free": (None, [vp]),
        "ora_lr_set_strength": (None, [vp, C.c_double]),
        "ora_lr_set_scale": (None, [vp, C.c_double]),
        "ora_lr_set_truncation": (None, [vp, C.POINTER(C.c_int)]),
        "ora_lr_set_unit_cell": (None, [vp, _dp]),
        "ora_lr_set_tensor": (C.c_int, [vp, C.c_int, _dp]),
        "ora_lr_get_tensor": (C.c_int, [vp, C.c_int, _dp]),
        "ora_lr_apply": (C.c_int, [vp, vp]),
        "ora_fft3d": (C.c_int, [_dp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ora_dipole_tensor3d_periodic": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "ora_dipole_tensor3d_open": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "ora_dipole_direct_open": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _ip, C.c_double, _dp]),
        "ora_lr3_new": (vp, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "ora_lr3_free": (None, [vp]),
        "ora_lr3_set_strength": (None, [vp, C.c_double]),
        "ora_lr3_set_tensor": (C.c_int, [vp, C.c_int, _dp]),
        "ora_lr3_apply": (C.c_int, [vp, vp]),
        "ora_sim_set_lr3": (None, [vp, vp]),
        "ora_lr2d_new": (vp, [C.c_int, C.c_int, C.c_int]),
        "ora_lr2d_free": (None, [vp]),
        "ora_lr2d_set_scale": (None, [vp, C.c_double]),
        "ora_lr2d_set_layer_scales": (None, [vp, _dp, _dp]),
        "ora_lr2d_set_tensor": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _dp]),
        "ora_lr2d_get_tensor": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _dp]),
        "ora_lr2d_apply": (C.c_int, [vp, vp, C.c_char_p]),
        "ora_dipole2d_tensor": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_int), C.c_int, _dp]),
        "ora_mag2d_element": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]),
        "ora_mag2d_tensor": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_int), _dp]),
        "ora_sim_set_lr2d": (None, [vp, vp, C.c_char_p]),
        "ora_llg_new": (vp, []),
        "ora_llg_free": (None, [vp]),
        "ora_llg_set_flags": (None, [vp, C.c_int, C.c_int]),
        "ora_llg_set_quaternion": (None, [vp, C.c_int]),
        "ora_llg_apply": (C.c_int, [vp, vp, C.c_double, vp, vp, C.c_int]),
        "ora_sim_new": (vp, [vp, vp, vp, vp, vp, vp]),
        "ora_sim_free": (None, [vp]),
        "ora_sim_set_seed": (None, [vp, C.c_uint64]),
        "ora_sim_calc_fields": (C.c_int, [vp, vp]),
        "ora_sim_heun": (C.c_int, [vp, C.c_int]),
        "ora_sim_rk": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "ora_sim_set_threads": (None, [C.c_int]),
    }
    for k, (res, args) in sig.items():
        f = getattr(L, k)
        f.restype = res
        f.argtypes = args
    _libs[name] = L
    return L


# ---------------------------------------------------------------- helpers
def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.ora_philox4x32_10(c, k, o)
    return [o[i] for i in range(4)]


def thermal_normals(seed, step, site):
    out = np.zeros(3)
    lib().ora_thermal_normals(seed, step, site, out)
    return out


def random_unit_spins(seed, n):
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().ora_random_unit_spins(seed, n, x, y, z)
    return x, y, z


def fft1d(a, sign=-1):
    d = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64).copy()
    lib().ora_fft1d(d, len(a), sign)
    return d.view(np.complex128)


def dft_naive(a, sign=-1):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    out = np.zeros_like(a)
    lib().ora_dft1d_naive(a.view(np.float64), out.view(np.float64), len(a), sign)
    return out


def fft2d_layers(a, sign=-1):
    """a: complex array shaped (nz, ny, nx); per-layer 2D transform."""
    nz, ny, nx = a.shape
    d = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().ora_fft2d_layers(d.view(np.float64).reshape(-1), nx, ny, nz, sign)
    return d


def dipole_tensor(nx, ny, nz, gmax=(1000, 1000, 1000, 1000), ABC=None, threads=True):
    """Returns dict XX..ZZ of arrays shaped (2nz-1, ny, nx); layer c holds z = c-nz+1."""
    L = lib(threads)
    m = nx * ny * (2 * nz - 1)
    outs = [np.zeros(m) for _ in range(6)]
    g = (C.c_int * 4)(*gmax)
    A = np.ascontiguousarray(np.eye(3).reshape(-1) if ABC is None else np.asarray(ABC, dtype=np.float64).reshape(-1))
    rc = L.ora_dipole_tensor(nx, ny, nz, g, A, *outs)
    assert rc == 0
    return {k: o.reshape(2 * nz - 1, ny, nx) for k, o in zip(("XX", "XY", "XZ", "YY", "YZ", "ZZ"), outs)}


def dipole2d_tensor(nx, ny, nz, ABC=None, offset=None, trunc=(25, 25, 25, 25), physical=False, threads=True):
    """Dipole2D makeData: array [9, nz, nz, ny, nx] (XX..ZZ, labels [dest][src]).
    ABC: [nz][3][3] cell vectors per layer (identity), offset: [nz][3] (zeros)."""
    L = lib(threads)
    A = np.ascontiguousarray(np.tile(np.eye(3), (nz, 1, 1)) if ABC is None else np.asarray(ABC, dtype=np.float64)).reshape(-1)
    off = np.ascontiguousarray(np.zeros((nz, 3)) if offset is None else np.asarray(offset, dtype=np.float64)).reshape(-1)
    out = np.zeros(9 * nz * nz * ny * nx)
    rc = L.ora_dipole2d_tensor(nx, ny, nz, A, off, (C.c_int * 4)(*trunc), int(physical), out)
    assert rc == 0
    return out.reshape(9, nz, nz, ny, nx)


MAG2D_NAMES = ("XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ")


def mag2d_element(ab, r, d1, d2):
    """Magnetostatics2D.N<ab>(rx, ry, rz, source prism d1, destination prism d2)."""
    k = MAG2D_NAMES.index(ab) if isinstance(ab, str) else int(ab)
    return lib().ora_mag2d_element(k, float(r[0]), float(r[1]), float(r[2]),
                                   np.ascontiguousarray(d1, dtype=np.float64), np.ascontiguousarray(d2, dtype=np.float64))


def mag2d_tensor(nx, ny, nz, ABC=None, grain=None, trunc=(25, 25, 25, 25), threads=True):
    """Magnetostatics2D makeData: array [9, nz, nz, ny, nx] (XX..ZZ, labels [source][dest]).
    ABC: [nz][3][3] cell vectors per layer (identity), grain: [nz][3] prism sizes (ones)."""
    L = lib(threads)
    A = np.ascontiguousarray(np.tile(np.eye(3), (nz, 1, 1)) if ABC is None else np.asarray(ABC, dtype=np.float64)).reshape(-1)
    gr = np.ascontiguousarray(np.ones((nz, 3)) if grain is None else np.asarray(grain, dtype=np.float64)).reshape(-1)
    out = np.zeros(9 * nz * nz * ny * nx)
    rc = L.ora_mag2d_tensor(nx, ny, nz, A, gr, (C.c_int * 4)(*trunc), out)
    assert rc == 0
    return out.reshape(9, nz, nz, ny, nx)


# ---------------------------------------------------------------- objects
class SpinSystem:
    def __init__(self, nx, ny=1, nz=1, threads=False, _h=None):
        self.L = lib(threads)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        self.h = _h if _h is not None else self.L.ora_sys_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_sys_free(self.h)
            self.h = None

    def copy(self):
        s = SpinSystem(self.nx, self.ny, self.nz, _h=self.L.ora_sys_copy(self.h))
        s.L = self.L
        return s

    def setSpins(self, x, y, z):
        self.L.ora_sys_set_spins(self.h, *(np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (x, y, z)))

    def setSpin(self, site, s, m=-1.0):
        self.L.ora_sys_set_spin(self.h, site, s[0], s[1], s[2], m)

    def spins(self):
        o = [np.zeros(self.n) for _ in range(4)]
        self.L.ora_sys_get_spins(self.h, *o)
        return o

    def setParams(self, alpha=None, gamma=None, dt=None, time=None):
        p = self.params()
        self.L.ora_sys_set_params(self.h, p[0] if alpha is None else alpha, p[1] if gamma is None else gamma,
                                  p[2] if dt is None else dt, p[3] if time is None else time)

    def params(self):
        o = np.zeros(4)
        self.L.ora_sys_get_params(self.h, o)
        return o

    def setSiteAlpha(self, a):
        self._sa = None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self.L.ora_sys_set_site_alpha(self.h, None if a is None else self._sa.ctypes.data)

    def setSiteGamma(self, g):
        self._sg = None if g is None else np.ascontiguousarray(g, dtype=np.float64)
        self.L.ora_sys_set_site_gamma(self.h, None if g is None else self._sg.ctypes.data)

    def time(self):
        return self.params()[3]

    def setAlpha(self, a):
        self.setParams(alpha=a)

    def setGamma(self, g):
        self.setParams(gamma=g)

    def setTimeStep(self, dt):
        self.setParams(dt=dt)

    def setTime(self, t):
        self.setParams(time=t)

    def timeStep(self):
        return self.params()[2]

    def field(self, name):
        o = [np.zeros(self.n) for _ in range(3)]
        rc = self.L.ora_sys_get_field(self.h, name.encode(), *o)
        if rc != 0:
            raise KeyError(name)
        return o

    def setField(self, name, hx, hy, hz):
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (hx, hy, hz)]
        self.L.ora_sys_set_field(self.h, name.encode(), *a)

    def resetFields(self):
        self.L.ora_sys_reset_fields(self.h)

    def sumFields(self):
        self.L.ora_sys_sum_fields(self.h)

    def netMoment(self):
        o = np.zeros(8)
        self.L.ora_sys_net_moment(self.h, o)
        return o

    def fieldDot(self, name):
        return self.L.ora_sys_field_dot(self.h, name.encode())


class Exchange:
    def __init__(self, nx, ny=1, nz=1, threads=False):
        self.L = lib(threads)
        self.h = self.L.ora_ex_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_ex_free(self.h)

    def setPeriodicXYZ(self, px, py, pz):
        self.L.ora_ex_set_pbc(self.h, int(px), int(py), int(pz))

    def normalizeMoments(self, yes):
        self.L.ora_ex_set_normalize(self.h, int(yes))

    def setScale(self, s):
        self.L.ora_ex_set_scale(self.h, s)

    def add(self, a, b, J=1.0):
        return self.L.ora_ex_add(self.h, a[0], a[1], a[2], b[0], b[1], b[2], J)

    def addPath(self, f, t, J):
        self.L.ora_ex_add_path(self.h, f, t, J)

    def addPaths(self, f, t, J):
        f = np.ascontiguousarray(f, dtype=np.int32)
        t = np.ascontiguousarray(t, dtype=np.int32)
        j = np.ascontiguousarray(np.broadcast_to(np.asarray(J, dtype=np.float64), f.shape))
        self.L.ora_ex_add_paths(self.h, f, t, j, int(f.size))

    def apply(self, ss):
        self.L.ora_ex_apply(self.h, ss.h)


class Anisotropy:
    def __init__(self, nx, ny=1, nz=1, threads=False):
        self.L = lib(threads)
        self.h = self.L.ora_ani_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_ani_free(self.h)

    def setScale(self, s):
        self.L.ora_ani_set_scale(self.h, s)

    def add(self, site, axis, K):
        return self.L.ora_ani_add(self.h, site, axis[0], axis[1], axis[2], K)

    def addMany(self, sites, axes, K):
        s = np.ascontiguousarray(sites, dtype=np.int32)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(axes, dtype=np.float64), (s.size, 3)))
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(K, dtype=np.float64), s.shape))
        self.L.ora_ani_add_many(self.h, s, a, k, int(s.size))

    def apply(self, ss):
        self.L.ora_ani_apply(self.h, ss.h)


class AnisotropyUniaxial:
    """Anisotropy.Uniaxial: per-op site, axis (normalised), K1, K2; slot "UniaxialAnisotropy"."""

    def __init__(self, nx, ny=1, nz=1, threads=False):
        self.L = lib(threads)
        self.h = self.L.ora_uni_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_uni_free(self.h)

    def setScale(self, s):
        self.L.ora_uni_set_scale(self.h, float(s))

    def add(self, site0, axis, K1, K2=0.0):
        rc = self.L.ora_uni_add(self.h, int(site0), *[float(v) for v in axis], float(K1), float(K2))
        assert rc >= 0
        return rc == 1

    def apply(self, ss):
        self.L.ora_uni_apply(self.h, ss.h)


class AppliedField:
    """Uniform applied field (slot "Zeeman", overwritten by apply) or, with
    heterogeneous=True, per-site fields (slot "Zeeman_Heterogeneous", accumulated)."""

    def __init__(self, nx, ny=1, nz=1, heterogeneous=False, threads=False):
        self.L = lib(threads)
        self.h = self.L.ora_af_new(nx, ny, nz, int(heterogeneous))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_af_free(self.h)

    def set(self, x, y=0.0, z=0.0):
        self.L.ora_af_set(self.h, float(x), float(y), float(z))

    def setScale(self, s):
        self.L.ora_af_set_scale(self.h, float(s))

    def setSite(self, site0, x, y, z):
        assert self.L.ora_af_set_site(self.h, int(site0), float(x), float(y), float(z)) == 0

    def apply(self, ss):
        self.L.ora_af_apply(self.h, ss.h)


class Thermal:
    def __init__(self, nx, ny=1, nz=1, threads=False):
        self.L = lib(threads)
        self.h = self.L.ora_th_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_th_free(self.h)

    def setTemperature(self, T):
        self.L.ora_th_set_temperature(self.h, T)

    def setScale(self, s):
        self.L.ora_th_set_scale(self.h, s)

    def scaleSite(self, site, v):
        self.L.ora_th_scale_site(self.h, site, v)

    def apply(self, ss, normals=None, seed=0, step=0):
        if normals is not None:
            self._n = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1)
            self.L.ora_th_apply(self.h, ss.h, self._n.ctypes.data, seed, step)
        else:
            self.L.ora_th_apply(self.h, ss.h, None, seed, step)


class Dipole:
    NAMES = ("XX", "XY", "XZ", "YY", "YZ", "ZZ")

    def __init__(self, nx, ny=1, nz=1, threads=False):
        self.L = lib(threads)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = self.L.ora_lr_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_lr_free(self.h)

    def setStrength(self, g):
        self.L.ora_lr_set_strength(self.h, g)

    def setScale(self, s):
        self.L.ora_lr_set_scale(self.h, s)

    def setTruncation(self, gmax):
        if np.isscalar(gmax):
            gmax = (-1,) * 4 if gmax == float("inf") else (int(gmax),) * 4
        self.L.ora_lr_set_truncation(self.h, (C.c_int * 4)(*gmax))

    def setUnitCell(self, A, B, Cc):
        self.L.ora_lr_set_unit_cell(self.h, np.ascontiguousarray(np.array([A, B, Cc], dtype=np.float64).reshape(-1)))

    def tensor(self, name):
        t = np.zeros(self.nx * self.ny * (2 * self.nz - 1))
        self.L.ora_lr_get_tensor(self.h, self.NAMES.index(name), t)
        return t.reshape(2 * self.nz - 1, self.ny, self.nx)

    def setTensor(self, name, t):
        self.L.ora_lr_set_tensor(self.h, self.NAMES.index(name), np.ascontiguousarray(t, dtype=np.float64).reshape(-1))

    def apply(self, ss):
        self.L.ora_lr_apply(self.h, ss.h)


def fft3d(a, sign=-1):
    """a: complex (nz, ny, nx); full 3D transform, unnormalised."""
    nz, ny, nx = a.shape
    d = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().ora_fft3d(d.view(np.float64).reshape(-1), nx, ny, nz, sign)
    return d


def dipole_tensor3d(nx, ny, nz, R=None, open=False, threads=True):
    """Six arrays (XX XY XZ YY YZ ZZ): periodic lattice sums to radius R shaped (nz, ny, nx),
    or the open-boundary kernel on the doubled grid shaped (2nz, 2ny, 2nx)."""
    L = lib(threads)
    if open:
        m = 8 * nx * ny * nz
        outs = [np.zeros(m) for _ in range(6)]
        assert L.ora_dipole_tensor3d_open(nx, ny, nz, *outs) == 0
        shp = (2 * nz, 2 * ny, 2 * nx)
    else:
        m = nx * ny * nz
        outs = [np.zeros(m) for _ in range(6)]
        assert L.ora_dipole_tensor3d_periodic(nx, ny, nz, int(R), *outs) == 0
        shp = (nz, ny, nx)
    return {k: o.reshape(shp) for k, o in zip(("XX", "XY", "XZ", "YY", "YZ", "ZZ"), outs)}


def dipole_direct_open(dims, x, y, z, sites, g, threads=True):
    """Open-boundary dipole field at the site indices `sites`, shaped (k, 3), by the direct
    real-space sum (no FFT)."""
    nx, ny, nz = dims
    sites = np.ascontiguousarray(sites, dtype=np.int32)
    out = np.zeros(3 * sites.size)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
    assert lib(threads).ora_dipole_direct_open(nx, ny, nz, *a, int(sites.size), sites, float(g), out) == 0
    return out.reshape(-1, 3)


class LongRange3D:
    """Nine-tensor 3D FFT convolution; open=True zero-pads to the doubled grid."""
    NAMES = ("XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ")

    def __init__(self, nx, ny=1, nz=1, open=False, threads=False):
        self.L = lib(threads)
        self.h = self.L.ora_lr3_new(nx, ny, nz, int(open))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_lr3_free(self.h)

    def setStrength(self, g):
        self.L.ora_lr3_set_strength(self.h, g)

    def setTensor(self, name, t):
        self.L.ora_lr3_set_tensor(self.h, self.NAMES.index(name), np.ascontiguousarray(t, dtype=np.float64).reshape(-1))

    def setDipole(self, T6):
        for a, b in (("XX", "XX"), ("XY", "XY"), ("XZ", "XZ"), ("YX", "XY"), ("YY", "YY"), ("YZ", "YZ"),
                     ("ZX", "XZ"), ("ZY", "YZ"), ("ZZ", "ZZ")):
            self.setTensor(a, T6[b])

    def apply(self, ss):
        self.L.ora_lr3_apply(self.h, ss.h)


class LongRange2D:
    """A tensor per ordered layer pair (labels [dest][src], read as [src][dest] by apply),
    per-layer source-moment and destination-field scales, an operator scale."""
    NAMES = ("XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ")

    def __init__(self, nx, ny=1, nz=1, slot="LongRange2D", threads=False):
        self.L = lib(threads)
        self.nx, self.ny, self.nz, self.slot = nx, ny, nz, slot
        self.h = self.L.ora_lr2d_new(nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_lr2d_free(self.h)

    def setScale(self, s):
        self.L.ora_lr2d_set_scale(self.h, s)

    def setLayerScales(self, src_moment, dest_field):
        self.L.ora_lr2d_set_layer_scales(self.h, np.ascontiguousarray(src_moment, dtype=np.float64),
                                         np.ascontiguousarray(dest_field, dtype=np.float64))

    def setTensors(self, T):
        """T: [9, nz, nz, ny, nx]"""
        T = np.asarray(T, dtype=np.float64)
        for ab in range(9):
            for d in range(self.nz):
                for s_ in range(self.nz):
                    assert self.L.ora_lr2d_set_tensor(self.h, d, s_, ab, np.ascontiguousarray(T[ab, d, s_]).reshape(-1)) == 0

    def apply(self, ss):
        self.L.ora_lr2d_apply(self.h, ss.h, self.slot.encode())


class LLG:
    def __init__(self, threads=False, quaternion=False):
        self.L = lib(threads)
        self.h = self.L.ora_llg_new()
        if quaternion:
            self.L.ora_llg_set_quaternion(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_llg_free(self.h)

    def setFlags(self, thermal_only_first_term=True, disable_renormalization=False):
        self.L.ora_llg_set_flags(self.h, int(thermal_only_first_term), int(disable_renormalization))

    def apply(self, frm, scale=1.0, dmdt=None, to=None, advance=True):
        dmdt = frm if dmdt is None else dmdt
        to = frm if to is None else to
        self.L.ora_llg_apply(self.h, frm.h, scale, dmdt.h, to.h, int(advance))


class Sim:
    """A fixed operator set driven by the PredictorCorrector / RungeKutta step."""

    def __init__(self, ss, ex=None, ani=None, dip=None, th=None, llg=None, seed=0):
        self.L = ss.L
        self.keep = (ss, ex, ani, dip, th, llg)
        self.llg = llg if llg is not None else LLG()
        g = lambda o: o.h if o is not None else None
        self.h = self.L.ora_sim_new(ss.h, g(ex), g(ani), g(dip), g(th), self.llg.h)
        self.L.ora_sim_set_seed(self.h, seed)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_sim_free(self.h)

    def addUniaxial(self, u):
        self.keep_uni = getattr(self, "keep_uni", []) + [u]
        self.L.ora_sim_add_uni(self.h, u.h)

    def addAppliedField(self, af):
        self.keep_af = getattr(self, "keep_af", []) + [af]
        self.L.ora_sim_add_af(self.h, af.h)

    def setLongRange3D(self, lr3):
        self.keep3 = lr3
        self.L.ora_sim_set_lr3(self.h, lr3.h if lr3 is not None else None)

    def setLongRange2D(self, lr2d):
        self.keep2 = lr2d
        self.L.ora_sim_set_lr2d(self.h, lr2d.h if lr2d is not None else None, (lr2d.slot if lr2d else "").encode())

    def calcFields(self, ss):
        self.L.ora_sim_calc_fields(self.h, ss.h)

    def heun(self, n=1):
        self.L.ora_sim_heun(self.h, n)

    def rk(self, name, n=1):
        rc = self.L.ora_sim_rk(self.h, name.encode(), n)
        if rc != 0:
            raise ValueError(name)

    @staticmethod
    def set_threads(n):
        lib(True).ora_sim_set_threads(n)
