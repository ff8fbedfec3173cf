"""Synthetic seeded workloads on the MagLua object API.

`build(api, spec)` assembles SpinSystem + Exchange + Anisotropy + Dipole
(+ Thermal) the way a MagLua script does (doc/manual/BasicMicromagneticSimulation.lua):
nearest-neighbour exchange paths added from every site to each neighbour,
uniaxial anisotropy on every site, the layered dipole, seeded random unit
spins.  `api` is any module exposing the MagLua class names with the bulk
helpers addPaths/addMany, so the same spec drives this package and any other
implementation of the API.
"""
from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class FilmSpec:
    nx: int
    ny: int
    nz: int
    J: float = 1.0
    K: float = 0.05
    axis: tuple = (0.0, 0.0, 1.0)
    g: float = 0.02
    pbc: tuple = (1, 1, 0)
    alpha: float = 0.1
    gamma: float = 1.0
    dt: float = 0.002
    T: float = 0.0
    seed: int = 0           # initial spins
    noise_seed: int = 0     # key of the counter-based thermal noise stream

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def name(self):
        t = f", T={self.T}" if self.T else ""
        return (f"{self.nx}x{self.ny}x{self.nz} film, J={self.J}, K={self.K}, g={self.g}, "
                f"alpha={self.alpha}, dt={self.dt}{t}")


# 32x32x1 periodic square lattice, FP64 CPU case
FILM_32 = FilmSpec(32, 32, 1, J=1.0, K=0.1, g=0.01, pbc=(1, 1, 0), alpha=0.1, dt=0.005)
# 256x256x8 thin film, periodic in-plane, open in z
FILM_256_T0 = FilmSpec(256, 256, 8, J=1.0, K=0.05, g=0.02, alpha=0.1, dt=0.002, T=0.0)
FILM_256_T05 = replace(FILM_256_T0, T=0.5)
# 512x512x4 film with strong dipolar coupling
FILM_512 = FilmSpec(512, 512, 4, J=1.0, K=0.02, g=0.1, alpha=0.1, dt=0.002, T=0.2)


def nn_paths(nx, ny, nz, pbc):
    """(from, to) of ex:add({site}, {site +- e_d}, J) over sites in index order and
    neighbours +x,-x,+y,-y,+z,-z; paths leaving a non-periodic side are dropped
    (l_addpath, cpu/exchange/spinoperationexchange.cpp:396-411); a dimension of
    size 1 contributes no paths."""
    site = np.arange(nx * ny * nz, dtype=np.int64)
    c0 = [site % nx, (site // nx) % ny, site // (nx * ny)]
    dims = (nx, ny, nz)
    nbs, oks = [], []
    for ax in range(3):
        if dims[ax] == 1:
            continue
        for s in (+1, -1):
            c = [v.copy() for v in c0]
            c[ax] = c[ax] + s
            if pbc[ax]:
                c[ax] %= dims[ax]
                ok = np.ones(site.size, dtype=bool)
            else:
                ok = (c[ax] >= 0) & (c[ax] < dims[ax])
                c[ax] = np.clip(c[ax], 0, dims[ax] - 1)
            nbs.append(c[0] + nx * (c[1] + ny * c[2]))
            oks.append(ok)
    if not nbs:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    nb = np.stack(nbs, axis=1)            # (n, ndir): site-major, direction order inside
    ok = np.stack(oks, axis=1)
    frm = np.broadcast_to(site[:, None], nb.shape)
    return frm[ok].astype(np.int32), nb[ok].astype(np.int32)


def build(api, spec, with_dipole=True, with_thermal=None, threads=None):
    """Returns dict(ss, ex, ani, dip, temp, llg) for `spec` on the classes of `api`."""
    kw = {} if threads is None else {"threads": threads}
    ss = api.SpinSystem(spec.nx, spec.ny, spec.nz, **kw)
    ss.setAlpha(spec.alpha)
    ss.setGamma(spec.gamma)
    ss.setTimeStep(spec.dt)
    ex = api.Exchange(spec.nx, spec.ny, spec.nz, **kw)
    ex.setPeriodicXYZ(*spec.pbc)
    f, t = nn_paths(spec.nx, spec.ny, spec.nz, spec.pbc)
    ex.addPaths(f, t, spec.J)
    ani = api.Anisotropy(spec.nx, spec.ny, spec.nz, **kw)
    ani.addMany(np.arange(spec.n, dtype=np.int32), np.asarray(spec.axis, dtype=np.float64), spec.K)
    dip = None
    if with_dipole and spec.g != 0:
        dip = api.Dipole(spec.nx, spec.ny, spec.nz, **kw)
        dip.setStrength(spec.g)
    temp = None
    if with_thermal if with_thermal is not None else spec.T != 0:
        temp = api.Thermal(spec.nx, spec.ny, spec.nz, **kw)
        temp.setTemperature(spec.T)
        if hasattr(temp, "setSeed"):        # a reproducible stream (an unseeded Thermal draws its key)
            temp.setSeed(spec.noise_seed)
    llg = api.LLG.Cartesian() if hasattr(api.LLG, "Cartesian") else api.LLG(**kw)
    return dict(ss=ss, ex=ex, ani=ani, dip=dip, temp=temp, llg=llg)
