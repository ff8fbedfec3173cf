"""LongRange2D and Dipole2D on the GPU against the FP64 oracle.

LongRange2D (cpu/longrange2d/spinoperationlongrange2d.cpp) keeps a full tensor per ordered
layer pair and reads the pair stored under [src][dest] for destination dest
(:463-526); Dipole2D (cpu/dipole2d/dip2d_luafuncs.lua:537-668) fills those tensors
with truncated dipole box sums.  The oracle's apply is pinned by a direct real-space
convolution and its Dipole2D tensors by closed forms and the layered lattice sum
(tests/test_oracle.py).  FP32 on the GPU: fields agree to a few 1e-7 relative
(tolerances sit ~10x above the measured values printed by each test)."""
import ctypes

import numpy as np
import pytest

import maglua_amd as M
from maglua_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def flat(t):
    return np.stack([np.asarray(v).reshape(-1) for v in t])


def pair(nx, ny, nz, T, sms, dfs, scale, seed, cls=M.LongRange2D, oslot="LongRange2D"):
    g = cls(nx, ny, nz)
    g.setNewDataRequired(False)
    for ab, name in enumerate(M.LongRange2D.NAMES):
        for d in range(nz):
            for s in range(nz):
                g.setTensorArray(d + 1, s + 1, name, T[ab, d, s])
    for k in range(nz):
        g.setSourceMomentScale(k + 1, sms[k])
        g.setDestinationFieldScale(k + 1, dfs[k])
    g.setScale(scale)
    o = O.LongRange2D(nx, ny, nz, slot=oslot, threads=True)
    o.setTensors(T)
    o.setLayerScales(sms, dfs)
    o.setScale(scale)
    x, y, z = O.random_unit_spins(seed, nx * ny * nz)
    gs, os_ = M.SpinSystem(nx, ny, nz), O.SpinSystem(nx, ny, nz, threads=True)
    gs.setSpins(x, y, z)
    os_.setSpins(x, y, z)
    return g, o, gs, os_


@pytest.mark.parametrize("dims", [(16, 8, 3), (32, 64, 2), (64, 32, 8), (40, 30, 3), (32, 32, 1), (256, 256, 4)])
def test_lr2d_random_tensors_against_oracle(dims):
    """Random tensors (no symmetry: T[s][d] != T[d][s]), random layer and operator scales.
    Covers compile-time column lengths 32..256 and the runtime-shaped kernels (8, 30)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx + 7 * ny + 31 * nz)
    T = rng.standard_normal((9, nz, nz, ny, nx)) / np.sqrt(nx * ny)
    sms, dfs = rng.uniform(0.5, 2.0, nz), rng.uniform(0.5, 2.0, nz)
    g, o, gs, os_ = pair(nx, ny, nz, T, sms, dfs, 0.3, 5)
    g.apply(gs)
    o.apply(os_)
    hg, ho = flat(gs.field("LongRange2D")), flat(os_.field("LongRange2D"))
    err = float(np.abs(hg - ho).max() / np.abs(ho).max())
    print(f"LongRange2D {nx}x{ny}x{nz} random tensors: max rel err = {err:.3e}", flush=True)
    assert err <= 5e-6


def test_lr2d_flags_follow_the_reference():
    """No make-data function: apply raises until setNewDataRequired(false); a tensorArray
    edit reaches the field only after setCompileRequired(true)."""
    nx, ny, nz = 16, 16, 2
    lr = M.LongRange2D(nx, ny, nz)
    ss = M.SpinSystem(nx, ny, nz)
    ss.setSpins(*O.random_unit_spins(1, nx * ny * nz))
    with pytest.raises(M.MagLuaError, match="make data function not set"):
        lr.apply(ss)
    lr.setNewDataRequired(False)
    lr.apply(ss)
    assert np.abs(flat(ss.field("LongRange2D"))).max() == 0.0
    lr.tensorArray(1, 1, "XX").set(1, 1, 1.0)          # self term: H_x = m_x on layer 1
    lr.apply(ss)
    assert np.abs(flat(ss.field("LongRange2D"))).max() == 0.0   # not compiled yet
    lr.setCompileRequired(True)
    lr.apply(ss)
    h = flat(ss.field("LongRange2D"))
    x = np.asarray(ss.spins()[0]).reshape(-1)
    n1 = nx * ny
    assert np.allclose(h[0, :n1], x[:n1], atol=1e-6) and np.abs(h[0, n1:]).max() <= 1e-6
    assert np.abs(h[1:]).max() <= 1e-6
    assert lr.getMatrix(1, 1, 1, 1, "XX") == 1.0
    with pytest.raises(M.MagLuaError, match="deprecated|depreciated"):
        lr.setStrength(1.0)
    with pytest.raises(M.MagLuaError, match="Invalid Layer"):
        lr.setSourceMomentScale(3, 1.0)


@pytest.mark.parametrize("physical", [False, True])
def test_dipole2d_against_oracle(physical):
    """Dipole2D with its defaults (truncation 25) and with a lifted third layer and an offset
    second layer; the tensors and the field against the oracle's makeData and apply."""
    nx, ny, nz = 32, 32, 3
    d = M.Dipole2D(nx, ny, nz, physical=physical)
    d.setUnitCell(2, [1, 0, 0], [0, 1, 0], [0.1, 0.0, 1.5])
    d.setOffset(2, [0.3, -0.2, 0.1])
    d.setTruncation(25, 20, 12, 2)
    ss = M.SpinSystem(nx, ny, nz)
    x, y, z = O.random_unit_spins(9, nx * ny * nz)
    ss.setSpins(x, y, z)
    d.apply(ss)
    ABC = np.tile(np.eye(3), (nz, 1, 1))
    ABC[1, 2] = (0.1, 0.0, 1.5)
    off = np.zeros((nz, 3))
    off[1] = (0.3, -0.2, 0.1)
    T = O.dipole2d_tensor(nx, ny, nz, ABC, off, (25, 20, 12, 2), physical=physical)
    got = np.array([[[d.tensorArray(a + 1, b + 1, n).toArray() for b in range(nz)] for a in range(nz)]
                    for n in M.LongRange2D.NAMES])
    assert np.abs(got - T).max() <= 1e-12 * np.abs(T).max()
    assert d.getMatrix(1, 2, 1, 1, "ZZ") == T[8, 0, 1, 0, 0]
    o = O.LongRange2D(nx, ny, nz, slot="Dipole2D", threads=True)
    o.setTensors(T)
    os_ = O.SpinSystem(nx, ny, nz, threads=True)
    os_.setSpins(x, y, z)
    o.apply(os_)
    hg, ho = flat(ss.field("Dipole2D")), flat(os_.field("Dipole2D"))
    err = float(np.abs(hg - ho).max() / np.abs(ho).max())
    print(f"Dipole2D physical={physical} {nx}x{ny}x{nz}: max rel err = {err:.3e}", flush=True)
    assert err <= 5e-6


def test_dipole2d_physical_single_layer_equals_layered_dipole():
    """nz = 1: the physical Dipole2D box sum is the layered Dipole's lattice sum with the
    same in-plane limits (no radial cut); both run on the GPU."""
    nx, ny = 32, 32
    d2 = M.Dipole2D(nx, ny, 1, physical=True)
    d2.setTruncation(0, 40, 40, 0)
    dp = M.Dipole(nx, ny, 1)
    assert M.lib().msd_dipole_set_truncation(dp.h, (ctypes.c_int * 4)(10 ** 6, 0, 40, 40)) == 0
    ss = M.SpinSystem(nx, ny, 1)
    ss.setSpins(*O.random_unit_spins(4, nx * ny))
    d2.apply(ss)
    dp.apply(ss)
    a, b = flat(ss.field("Dipole2D")), flat(ss.field("Dipole"))
    err = float(np.abs(a - b).max() / np.abs(b).max())
    print(f"Dipole2D physical vs layered Dipole 32x32x1: max rel err = {err:.3e}", flush=True)
    assert err <= 5e-6


def test_heun_with_dipole2d_against_oracle():
    """The fused Heun driver with a Dipole2D operator (physical form, truncation 10, scale
    0.05) on a 32x32x2 film: 30 steps against the oracle's Heun with the same tensors."""
    spec = W.FilmSpec(32, 32, 2, J=1.0, K=0.05, g=0.0, alpha=0.1, dt=0.005, T=0.0)
    g = W.build(M, spec, with_dipole=False)
    o = W.build(O, spec, with_dipole=False, threads=True)
    x, y, z = O.random_unit_spins(3, spec.n)
    g["ss"].setSpins(x, y, z)
    o["ss"].setSpins(x, y, z)
    d = M.Dipole2D(spec.nx, spec.ny, spec.nz, physical=True)
    d.setTruncation(10, 10, 10, 10)
    d.setScale(0.05)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], d, None, g["llg"])
    heun.step(30)
    od = O.LongRange2D(spec.nx, spec.ny, spec.nz, slot="Dipole2D", threads=True)
    od.setTensors(O.dipole2d_tensor(spec.nx, spec.ny, spec.nz, trunc=(10, 10, 10, 10), physical=True))
    od.setScale(0.05)
    sim = O.Sim(o["ss"], o["ex"], o["ani"], None, None, o["llg"])
    sim.setLongRange2D(od)
    sim.heun(30)
    a = flat(g["ss"].spins()[:3])
    b = flat(o["ss"].spins()[:3])
    err = float(np.abs(a - b).max())
    print(f"Heun + Dipole2D 32x32x2, 30 steps: max|dm| = {err:.3e}", flush=True)
    assert err <= 2e-6
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-12


def test_dipole2d_films_in_one_ensemble_match_solo_steps():
    """Two Dipole2D films and a layered-Dipole film in one Ensemble: the per-pair films
    share one column launch and every member ends bit-identical to its solo run."""
    def film(seed, kind):
        spec = W.FilmSpec(32, 32, 2, J=1.0, K=0.05, g=0.02, alpha=0.1, dt=0.005, T=0.0)
        g = W.build(M, spec, with_dipole=(kind == "dipole"))
        g["ss"].setSpins(*O.random_unit_spins(seed, spec.n))
        d = g["dip"]
        if kind != "dipole":
            d = M.Dipole2D(spec.nx, spec.ny, spec.nz, physical=(kind == "physical"))
            d.setTruncation(8, 8, 8, 8)
            d.setScale(0.03)
        return g, d, M.Heun(g["ss"], g["ex"], g["ani"], d, None, g["llg"])
    kinds = ["physical", "reference", "dipole"]
    ens = [film(20 + k, kind) for k, kind in enumerate(kinds)]
    E = M.Ensemble([h for (_, _, h) in ens])
    E.step(12)
    for k, (kind, (g, _d, _h)) in enumerate(zip(kinds, ens)):
        solo_g, _sd, solo = film(20 + k, kind)
        solo.step(12)
        a, b = flat(g["ss"].spins()[:3]), flat(solo_g["ss"].spins()[:3])
        print(f"ensemble member {kind}: max|ensemble - solo| = {np.abs(a - b).max():.3e}", flush=True)
        assert np.array_equal(a, b)


def _mag2d(nx, ny, nz, trunc):
    m = M.Magnetostatics2D(nx, ny, nz)
    m.setUnitCell(1, [1, 0, 0], [0.2, 1.1, 0], [0.1, 0.0, 1.5])
    m.setGrainSize(1, 1.0, 0.9, 1.2)
    m.setGrainSize(2, 0.8, 1.0, 0.5)
    m.setTruncation(*trunc)
    ABC = np.tile(np.eye(3), (nz, 1, 1))
    ABC[0] = ((1, 0, 0), (0.2, 1.1, 0), (0.1, 0.0, 1.5))
    grain = np.ones((nz, 3))
    grain[0], grain[1] = (1.0, 0.9, 1.2), (0.8, 1.0, 0.5)
    sms = np.array([1.0 / m.grainSize(k + 1)[3] for k in range(nz)])
    dfs = np.full(nz, 4 * np.pi)
    for k in range(nz):
        m.setSourceMomentScale(k + 1, sms[k])
        m.setDestinationFieldScale(k + 1, dfs[k])
    return m, O.mag2d_tensor(nx, ny, nz, ABC, grain, trunc), sms, dfs


def test_magnetostatics2d_against_oracle():
    """Magnetostatics2D on a 32x32x2 film with a sheared cell, a lifted second layer,
    unequal prisms and truncation (12, 10, 8, 1), the 1/V and 4 pi scales set as a script
    would: the tensors (the library's asinh/atan form of Newell's f and g) against the
    oracle's makeData (the reference's log/atan2 form), and the field against the oracle's
    apply."""
    nx, ny, nz = 32, 32, 2
    m, T, sms, dfs = _mag2d(nx, ny, nz, (12, 10, 8, 1))
    ss = M.SpinSystem(nx, ny, nz)
    x, y, z = O.random_unit_spins(11, nx * ny * nz)
    ss.setSpins(x, y, z)
    m.apply(ss)
    got = np.array([[[m.tensorArray(a + 1, b + 1, n).toArray() for b in range(nz)] for a in range(nz)]
                    for n in M.LongRange2D.NAMES])
    terr = float(np.abs(got - T).max() / np.abs(T).max())
    print(f"Magnetostatics2D tensors vs oracle: max rel diff = {terr:.3e}", flush=True)
    assert terr <= 1e-10
    assert m.getMatrix(1, 2, "ZZ", 0, 0) == got[8, 0, 1, 0, 0]
    o = O.LongRange2D(nx, ny, nz, slot="Magnetostatics2D", threads=True)
    o.setTensors(T)
    o.setLayerScales(sms, dfs)
    os_ = O.SpinSystem(nx, ny, nz, threads=True)
    os_.setSpins(x, y, z)
    o.apply(os_)
    hg, ho = flat(ss.field("Magnetostatics2D")), flat(os_.field("Magnetostatics2D"))
    err = float(np.abs(hg - ho).max() / np.abs(ho).max())
    print(f"Magnetostatics2D 32x32x2 field: max rel err = {err:.3e}", flush=True)
    assert err <= 5e-6


def test_heun_with_magnetostatics2d_against_oracle():
    """The fused Heun driver with a Magnetostatics2D operator (scaled down by 0.02 beside
    exchange J=1) on a 32x32x2 film: 30 steps against the oracle's Heun with the same
    tensors."""
    spec = W.FilmSpec(32, 32, 2, J=1.0, K=0.05, g=0.0, alpha=0.1, dt=0.005, T=0.0)
    g = W.build(M, spec, with_dipole=False)
    o = W.build(O, spec, with_dipole=False, threads=True)
    x, y, z = O.random_unit_spins(6, spec.n)
    g["ss"].setSpins(x, y, z)
    o["ss"].setSpins(x, y, z)
    m, T, sms, dfs = _mag2d(spec.nx, spec.ny, spec.nz, (8, 8, 8, 1))
    m.setScale(0.02)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], m, None, g["llg"])
    heun.step(30)
    od = O.LongRange2D(spec.nx, spec.ny, spec.nz, slot="Magnetostatics2D", threads=True)
    od.setTensors(T)
    od.setLayerScales(sms, dfs)
    od.setScale(0.02)
    sim = O.Sim(o["ss"], o["ex"], o["ani"], None, None, o["llg"])
    sim.setLongRange2D(od)
    sim.heun(30)
    err = float(np.abs(flat(g["ss"].spins()[:3]) - flat(o["ss"].spins()[:3])).max())
    print(f"Heun + Magnetostatics2D 32x32x2, 30 steps: max|dm| = {err:.3e}", flush=True)
    assert err <= 2e-6
