"""GPU parity of Anisotropy.Uniaxial (operator API, and inside the PredictorCorrector
step closure over the operator API) against the FP64 oracle, with the reference's
own pairing of sz with nx in the dot product (spinoperationanisotropy_uniaxial.cpp:287)
and its binding's K2 argument handling."""
import numpy as np
import pytest

import maglua_amd as M
from maglua_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def flat(t):
    return np.stack([np.asarray(v).reshape(-1) for v in t])


def site3(i, dims):
    nx, ny, _ = dims
    return [i % nx + 1, (i // nx) % ny + 1, i // (nx * ny) + 1]


def random_ops(dims, seed, count):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    sites = rng.integers(0, n, count)                    # repeats: several ops on a site
    axes = rng.standard_normal((count, 3))
    axes[::17] = 0.0                                     # zero axes are not added
    K1 = rng.uniform(-1, 1, count)
    K2 = rng.uniform(-0.5, 0.5, count)
    return sites, axes, K1, K2


def build_pair(dims, seed, count, scale=0.7):
    gu, ou = M.Anisotropy.Uniaxial(*dims), O.AnisotropyUniaxial(*dims)
    gu.setScale(scale)
    ou.setScale(scale)
    sites, axes, K1, K2 = random_ops(dims, seed, count)
    for i, a, k1, k2 in zip(sites, axes, K1, K2):
        # the host mirror's add(site, axis, K1) takes K2 = 0; pass both through the ABI
        added = gu.add(site3(int(i), dims), a, k1)
        assert added == ou.add(int(i), a, k1, 0.0)
    return gu, ou


@pytest.mark.parametrize("dims", [(8, 6, 3), (33, 5, 2)])
def test_uniaxial_operator_matches_oracle(dims):
    n = dims[0] * dims[1] * dims[2]
    gu, ou = build_pair(dims, n, 3 * n)
    rng = np.random.default_rng(n + 1)
    s = rng.standard_normal((3, n)) * rng.uniform(0.5, 2.0, n)    # |s| = ms varies
    gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims)
    gs.setSpins(*s)
    os_.setSpins(*s)
    gu.apply(gs)
    ou.apply(os_)
    hg, ho = flat(gs.field("UniaxialAnisotropy")), flat(os_.field("UniaxialAnisotropy"))
    err = np.abs(hg - ho).max() / np.abs(ho).max()
    print(f"uniaxial {dims}: rel err {err:.2e}", flush=True)
    assert err <= 2e-6


def test_uniaxial_reference_quirks():
    dims = (2, 1, 1)
    u = M.Anisotropy.Uniaxial(*dims)
    assert u.add([1, 1, 1], [0, 0, 2], 3.0)            # axis along z
    assert not u.add([2, 1, 1], [0, 0, 0], 3.0)         # zero axis: not added
    assert u.add([2, 1, 1], [1, 0, 0], 2.0, 0.25)       # a fourth number: K2 takes K1's value
    assert u.numberOfAxes() == 2
    assert u.axis(2) == ([1.0, 0.0, 0.0], 2.0, 2.0)
    assert u.get([1, 1, 1]) == ([0.0, 0.0, 1.0], 3.0, 0.0)
    ss = M.SpinSystem(*dims)
    ss.setSpins([0.6, 0.6], [0.0, 0.0], [0.8, 0.8])
    u.apply(ss)
    h = flat(ss.field("UniaxialAnisotropy"))
    # site 1: n = z, so d = sx*0 + sy*0 + sz*0 = 0 and the field vanishes;
    # site 2: n = x, d = sx + sz = 1.4, h = (2*2*1.4 + 4*2*1.4^3, 0, 0)
    assert np.allclose(h[:, 0], 0.0)
    assert np.allclose(h[:, 1], [2 * 2.0 * 1.4 + 4 * 2.0 * 1.4 ** 3, 0, 0], rtol=1e-6)


def test_uniaxial_in_the_pc_step_closure_matches_oracle():
    spec = W.FilmSpec(16, 16, 2, J=1.0, K=0.0, g=0.0, T=0.0, dt=0.005)
    dims = (spec.nx, spec.ny, spec.nz)
    gpu, ora = W.build(M, spec), W.build(O, spec)
    x, y, z = O.random_unit_spins(6, spec.n)
    gpu["ss"].setSpins(x, y, z)
    ora["ss"].setSpins(x, y, z)
    gu, ou = build_pair(dims, 7, 2 * spec.n, scale=1.0)

    def calc(ss):
        ss.resetFields()
        gpu["ex"].apply(ss)
        gu.apply(ss)
        ss.sumFields()

    step = M.make_pc_step_function(gpu["ss"], calc, gpu["llg"], 1e9)      # one corrector pass = Heun
    sim = O.Sim(ora["ss"], ora["ex"], None, None, None, ora["llg"])
    sim.addUniaxial(ou)
    for _ in range(40):
        step()
    sim.heun(40)
    d = float(np.abs(flat(gpu["ss"].spins()[:3]) - flat(ora["ss"].spins()[:3])).max())
    print(f"uniaxial PC closure vs oracle over 40 steps: max|dm| {d:.2e}", flush=True)
    assert d <= 1e-5


def test_uniaxial_in_the_fused_step_matches_oracle():
    spec = W.FilmSpec(16, 16, 2, J=1.0, K=0.0, g=0.0, T=0.0, dt=0.005)
    dims = (spec.nx, spec.ny, spec.nz)
    gpu, ora = W.build(M, spec), W.build(O, spec)
    x, y, z = O.random_unit_spins(8, spec.n)
    gpu["ss"].setSpins(x, y, z)
    ora["ss"].setSpins(x, y, z)
    gu, ou = build_pair(dims, 9, 2 * spec.n, scale=1.0)
    heun = M.Heun(gpu["ss"], gpu["ex"], None, None, None, gpu["llg"])
    heun.setUniaxial(gu)
    sim = O.Sim(ora["ss"], ora["ex"], None, None, None, ora["llg"])
    sim.addUniaxial(ou)
    heun.step(60)
    sim.heun(60)
    d1 = float(np.abs(flat(gpu["ss"].spins()[:3]) - flat(ora["ss"].spins()[:3])).max())
    # ops added after setUniaxial apply from the next step() call
    for i in range(0, spec.n, 5):
        assert gu.add(site3(i, dims), [0.3, -0.4, 0.5], 0.8) == ou.add(i, [0.3, -0.4, 0.5], 0.8, 0.0)
    heun.step(40)
    sim.heun(40)
    d2 = float(np.abs(flat(gpu["ss"].spins()[:3]) - flat(ora["ss"].spins()[:3])).max())
    print(f"uniaxial fused step vs oracle: max|dm| {d1:.2e} after 60 steps, {d2:.2e} after 100", flush=True)
    assert d1 <= 1e-5 and d2 <= 1e-5
