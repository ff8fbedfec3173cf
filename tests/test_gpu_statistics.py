"""T > 0 parity: statistics of independent runs, GPU FP32 against the FP64 oracle.

The thermal field is random, so at T > 0 trajectories are compared in distribution:
16 independent runs on each side, with different noise streams (the GPU keys its
Philox stream by 100 + i, the oracle by 1000 + i), must give ensemble means of
|M| per site and energy per site within 2 sigma of each other (sigma = the
combined standard error of the two means).  Energy per site is
-1/2 sum_i m_i . (H_exchange + H_anisotropy + H_dipole): every term is bilinear in
the spins, so this is the Hamiltonian (the reference's energy scripts sum
fieldArray:dot(spinArray) per slot, cpu/magnetostatics3d/EnergyTest.lua:64-67).
The seeds are fixed, so each test is deterministic.

The paramagnet (J = K = g = 0, applied field B) checks the thermal field against the
closed form: <m.B^> = L(ms |B| / T_eff) with the Langevin function L(x) = coth x - 1/x.
The reference's noise, sigma^2 = 2 alpha T / (ms dt gamma) on the precession term
of an LL equation with prefactor gamma / (1 + alpha^2)
(cpu/thermal/spinoperationthermal.cpp:135-191, cpu/llg/llgcartesian.cpp:159-187),
makes T_eff = T / (1 + alpha^2): the Fokker-Planck stationary state of that
equation.  The oracle test pins the same value on the CPU."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import maglua_amd as M
from maglua_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RUNS = 16
POOL = min(16, os.cpu_count() or 1)


def langevin(x):
    return 1.0 / np.tanh(x) - 1.0 / x


def _energy(ss, ops, slots):
    ss.resetFields()
    for op in ops:
        op.apply(ss)
    return -0.5 * sum(ss.fieldDot(s) for s in slots)


def _sample(ss, ops, slots, n):
    x, y, z, _m = ss.spins()
    mag = float(np.sqrt(np.sum(x) ** 2 + np.sum(y) ** 2 + np.sum(z) ** 2)) / n
    return mag, _energy(ss, ops, slots) / n


def _ordered(n):
    return np.zeros(n), np.zeros(n), np.ones(n)


def _oracle_run(spec, lr3, seed, eq, every, samples):
    o = W.build(O, spec, with_dipole=not lr3)
    o["ss"].setSpins(*_ordered(spec.n))
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"], seed=seed)
    ops = [o["ex"], o["ani"]]
    slots = ["Exchange", "Anisotropy"]
    if lr3:
        lr = O.LongRange3D(spec.nx, spec.ny, spec.nz)
        lr.setDipole(O.dipole_tensor3d(spec.nx, spec.ny, spec.nz, R=lr3))
        lr.setStrength(spec.g)
        sim.setLongRange3D(lr)
        ops.append(lr)
        slots.append("LongRange3D")
    else:
        ops.append(o["dip"])
        slots.append("Dipole")
    sim.heun(eq)
    vals = []
    for _ in range(samples):
        sim.heun(every)
        vals.append(_sample(o["ss"], ops, slots, spec.n))
    return np.mean(vals, axis=0)


def _oracle_runs(spec, lr3, eq, every, samples):
    with ThreadPoolExecutor(POOL) as ex:
        futs = [ex.submit(_oracle_run, spec, lr3, 1000 + i, eq, every, samples) for i in range(RUNS)]
        return np.array([f.result() for f in futs])


def _gpu_member(spec, lr3, seed):
    g = W.build(M, spec, with_dipole=not lr3)
    g["ss"].setSpins(*_ordered(spec.n))
    g["temp"].setSeed(seed)
    ops = [g["ex"], g["ani"]]
    slots = ["Exchange", "Anisotropy"]
    if lr3:
        lr = M.LongRange3D(spec.nx, spec.ny, spec.nz)
        lr.setRadius(lr3)
        lr.setStrength(spec.g)
        g["lr"] = lr
        ops.append(lr)
        slots.append("LongRange3D")
    else:
        ops.append(g["dip"])
        slots.append("Dipole")
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["lr"] if lr3 else g["dip"], g["temp"], g["llg"])
    return g, heun, ops, slots


def _compare(name, gpu, ora):
    """Both columns (|M|/n, E/n): |mean_gpu - mean_oracle| <= 2 sigma."""
    for k, what in enumerate(("|M|/n", "E/n")):
        mg, mo = gpu[:, k].mean(), ora[:, k].mean()
        se = np.sqrt(gpu[:, k].var(ddof=1) / len(gpu) + ora[:, k].var(ddof=1) / len(ora))
        print(f"{name} {what}: GPU {mg:.5f}  oracle {mo:.5f}  diff {mg - mo:+.2e}  2sigma {2 * se:.2e}", flush=True)
        assert abs(mg - mo) <= 2 * se, (name, what)


def test_film_t05_ensemble_means_within_2sigma():
    """The 256x256x8 T = 0.5 film's parameters (J=1, K=0.05 z, g=0.02, alpha=0.1,
    dt=0.002, periodic in-plane, open z) on a 16x16x4 film: 2000 steps from the
    ordered state, then 20 samples 100 steps apart, time-averaged per run."""
    spec = W.FilmSpec(16, 16, 4, J=1.0, K=0.05, g=0.02, alpha=0.1, dt=0.002, T=0.5)
    eq, every, samples = 2000, 100, 20
    gpu = []
    for i in range(RUNS):
        g, heun, ops, slots = _gpu_member(spec, 0, 100 + i)
        heun.step(eq)
        vals = []
        for _ in range(samples):
            heun.step(every)
            vals.append(_sample(g["ss"], ops, slots, spec.n))
        gpu.append(np.mean(vals, axis=0))
    _compare("film 16x16x4 T=0.5", np.array(gpu), _oracle_runs(spec, 0, eq, every, samples))


@pytest.mark.parametrize("T", [0.1, 2.0])
def test_cube_ensemble_means_within_2sigma(T):
    """The batched replicas' parameters (J=1, K=0.1, g=0.01 LongRange3D lattice sum,
    periodic) on 8x8x8 cubes at both ends of their temperature range: 16 replicas in
    one Ensemble on the GPU, 1500 steps from the ordered state, then 20 samples 50
    steps apart."""
    spec = W.FilmSpec(8, 8, 8, J=1.0, K=0.1, g=0.01, pbc=(1, 1, 1), alpha=0.1, dt=0.005, T=T)
    eq, every, samples, R = 1500, 50, 20, 8
    members = [_gpu_member(spec, R, 100 + i) for i in range(RUNS)]
    E = M.Ensemble([h for (_, h, _, _) in members])
    E.step(eq)
    acc = [[] for _ in members]
    for _ in range(samples):
        E.step(every)
        for k, (g, _h, ops, slots) in enumerate(members):
            acc[k].append(_sample(g["ss"], ops, slots, spec.n))
    gpu = np.array([np.mean(a, axis=0) for a in acc])
    _compare(f"cube 8^3 T={T}", gpu, _oracle_runs(spec, R, eq, every, samples))


def test_paramagnet_langevin():
    """J = K = g = 0, B = 1 along z, T = 1, alpha = 0.5, dt = 0.001 on 128x128 spins:
    <m_z> = L((1 + alpha^2) B / T) = 0.3789 (not L(B / T) = 0.3130).  20000 steps to
    equilibrate, then 40 samples 2000 steps (one time unit) apart."""
    nx, ny, alpha, dt, T, B = 128, 128, 0.5, 0.001, 1.0, 1.0
    n = nx * ny
    ss = M.SpinSystem(nx, ny, 1)
    ss.setAlpha(alpha)
    ss.setTimeStep(dt)
    ss.setSpins(*_ordered(n))
    th = M.Thermal(nx, ny, 1, seed=12345)
    th.setTemperature(T)
    af = M.AppliedField(nx, ny, 1)
    af.set(0.0, 0.0, B)
    heun = M.Heun(ss, None, None, None, th, M.LLG.Cartesian())
    heun.addAppliedField(af)
    heun.step(20000)
    mz = []
    for _ in range(40):
        heun.step(2000)
        mz.append(ss.netMoment()[2] / n)
    mean, se = float(np.mean(mz)), float(np.std(mz, ddof=1) / np.sqrt(len(mz)))
    want, naive = langevin((1 + alpha ** 2) * B / T), langevin(B / T)
    print(f"paramagnet <mz> = {mean:.4f} +- {se:.4f}; L((1+a^2)B/T) = {want:.4f}; L(B/T) = {naive:.4f}", flush=True)
    # the statistical error plus the step's bias measured on the oracle (< 1e-3 at dt = 0.001)
    assert abs(mean - want) <= 4 * se + 1.5e-3
    assert abs(mean - naive) > 0.05
