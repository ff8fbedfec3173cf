"""GPU parity of AppliedField and AppliedField.Heterogeneous (operator API and the
fused step driver) against the FP64 oracle, and a spin precessing in a uniform
applied field against the closed-form damped precession."""
import numpy as np
import pytest

import maglua_amd as M
from maglua_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def flat(t):
    return np.stack([np.asarray(v).reshape(-1) for v in t])


def test_applied_field_operator_slots():
    dims = (4, 3, 2)
    ss = M.SpinSystem(*dims)
    ss.randomize(1)
    af = M.AppliedField(*dims)
    af.set([0.5, -1.0])                 # a missing component is 0
    assert af.get() == (0.5, -1.0, 0.0)
    af.add(0.0, 0.0, 2.0)
    af.setScale(3.0)
    af.apply(ss)
    af.apply(ss)                        # the slot is overwritten, not accumulated
    assert np.array_equal(flat(ss.field("Zeeman")), np.array([[1.5] * 24, [-3.0] * 24, [6.0] * 24]))
    assert (af.x(), af.y(), af.z()) == (0.5, -1.0, 2.0)
    hf = M.AppliedField.Heterogeneous(*dims)
    hf.set([2, 3, 1], [1.0, 2.0, -0.5])  # 1-based site (2,3,1) is index 1 + 4*2 = 9
    assert hf.get([2, 3, 1]) == [1.0, 2.0, -0.5]
    hf.setScale(2.0)
    hf.apply(ss)
    hf.apply(ss)                        # accumulates until resetFields
    want = np.zeros((3, 24))
    want[:, 9] = [4.0, 8.0, -2.0]
    assert np.array_equal(flat(ss.field("Zeeman_Heterogeneous")), want)
    ss.resetFields()
    hf.apply(ss)
    assert np.array_equal(flat(ss.field("Zeeman_Heterogeneous")), want / 2)
    with pytest.raises(M.MagLuaError):
        hf.set([5, 1, 1], [1.0, 0.0, 0.0])


def test_fused_heun_with_applied_fields_matches_oracle():
    spec = W.FilmSpec(16, 16, 2, J=1.0, K=0.05, g=0.0, T=0.0, dt=0.005)
    dims = (spec.nx, spec.ny, spec.nz)
    gpu, ora = W.build(M, spec), W.build(O, spec)
    x, y, z = O.random_unit_spins(4, spec.n)
    gpu["ss"].setSpins(x, y, z)
    ora["ss"].setSpins(x, y, z)
    rng = np.random.default_rng(2)
    sites = rng.choice(spec.n, 40, replace=False)
    vals = rng.standard_normal((40, 3)) * 0.5
    ag, ao = M.AppliedField(*dims), O.AppliedField(*dims)
    hg, ho = M.AppliedField.Heterogeneous(*dims), O.AppliedField(*dims, heterogeneous=True)
    for a in (ag, ao):
        a.set(0.3, -0.2, 0.5)
    for i, v in zip(sites, vals):
        x0, r = int(i) % spec.nx, int(i) // spec.nx
        hg.set([x0 + 1, r % spec.ny + 1, r // spec.ny + 1], list(v))
        ho.setSite(int(i), *v)
    hg.setScale(1.5)
    ho.setScale(1.5)
    heun = M.Heun(gpu["ss"], gpu["ex"], gpu["ani"], None, None, gpu["llg"])
    heun.addAppliedField(ag)
    heun.addAppliedField(hg)
    sim = O.Sim(ora["ss"], ora["ex"], ora["ani"], None, None, ora["llg"])
    sim.addAppliedField(ao)
    sim.addAppliedField(ho)
    heun.step(100)
    sim.heun(100)
    d1 = float(np.abs(flat(gpu["ss"].spins()[:3]) - flat(ora["ss"].spins()[:3])).max())
    # a change after adding applies from the next step() call
    for a in (ag, ao):
        a.set(0.0, 0.0, 1.0)
    heun.step(50)
    sim.heun(50)
    d2 = float(np.abs(flat(gpu["ss"].spins()[:3]) - flat(ora["ss"].spins()[:3])).max())
    print(f"applied fields: max|dm| {d1:.2e} after 100 steps, {d2:.2e} after 150", flush=True)
    assert d1 <= 1e-5 and d2 <= 1e-5


def test_spin_precession_in_an_applied_field():
    """One spin, uniform field B along z, alpha 0.1: the fused Heun steps converge to
    the closed-form damped precession at second order."""
    B, alpha, gamma, th0 = 1.0, 0.1, 1.0, 1.0

    def exact(t):
        gp = gamma / (1 + alpha ** 2)
        th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-alpha * gp * B * t))
        return np.array([np.sin(th) * np.cos(gp * B * t), np.sin(th) * np.sin(gp * B * t), np.cos(th)])

    errs = []
    for dt in (0.02, 0.01):
        ss = M.SpinSystem(1, 1, 1)
        ss.setSpins([np.sin(th0)], [0.0], [np.cos(th0)])
        ss.setAlpha(alpha)
        ss.setGamma(gamma)
        ss.setTimeStep(dt)
        af = M.AppliedField(ss)
        af.set(0.0, 0.0, B)
        heun = M.Heun(ss, None, None, None, None, M.LLG.Cartesian())
        heun.addAppliedField(af)
        heun.step(int(round(2.0 / dt)))
        s = flat(ss.spins()[:3])[:, 0]
        errs.append(float(np.abs(s - exact(2.0)).max()))
    print(f"precession errors {errs}", flush=True)
    assert 3.4 < errs[0] / errs[1] < 4.6 and errs[1] < 2e-5


def test_adaptive_rk_steps_follow_the_closed_form():
    """make_adapt_step_function around the RK2 closure on the operator API (it passes
    the current time step to its stage systems, RungeKutta.lua:198): a spin in a
    uniform AppliedField reaches the closed-form state at t = 2 to the tolerance's
    accuracy, with rejected and accepted steps along the way."""
    B, alpha, gamma, th0 = 1.0, 0.1, 1.0, 1.0
    gp = gamma / (1 + alpha ** 2)
    th = 2 * np.arctan(np.tan(th0 / 2) * np.exp(-alpha * gp * B * 2.0))
    want = np.array([np.sin(th) * np.cos(gp * B * 2.0), np.sin(th) * np.sin(gp * B * 2.0), np.cos(th)])
    ss = M.SpinSystem(1, 1, 1)
    ss.setSpins([np.sin(th0)], [0.0], [np.cos(th0)])
    ss.setAlpha(alpha)
    ss.setGamma(gamma)
    ss.setTimeStep(0.5)
    af = M.AppliedField(ss)
    af.set(0.0, 0.0, B)
    llg = M.LLG.Cartesian()

    def calc(s):
        s.resetFields()
        af.apply(s)
        s.sumFields()

    step = M.make_rk_step_function(ss, "RK2", calc, llg)
    adapt = M.make_adapt_step_function(ss, step, 1e-5, llg)
    results = []
    while ss.time() < 2.0 - 1e-9 and len(results) < 5000:
        if ss.time() + ss.timeStep() > 2.0:
            ss.setTimeStep(2.0 - ss.time())
        results.append(adapt())
    err = float(np.abs(flat(ss.spins()[:3])[:, 0] - want).max())
    print(f"adaptive: {results.count(True)} accepted, {results.count(False)} rejected, err {err:.2e}", flush=True)
    assert abs(ss.time() - 2.0) < 1e-9
    assert results.count(False) > 0 and results.count(True) > 5
    assert err < 1e-4
