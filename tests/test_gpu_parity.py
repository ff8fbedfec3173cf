"""GPU parity: libmsd.so kernels (through the C ABI) against the FP64 oracle on
the same seeded inputs.  Floating-point tolerances are stated per test; the
integer Philox stream is bit-exact."""
import os

import numpy as np
import pytest

import maglua_amd as M
from maglua_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def flat(t):
    return np.stack([np.asarray(v).reshape(-1) for v in t])


def twin(spec, seed=None, dipole=True, thermal=None):
    """The same workload on the GPU and on the oracle, from the same FP64 state."""
    g = W.build(M, spec, with_dipole=dipole, with_thermal=thermal)
    o = W.build(O, spec, with_dipole=dipole, with_thermal=thermal)
    x, y, z = O.random_unit_spins(spec.seed if seed is None else seed, spec.n)
    g["ss"].setSpins(x, y, z)
    o["ss"].setSpins(x, y, z)
    return g, o


# ------------------------------------------------------------------- RNG
def test_philox_device_known_answers():
    from tests.test_oracle import KAT
    for ctr, key, want in KAT:
        out = M.philox_device(np.array([ctr], dtype=np.uint32), key)
        assert tuple(int(v) for v in out[0]) == want


def test_philox_device_matches_oracle_stream():
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2 ** 32, size=(257, 4), dtype=np.uint64).astype(np.uint32)
    out = M.philox_device(ctr, (123, 456))
    for i in range(0, 257, 16):
        assert list(out[i]) == O.philox([int(v) for v in ctr[i]], (123, 456))


def test_random_spins_match_oracle():
    ss = M.SpinSystem(17, 9, 3)
    ss.randomize(5)
    x, y, z, m = ss.spins()
    ox, oy, oz = O.random_unit_spins(5, 17 * 9 * 3)
    assert np.abs(flat((x, y, z)) - np.stack([ox, oy, oz])).max() < 1e-7
    assert np.allclose(m, 1.0, atol=1e-7)


# ------------------------------------------------------------ operators
@pytest.mark.parametrize("dims,pbc", [((8, 8, 3), (1, 1, 0)), ((5, 7, 2), (1, 1, 1)), ((16, 1, 1), (1, 0, 0))])
def test_exchange_op(dims, pbc):
    nx, ny, nz = dims
    n = nx * ny * nz
    rng = np.random.default_rng(1)
    s = rng.standard_normal((3, n)) * rng.uniform(0.5, 2.0, n)   # non-unit moments exercise normalisation
    gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims)
    gs.setSpins(*s)
    os_.setSpins(*s)
    f, t = W.nn_paths(nx, ny, nz, pbc)
    J = rng.uniform(-1, 2, f.size)
    ge, oe = M.Exchange(*dims), O.Exchange(*dims)
    ge.addPaths(f, t, J)
    oe.addPaths(f, t, J)
    ge.apply(gs)
    oe.apply(os_)
    hg, ho = flat(gs.field("Exchange")), flat(os_.field("Exchange"))
    assert np.abs(hg - ho).max() <= 2e-6 * np.abs(ho).max()


def test_anisotropy_op_modes():
    dims = (6, 5, 2)
    n = 60
    rng = np.random.default_rng(2)
    s = rng.standard_normal((3, n))
    for mode in ("uniform", "per-site", "stacked"):
        gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims)
        gs.setSpins(*s)
        os_.setSpins(*s)
        ga, oa = M.Anisotropy(*dims), O.Anisotropy(*dims)
        if mode == "uniform":
            ga.addMany(np.arange(n), (0.2, 0.1, 1.0), 0.3)
            oa.addMany(np.arange(n), (0.2, 0.1, 1.0), 0.3)
        else:
            ax = rng.standard_normal((n, 3))
            K = rng.uniform(-1, 1, n)
            ga.addMany(np.arange(n), ax, K)
            oa.addMany(np.arange(n), ax, K)
            if mode == "stacked":
                ga.addMany([3, 3], [(1, 0, 0), (0, 1, 1)], [0.5, -0.2])
                oa.addMany([3, 3], [(1, 0, 0), (0, 1, 1)], [0.5, -0.2])
        ga.apply(gs)
        oa.apply(os_)
        hg, ho = flat(gs.field("Anisotropy")), flat(os_.field("Anisotropy"))
        assert np.abs(hg - ho).max() <= 2e-6 * max(1.0, np.abs(ho).max()), mode


def test_thermal_op_same_stream_as_oracle():
    dims = (16, 8, 2)
    n = 256
    rng = np.random.default_rng(3)
    s = rng.standard_normal((3, n))
    gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims)
    for ss in (gs, os_):
        ss.setSpins(*s)
        ss.setAlpha(0.2)
        ss.setTimeStep(0.01)
    gt = M.Thermal(*dims, seed=99)
    gt.setTemperature(0.8)
    gt.scaleSite((3, 1, 1), 0.5)
    ot = O.Thermal(*dims)
    ot.setTemperature(0.8)
    ot.scaleSite(2, 0.5)
    for draw in range(3):
        gt.apply(gs)
        ot.apply(os_, seed=99, step=draw)
        hg, ho = flat(gs.field("Thermal")), flat(os_.field("Thermal"))
        assert np.abs(hg - ho).max() <= 5e-6 * np.abs(ho).max(), draw
    assert gt.draw() == 3


@pytest.mark.parametrize("dims", [(8, 8, 1), (8, 8, 2), (12, 6, 3), (20, 10, 4), (32, 32, 1), (10, 15, 8), (64, 48, 5)])
def test_dipole_op(dims):
    nx, ny, nz = dims
    n = nx * ny * nz
    rng = np.random.default_rng(nx * 100 + ny)
    s = rng.standard_normal((3, n))
    gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims, threads=True)
    gs.setSpins(*s)
    os_.setSpins(*s)
    gd, od = M.Dipole(*dims), O.Dipole(*dims, threads=True)
    for d in (gd, od):
        d.setStrength(0.3)
        d.setTruncation(60)
    gd.apply(gs)
    od.apply(os_)
    hg, ho = flat(gs.field("Dipole")), flat(os_.field("Dipole"))
    # FP32 spectra and transforms against FP64: relative to the field scale
    assert np.abs(hg - ho).max() <= 2e-5 * np.abs(ho).max()


def test_dipole_op_full_size_film():
    """256x256x8 at the default truncation (1000): one evaluation, FP32 vs FP64."""
    spec = W.FILM_256_T0
    g = M.SpinSystem(spec.nx, spec.ny, spec.nz)
    o = O.SpinSystem(spec.nx, spec.ny, spec.nz, threads=True)
    x, y, z = O.random_unit_spins(0, spec.n)
    g.setSpins(x, y, z)
    o.setSpins(x, y, z)
    gd = M.Dipole(spec.nx, spec.ny, spec.nz)
    od = O.Dipole(spec.nx, spec.ny, spec.nz, threads=True)
    gd.setStrength(spec.g)
    od.setStrength(spec.g)
    gd.apply(g)
    od.apply(o)
    hg, ho = flat(g.field("Dipole")), flat(o.field("Dipole"))
    assert np.abs(hg - ho).max() <= 2e-5 * np.abs(ho).max()


def test_dipole_math_huge_against_oracle():
    """setTruncation(math.huge) on a 4x2x2 film: every tensor element (~1e8 image terms
    each, FP64 on the device) and one Dipole:apply against the oracle's ring sum."""
    dims = (4, 2, 2)
    gd, od = M.Dipole(*dims), O.Dipole(*dims, threads=True)
    gd.setTruncation(float("inf"))
    od.setTruncation(float("inf"))
    assert gd.truncation() == (float("inf"),) * 4
    T = O.dipole_tensor(*dims, gmax=(-1, -1, -1, -1))
    worst, big = 0.0, max(np.abs(v).max() for v in T.values())
    for name in M.Dipole.NAMES:
        for c in range(3):
            for y in range(2):
                for x in range(4):
                    v = gd.getMatrix(name, (x, y, c - 1))
                    worst = max(worst, abs(v - T[name][c, y, x]))
    print(f"math.huge tensor 4x2x2: max |diff| = {worst:.3e} (max |T| = {big:.3f})", flush=True)
    assert worst <= 1e-12 * big
    s = np.random.default_rng(5).standard_normal((3, 16))
    gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims)
    gs.setSpins(*s)
    os_.setSpins(*s)
    for d in (gd, od):
        d.setStrength(0.5)
    gd.apply(gs)
    od.apply(os_)
    hg, ho = flat(gs.field("Dipole")), flat(os_.field("Dipole"))
    err = float(np.abs(hg - ho).max() / np.abs(ho).max())
    print(f"math.huge Dipole:apply 4x2x2: max rel err = {err:.3e}", flush=True)
    assert err <= 2e-6


def test_dipole_matrix_get_set():
    gd, od = M.Dipole(8, 6, 2), O.Dipole(8, 6, 2)
    for d in (gd, od):
        d.setTruncation(30)
    T = {k: od.tensor(k) for k in od.NAMES}
    for name, off in (("XX", (1, 0, 0)), ("XZ", (3, 2, -1)), ("YZ", (-1, 2, 1)), ("ZZ", (0, 0, 0))):
        want = T[name][off[2] + 1, off[1] % 6, off[0] % 8]
        assert abs(gd.getMatrix(name, off) - want) <= 1e-12 * max(1.0, abs(want))
    # a custom tensor keeping the lattice-sum parities is used as given
    s = np.random.default_rng(0).standard_normal((3, 96))
    gs, os_ = M.SpinSystem(8, 6, 2), O.SpinSystem(8, 6, 2)
    gs.setSpins(*s)
    os_.setSpins(*s)
    for d in (gd,):
        d.setMatrix("ZZ", (0, 0, 1), 0.5)
        d.setMatrix("ZZ", (0, 0, -1), 0.5)
    T["ZZ"][2, 0, 0] = 0.5
    T["ZZ"][0, 0, 0] = 0.5
    od.setTensor("ZZ", T["ZZ"])
    gd.apply(gs)
    od.apply(os_)
    hg, ho = flat(gs.field("Dipole")), flat(os_.field("Dipole"))
    assert np.abs(hg - ho).max() <= 2e-5 * np.abs(ho).max()
    # a tensor without those parities is rejected, not silently mis-applied
    gd.setMatrix("XY", (1, 0, 0), 3.0)
    with pytest.raises(M.MagLuaError):
        gd.apply(gs)


def test_sum_and_reset_fields_semantics():
    spec = W.FilmSpec(8, 8, 2, T=0.0)
    g, o = twin(spec)
    for d in (g, o):
        ss = d["ss"]
        ss.resetFields()
        d["ex"].apply(ss)
        d["ani"].apply(ss)
        d["dip"].apply(ss)
        ss.sumFields()
    tg, to = flat(g["ss"].field("Total")), flat(o["ss"].field("Total"))
    assert np.abs(tg - to).max() <= 2e-6 * np.abs(to).max()
    g["ss"].resetFields()
    assert not g["ss"].slotUsed("Exchange")
    assert np.abs(flat(g["ss"].field("Exchange"))).max() == 0
    g["ss"].sumFields()
    assert np.abs(flat(g["ss"].field("Total"))).max() == 0


def test_llg_op_apply_forms():
    spec = W.FilmSpec(8, 8, 2, T=0.0, dt=0.01)
    g, o = twin(spec)
    for d, api in ((g, M), (o, O)):
        ss = d["ss"]
        ss.resetFields()
        d["ex"].apply(ss)
        d["ani"].apply(ss)
        d["dip"].apply(ss)
        ss.sumFields()
    P_g, P_o = g["ss"].copy(), o["ss"].copy()
    g["llg"].apply(g["ss"], 0.5, g["ss"], P_g)          # Lua form apply(from, scale, dmdt, to)
    o["llg"].apply(o["ss"], 0.5, o["ss"], P_o)
    assert np.abs(flat(P_g.spins()[:3]) - flat(P_o.spins()[:3])).max() < 1e-6
    assert abs(P_g.time() - 0.005) < 1e-15
    g["llg"].apply(g["ss"], False)                      # in place, time not advanced
    o["llg"].apply(o["ss"], 1.0, o["ss"], o["ss"], False)
    assert np.abs(flat(g["ss"].spins()[:3]) - flat(o["ss"].spins()[:3])).max() < 1e-6
    assert g["ss"].time() == 0.0


def test_reductions_fp64():
    ss = M.SpinSystem(64, 32, 3)
    ss.randomize(2)
    x, y, z, m = ss.spins()
    mom = ss.netMoment()
    want = [x.sum(), y.sum(), z.sum()]
    assert np.allclose(mom[:3], want, atol=1e-9)
    assert np.allclose(mom[4:7], [(x * x).sum(), (y * y).sum(), (z * z).sum()], rtol=1e-12)
    ss2 = ss.copy()
    assert ss.diff(ss2)[3] == 0.0
    ss2.setSpin((1, 1, 1), (0, 0, 1))
    d = ss.diff(ss2)
    assert abs(d[0] - abs(x[0, 0, 0])) < 1e-7 and abs(d[2] - abs(z[0, 0, 0] - 1)) < 1e-7
    ex = M.Exchange(ss)
    f, t = W.nn_paths(64, 32, 3, (1, 1, 1))
    ex.addPaths(f, t, 1.0)
    ss.resetFields()
    ex.apply(ss)
    hx, hy, hz = ss.field("Exchange")
    assert abs(ss.fieldDot("Exchange") - (hx * x + hy * y + hz * z).sum()) < 1e-6 * abs((hx * x).sum() + 1)


def test_errors_are_raised():
    a = M.SpinSystem(4, 4, 1)
    e = M.Exchange(5, 4, 1)
    with pytest.raises(M.MagLuaError, match="size mismatch"):
        e.apply(a)
    with pytest.raises(M.MagLuaError):
        M.Dipole(5, 4, 1)                                 # odd nx
    with pytest.raises(M.MagLuaError):
        M.Dipole(14, 4, 1)                                # factor 7
    with pytest.raises(M.MagLuaError):
        a.setSpin((5, 1), (1, 0, 0))


# ------------------------------------------------------------ integrators
def _max_dm(g, o):
    return float(np.abs(flat(g["ss"].spins()[:3]) - flat(o["ss"].spins()[:3])).max())


def test_fused_heun_film32_1000_steps():
    """32x32x1, J=1, K=0.1, g=0.01, T=0, Heun dt=0.005, alpha=0.1, 1000 steps from seed 0:
    FP32 GPU against the FP64 oracle, max|dm| <= 1e-5."""
    spec = W.FILM_32
    g, o = twin(spec)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"])
    for k in range(10):
        heun.step(100)
        sim.heun(100)
        err = _max_dm(g, o)
        print(f"film32 step {100 * (k + 1)}: max|dm| = {err:.3e}", flush=True)
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-9
    assert err <= 1e-5


def test_fused_heun_matches_golden_fixture():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "film32_heun.npz")
    gold = np.load(path)
    spec = W.FILM_32
    g = W.build(M, spec)
    g["ss"].setSpins(gold["x0"], gold["y0"], gold["z0"])
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    done = 0
    for k, step in enumerate(gold["steps"]):
        heun.step(int(step) - done)
        done = int(step)
        s = flat(g["ss"].spins()[:3])
        assert np.abs(s - gold["states"][k]).max() <= 1e-5, step


def test_fused_heun_film256_50_steps():
    spec = W.FILM_256_T0
    g, o = twin(spec)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    O.Sim.set_threads(8)
    og = W.build(O, spec, threads=True)
    og["ss"].setSpins(*O.random_unit_spins(spec.seed, spec.n))
    sim = O.Sim(og["ss"], og["ex"], og["ani"], og["dip"], og["temp"], og["llg"])
    heun.step(50)
    sim.heun(50)
    err = float(np.abs(flat(g["ss"].spins()[:3]) - flat(og["ss"].spins()[:3])).max())
    print(f"film256 50 steps: max|dm| = {err:.3e}", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("spec", [W.FilmSpec(32, 32, 1, T=0.3, dt=0.005, g=0.01, K=0.1),
                                  W.FilmSpec(24, 16, 3, T=0.5, dt=0.002)])
def test_fused_heun_thermal_same_noise(spec):
    """T > 0 with the same counter-based noise on both sides: trajectories agree
    over a short horizon (the FP32 Box-Muller differs from FP64 by ~1e-7)."""
    g, o = twin(spec)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"], seed=0)
    heun.step(20)
    sim.heun(20)
    err = _max_dm(g, o)
    print(f"thermal {spec.name()}: 20 steps max|dm| = {err:.3e}", flush=True)
    assert err <= 1e-5      # measured 7e-8
    assert g["temp"].draw() == 20


def test_graph_and_direct_launches_agree():
    spec = W.FilmSpec(32, 32, 2, T=0.4, dt=0.005)
    a = W.build(M, spec)
    b = W.build(M, spec)
    a["ss"].randomize(4)
    b["ss"].randomize(4)
    ha = M.Heun(a["ss"], a["ex"], a["ani"], a["dip"], a["temp"], a["llg"], graph=True)
    hb = M.Heun(b["ss"], b["ex"], b["ani"], b["dip"], b["temp"], b["llg"], graph=False)
    for _ in range(3):
        ha.step(7)
        hb.step(7)
        M.synchronize()
        assert np.array_equal(flat(a["ss"].spins()), flat(b["ss"].spins()))


def test_drivers_sharing_a_dipole_alternate():
    """Two step drivers on one Dipole object stepped in turn: each driver's cached spin
    spectrum lives in the Dipole's buffers, so each must transform its own state again
    after the other has stepped.  Both match drivers with Dipoles of their own."""
    spec = W.FilmSpec(32, 32, 2, T=0.0, dt=0.005, g=0.1)
    sys = [W.build(M, spec) for _ in range(4)]
    for k, g in enumerate(sys):
        g["ss"].randomize(10 + k % 2)
    shared = sys[0]["dip"]
    h = [M.Heun(g["ss"], g["ex"], g["ani"], shared if k < 2 else g["dip"], g["temp"], g["llg"])
         for k, g in enumerate(sys)]
    for _ in range(3):
        for k in range(4):
            h[k].step(5)
    M.synchronize()
    for k in range(2):
        assert np.array_equal(flat(sys[k]["ss"].spins()), flat(sys[k + 2]["ss"].spins())), k


@pytest.mark.parametrize("kind", ["RK2", "RK4", "BS3"])
def test_rk_driver_on_op_api(kind):
    spec = W.FilmSpec(16, 16, 2, T=0.0, dt=0.01)
    g, o = twin(spec)

    def calc(d):
        def f(ss):
            ss.resetFields()
            d["ex"].apply(ss)
            d["ani"].apply(ss)
            d["dip"].apply(ss)
            ss.sumFields()
        return f

    step = M.make_rk_step_function(g["ss"], kind, calc(g), g["llg"])
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], None, o["llg"])
    for _ in range(20):
        step()
    sim.rk(kind, 20)
    assert _max_dm(g, o) <= 1e-5
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-9


def test_pc_driver_on_op_api_equals_fused_heun():
    spec = W.FilmSpec(16, 16, 2, T=0.0, dt=0.01)
    a = W.build(M, spec)
    b = W.build(M, spec)
    a["ss"].randomize(3)
    b["ss"].randomize(3)

    def calc(ss):
        ss.resetFields()
        a["ex"].apply(ss)
        a["ani"].apply(ss)
        a["dip"].apply(ss)
        ss.sumFields()

    step = M.make_pc_step_function(a["ss"], calc, a["llg"], 1e9)     # one corrector pass = Heun
    heun = M.Heun(b["ss"], b["ex"], b["ani"], b["dip"], b["temp"], b["llg"])
    for _ in range(10):
        step()
    heun.step(10)
    assert float(np.abs(flat(a["ss"].spins()) - flat(b["ss"].spins())).max()) <= 1e-6
    assert abs(a["ss"].time() - b["ss"].time()) < 1e-12


# ------------------------------------------------------- row-tile shapes
ROW_TILES_256 = ["8:1024", "4:512", "8:512", "2:256", "4:1024"]


def _run_with_tile(spec, tile, steps, monkeypatch):
    if tile is None:
        monkeypatch.delenv("MSD_ROW_TILE", raising=False)
    else:
        monkeypatch.setenv("MSD_ROW_TILE", tile)
    g = W.build(M, spec)
    g["ss"].setSpins(*O.random_unit_spins(spec.seed, spec.n))
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    heun.step(steps)
    monkeypatch.delenv("MSD_ROW_TILE", raising=False)
    return flat(g["ss"].spins()), g


def test_row_tile_shapes_and_y_schedules(monkeypatch):
    """Full-column transforms: every compiled row-tile shape for nx=256 pairs the same
    real rows per complex line and runs the same butterflies per line, so all give
    bit-identical steps.  Split y transforms: shapes with the same rows per tile run the
    same radix-ty stage and the same sub-column transforms, so they agree bit for bit;
    every schedule is checked against the oracle."""
    spec = W.FilmSpec(256, 64, 2, T=0.3, dt=0.004, g=0.02, K=0.05)
    o = W.build(O, spec)
    o["ss"].setSpins(*O.random_unit_spins(spec.seed, spec.n))
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"], seed=0)
    sim.heun(12)
    want = flat(o["ss"].spins()[:3])
    full = [_run_with_tile(spec, t + ":0", 12, monkeypatch)[0] for t in ROW_TILES_256]
    for t, got in zip(ROW_TILES_256, full):
        print(f"row tile {t}, full columns: max|diff| = {np.abs(got - full[0]).max():.3e}", flush=True)
        assert np.array_equal(got, full[0]), t
    split = {t: _run_with_tile(spec, t + ":" + t.split(":")[0], 12, monkeypatch)[0] for t in ROW_TILES_256}
    ref, g = _run_with_tile(spec, None, 12, monkeypatch)
    assert g["dip"].schedule() == {"rows_per_tile": 8, "threads": 1024, "split": 0}
    assert np.array_equal(ref, full[0])
    assert np.array_equal(split["8:1024"], split["8:512"])
    assert np.array_equal(split["4:1024"], split["4:512"])
    for t, got in list(split.items()) + [("full", full[0])]:
        err = float(np.abs(got[:3] - want).max())
        print(f"row tile {t}, 256x64x2 T=0.3, 12 steps vs oracle: max|dm| = {err:.3e}", flush=True)
        assert err <= 2e-6        # measured 1e-7 .. 2e-7


def test_profile_leaves_state_one_step_ahead():
    """profile(n) repeats each launch of one step n times; the state, time and
    thermal draw end exactly one Heun step ahead."""
    spec = W.FilmSpec(64, 32, 2, T=0.5, dt=0.005)
    a = W.build(M, spec)
    b = W.build(M, spec)
    a["ss"].randomize(6)
    b["ss"].randomize(6)
    ha = M.Heun(a["ss"], a["ex"], a["ani"], a["dip"], a["temp"], a["llg"])
    hb = M.Heun(b["ss"], b["ex"], b["ani"], b["dip"], b["temp"], b["llg"])
    ha.step(2)
    prof = ha.profile(5)
    assert set(prof) == {"k_cols", "k_step_pred", "k_step_corr"}
    assert all(v > 0 for v in prof.values())
    ha.step(3)
    hb.step(6)
    assert np.array_equal(flat(a["ss"].spins()), flat(b["ss"].spins()))
    assert a["ss"].time() == b["ss"].time()
    assert a["temp"].draw() == b["temp"].draw() == 6


def test_autotune_picks_a_compiled_shape_and_leaves_the_trajectory():
    """The tuned driver steps bit-identically to a driver created with the chosen shape
    and schedule (MSD_ROW_TILE), and tuning leaves the state where it was."""
    spec = W.FilmSpec(256, 64, 2, T=0.3, dt=0.004, g=0.02, K=0.05)
    x, y, z = O.random_unit_spins(spec.seed, spec.n)
    g = W.build(M, spec)
    g["ss"].setSpins(x, y, z)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    before = flat(g["ss"].spins())
    best = heun.autotune(steps=4)
    print(f"autotune: {best}", flush=True)
    assert (best["rows_per_tile"], best["threads"]) in {(8, 1024), (4, 512), (8, 512), (2, 256), (4, 1024)}
    assert best["split"] in (0, best["rows_per_tile"])
    assert best["ms_per_step"] > 0
    assert g["temp"].draw() == 0 and g["ss"].time() == 0.0
    assert np.array_equal(flat(g["ss"].spins()), before)
    heun.step(8)
    os.environ["MSD_ROW_TILE"] = f'{best["rows_per_tile"]}:{best["threads"]}:{best["split"]}'
    try:
        h = W.build(M, spec)
    finally:
        del os.environ["MSD_ROW_TILE"]
    h["ss"].setSpins(x, y, z)
    M.Heun(h["ss"], h["ex"], h["ani"], h["dip"], h["temp"], h["llg"]).step(8)
    assert np.array_equal(flat(g["ss"].spins()), flat(h["ss"].spins()))
    assert g["ss"].time() == h["ss"].time()


# ------------------------------------------------------- LLG.Quaternion
@pytest.mark.parametrize("thermal_first", [True, False])
def test_quaternion_llg_op_matches_oracle(thermal_first):
    spec = W.FilmSpec(8, 8, 2, T=0.7, dt=0.03, alpha=0.3)
    g, o = twin(spec)
    ms = 0.5 + np.random.default_rng(4).random(spec.n)      # moment sizes other than 1 (|s| = ms)
    x, y, z = O.random_unit_spins(spec.seed, spec.n)
    for d in (g, o):
        d["ss"].setSpins(x * ms, y * ms, z * ms)
    lg, lo = M.LLG.Quaternion(), O.LLG(quaternion=True)
    lg.setThermalOnlyFirstTerm(thermal_first)
    lo.setFlags(thermal_only_first_term=thermal_first)
    for d, seed in ((g, None), (o, 5)):
        ss = d["ss"]
        ss.resetFields()
        d["ex"].apply(ss)
        d["ani"].apply(ss)
        d["dip"].apply(ss)
        if seed is None:
            d["temp"].setSeed(5)
            d["temp"].apply(ss)
        else:
            d["temp"].apply(ss, seed=5, step=0)
        ss.sumFields()
    P_g, P_o = g["ss"].copy(), o["ss"].copy()
    lg.apply(g["ss"], 0.5, g["ss"], P_g)
    lo.apply(o["ss"], 0.5, o["ss"], P_o)
    err = np.abs(flat(P_g.spins()) - flat(P_o.spins())).max()
    print(f"quaternion op (thermal_first={thermal_first}): max err {err:.2e}", flush=True)
    assert err < 2e-6
    assert abs(P_g.time() - 0.015) < 1e-15


def _oracle_pc_step(d, llg, calc):
    ss = d["ss"]
    P, Cc = d.setdefault("_P", ss.copy()), d.setdefault("_C", ss.copy())
    calc(ss)
    llg.apply(ss, 1.0, ss, P)
    llg.apply(ss, 0.5, ss, Cc)
    calc(P)
    llg.apply(Cc, 0.5, P, Cc)
    O.lib().ora_sys_copy_spins_to(Cc.h, ss.h)
    ss.setTime(Cc.time())


def test_quaternion_llg_pc_driver_matches_oracle():
    spec = W.FilmSpec(16, 16, 2, T=0.0, dt=0.01, alpha=0.2, g=0.02)
    g, o = twin(spec)

    def calc_of(d):
        def calc(ss):
            ss.resetFields()
            d["ex"].apply(ss)
            d["ani"].apply(ss)
            d["dip"].apply(ss)
            ss.sumFields()
        return calc

    lg, lo = M.LLG.new("Quaternion"), O.LLG(quaternion=True)
    step = M.make_pc_step_function(g["ss"], calc_of(g), lg, 1e9)
    for _ in range(30):
        step()
        _oracle_pc_step(o, lo, calc_of(o))
    err = _max_dm(g, o)
    print(f"quaternion PC driver, 30 steps: max|dm| {err:.2e}", flush=True)
    assert err <= 1e-5
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-12


@pytest.mark.parametrize("T,thermal_first,comp", [(0.0, True, True), (0.4, True, True), (0.4, False, False)])
def test_fused_heun_quaternion_matches_oracle(T, thermal_first, comp):
    """LLG.Quaternion in the fused Heun step (each turn taken as an increment, so the
    compensated state keeps it): 30 steps against the oracle's PredictorCorrector pass
    with LLGQuaternion::apply and the thermal kick, same counter-based noise.  The
    step is large enough (turns of ~0.1 rad) that LLG.Cartesian's trajectory lies far
    outside the tolerance, so the test tells the two integrators apart."""
    spec = W.FilmSpec(24, 16, 3, T=T, dt=0.02, alpha=0.3, K=0.1, g=0.05)
    g, o = twin(spec)
    c = W.build(O, spec)
    c["ss"].setSpins(*O.random_unit_spins(spec.seed, spec.n))
    lg, lo, lc = M.LLG.Quaternion(), O.LLG(quaternion=True), O.LLG()
    lg.setThermalOnlyFirstTerm(thermal_first)
    lo.setFlags(thermal_only_first_term=thermal_first)
    lc.setFlags(thermal_only_first_term=thermal_first)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], lg, compensated=comp)
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], lo, seed=0)
    cart = O.Sim(c["ss"], c["ex"], c["ani"], c["dip"], c["temp"], lc, seed=0)
    heun.step(30)
    sim.heun(30)
    cart.heun(30)
    err = _max_dm(g, o)
    apart = float(np.abs(flat(o["ss"].spins()[:3]) - flat(c["ss"].spins()[:3])).max())
    print(f"fused quaternion T={T} thermal_first={thermal_first} compensated={comp}: 30 steps "
          f"max|dm| {err:.2e} (Cartesian trajectory {apart:.2e} away)", flush=True)
    assert err <= 1e-5
    assert apart >= 100 * max(err, 1e-7)
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-12


@pytest.mark.parametrize("kind,T,dipole", [("RK2", 0.0, True), ("RK4", 0.3, True), ("K3", 0.0, False)])
def test_fused_rk_quaternion_matches_oracle(kind, T, dipole):
    """LLG.Quaternion in the fused Runge-Kutta stages: each chained sub-step turns the
    running state about the stage's axis by c dt (K3's c = -1 turns it back), against
    the oracle's RungeKutta.lua restatement with LLGQuaternion::apply, 20 steps."""
    spec = W.FilmSpec(32, 16, 2, T=T, dt=0.01)
    g, o = twin(spec, dipole=dipole)
    lg, lo = M.LLG.Quaternion(), O.LLG(quaternion=True)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], lg)
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], lo, seed=spec.noise_seed)
    heun.rk(kind, 20)
    sim.rk(kind, 20)
    err = _max_dm(g, o)
    print(f"fused quaternion {kind} T={T} dipole={dipole}: 20 steps max|dm| = {err:.3e}", flush=True)
    assert err <= 1e-5
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-12


def test_unseeded_thermals_get_distinct_streams():
    """An unseeded Thermal (or seed 0) is keyed from /dev/urandom, so two of them never
    share a noise stream; an RNG object or an explicit seed keys it reproducibly."""
    a, b, c = M.Thermal(4, 4, 1), M.Thermal(4, 4, 1), M.Thermal(4, 4, 1, seed=0)
    assert len({a.seed(), b.seed(), c.seed()}) == 3 and 0 not in (a.seed(), b.seed(), c.seed())
    assert M.Thermal(4, 4, 1, seed=77).seed() == 77
    assert M.Thermal(4, 4, 1, M.Random(91)).seed() == 91
    t = M.Thermal(4, 4, 1)
    t.setSeed(0, 5)
    assert t.seed() == 0 and t.draw() == 5


def test_ensemble_rejects_members_sharing_a_thermal():
    spec = W.FilmSpec(8, 8, 8, g=0.0, pbc=(1, 1, 1), T=0.5)
    g1 = W.build(M, spec, with_dipole=False)
    g2 = W.build(M, spec, with_dipole=False)
    h1 = M.Heun(g1["ss"], g1["ex"], g1["ani"], None, g1["temp"], g1["llg"])
    h2 = M.Heun(g2["ss"], g2["ex"], g2["ani"], None, g1["temp"], g2["llg"])
    with pytest.raises(M.MagLuaError, match="share a Thermal"):
        M.Ensemble([h1, h2])
    h3 = M.Heun(g2["ss"], g2["ex"], g2["ani"], None, g2["temp"], g2["llg"])
    M.Ensemble([h1, h3]).step(2)


def test_changing_temperature_and_dt_reuses_the_step_graph():
    """temp:set(T) and setTimeStep between steps (a temperature sweep script): the graph
    captured once replays the new values, and the steps equal direct launches bit for
    bit (so no stale value is replayed)."""
    spec = W.FilmSpec(32, 32, 2, T=0.4, dt=0.005)
    runs = []
    for graph in (True, False):
        g = W.build(M, spec)
        g["ss"].randomize(8)
        heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"], graph=graph)
        for k in range(12):
            g["temp"].setTemperature(0.4 + 0.1 * k)
            if k % 4 == 3:
                g["ss"].setTimeStep(0.005 * (1 + k / 10))
            heun.step(1)
        M.synchronize()
        runs.append((flat(g["ss"].spins()), g["ss"].time(), heun.graphBuilds()))
    print(f"graph builds over 12 steps with new T every step: {runs[0][2]}", flush=True)
    assert runs[0][2] == 1
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]


def test_span_graphs_take_new_values_and_match_direct_launches():
    """Runs of 16 steps and more replay a 16-step graph: with the temperature and the
    time step changed between runs, and run lengths that leave remainders for the
    one-step graph, the trajectory equals direct launches bit for bit and each graph
    is captured once."""
    spec = W.FilmSpec(64, 64, 2, T=0.3, dt=0.004)
    runs = []
    for graph in (True, False):
        g = W.build(M, spec)
        g["ss"].randomize(5)
        heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"], graph=graph)
        for k, n in enumerate((16, 37, 5, 48, 16)):
            g["temp"].setTemperature(0.3 + 0.2 * k)
            if k == 2:
                g["ss"].setTimeStep(0.005)
            heun.step(n)
        M.synchronize()
        runs.append((flat(g["ss"].spins()), g["ss"].time(), g["temp"].draw(), heun.graphBuilds()))
    print(f"span graphs: builds {runs[0][3]}, draws {runs[0][2]}", flush=True)
    assert runs[0][3] == 2                    # the 16-step graph and the one-step graph
    assert runs[0][2] == runs[1][2] == 16 + 37 + 5 + 48 + 16
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]


@pytest.mark.parametrize("kind,T,dipole", [("RK1", 0.0, True), ("RK2", 0.0, True), ("RK4", 0.0, True),
                                           ("BS3", 0.0, True), ("RK6", 0.0, False), ("RK4_38", 0.3, True),
                                           ("K3", 0.5, False)])
def test_fused_rk_driver_matches_oracle(kind, T, dipole):
    """The fused Runge-Kutta stages (one row kernel per stage, plus the column pass with
    a Dipole) against the oracle's RungeKutta.lua restatement, 20 steps; the thermal
    kick after the last stage on the same noise."""
    spec = W.FilmSpec(32, 16, 2, T=T, dt=0.01)
    g, o = twin(spec, dipole=dipole)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"], seed=spec.noise_seed)
    heun.rk(kind, 20)
    sim.rk(kind, 20)
    err = _max_dm(g, o)
    print(f"fused {kind} T={T} dipole={dipole}: 20 steps max|dm| = {err:.3e}", flush=True)
    assert err <= 1e-5
    assert abs(g["ss"].time() - o["ss"].time()) < 1e-12
    if T:
        assert g["temp"].draw() == 20


def test_fused_rk_graph_and_direct_launches_agree():
    spec = W.FilmSpec(32, 32, 2, T=0.4, dt=0.005)
    runs = []
    for graph in (True, False):
        g = W.build(M, spec)
        g["ss"].randomize(3)
        h = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"], graph=graph)
        h.rk("RK4", 9)
        h.step(3)        # Heun after RK: the same spin spectrum bookkeeping
        h.rk("BS3", 4)
        runs.append(flat(g["ss"].spins()))
    assert np.array_equal(runs[0], runs[1])


# ------------------------------------------------- split y-transform schedules
SPLIT_CASES = [((256, 256, 8), "8:1024:8"), ((256, 256, 8), "4:512:4"), ((256, 256, 8), "2:256:2"),
               ((512, 512, 4), "4:1024:4"), ((512, 512, 4), "2:512:2"), ((64, 64, 2), "8:256:8"),
               ((128, 128, 3), "8:512:8"), ((256, 128, 1), "4:1024:4")]


@pytest.mark.parametrize("dims,tile", SPLIT_CASES)
def test_split_schedule_dipole_against_oracle(dims, tile, monkeypatch):
    """Dipole:apply with the split y transform (the row kernels take a radix-ty stage
    across their tile's rows, the column kernel transforms sub-columns of ny/ty) against
    the FP64 oracle; sub-column lengths 8 .. 256, 1 to 8 layers."""
    monkeypatch.setenv("MSD_ROW_TILE", tile)
    nx, ny, nz = dims
    n = nx * ny * nz
    s = np.random.default_rng(nx + ny + nz).standard_normal((3, n))
    gs, os_ = M.SpinSystem(*dims), O.SpinSystem(*dims, threads=True)
    gs.setSpins(*s)
    os_.setSpins(*s)
    gd, od = M.Dipole(*dims), O.Dipole(*dims, threads=True)
    monkeypatch.delenv("MSD_ROW_TILE")
    ty = int(tile.split(":")[0])
    assert gd.schedule()["split"] == ty
    for d in (gd, od):
        d.setStrength(0.3)
        d.setTruncation(40)
    gd.apply(gs)
    od.apply(os_)
    hg, ho = flat(gs.field("Dipole")), flat(os_.field("Dipole"))
    err = float(np.abs(hg - ho).max() / np.abs(ho).max())
    print(f"split {tile} {dims}: Dipole:apply max rel err = {err:.3e}", flush=True)
    assert err <= 2e-6


@pytest.mark.parametrize("tile", ["8:256:8", "8:256:0"])
def test_split_schedule_heun_and_rk_against_oracle(tile, monkeypatch):
    """Fused Heun (T > 0, same noise) and fused RK4 stages on the split schedule against
    the oracle, 30 + 10 steps; the full-column schedule of the same film beside it."""
    spec = W.FilmSpec(64, 64, 2, T=0.3, dt=0.005, g=0.05)
    monkeypatch.setenv("MSD_ROW_TILE", tile)
    g, o = twin(spec)
    monkeypatch.delenv("MSD_ROW_TILE")
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"], seed=spec.noise_seed)
    heun.step(30)
    sim.heun(30)
    heun.rk("RK4", 10)
    sim.rk("RK4", 10)
    err = _max_dm(g, o)
    print(f"schedule {tile}: 30 Heun + 10 RK4 steps max|dm| = {err:.3e}", flush=True)
    assert err <= 2e-5
