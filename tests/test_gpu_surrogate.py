"""GPU parity of the dipolar-field surrogate (BF16 MFMA kernels through the C ABI)
against the FP64 network oracle (oracle/cnn_oracle.py), which rounds to BF16 at
the same places; then training, the gain over a frozen field, and the learned
step driver with its exact fallback."""
import numpy as np
import pytest

import maglua_amd as M
from maglua_amd import surrogate as S
from maglua_amd import workloads as W
from oracle import cnn_oracle as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def soa(a):
    """[nz,ny,nx,3] -> [3][n]"""
    return np.moveaxis(a, -1, 0).reshape(3, -1)


def aos(v, dims):
    nx, ny, nz = dims
    return np.moveaxis(np.asarray(v).reshape(3, nz, ny, nx), 0, -1)


def setup(dims, seed, s_in=60.0, s_out=0.04, sigma2=2e-4, periodic=True):
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    Ws = N.init_weights(seed)
    dm = rng.standard_normal((nz, ny, nx, 3)) * 0.01
    T = rng.standard_normal((nz, ny, nx, 3)) * 0.02
    sur = M.Surrogate(nx, ny, nz, periodic=periodic)
    for l in range(3):
        sur.setWeights(l, Ws[l])
    sur.setScales(s_in, s_out, sigma2)
    sur.setInput(soa(dm))
    sur.setTarget(soa(T))
    return sur, Ws, dm, T, (s_in, s_out, sigma2)


def _forward_errors(got, want):
    """Max and median |got - want| relative to max|want|.  The oracle rounds to BF16
    where the kernels do, so outputs agree to FP32 accumulation order (median measured
    5e-8 .. 1e-7); an intermediate activation that lies on a BF16 rounding boundary can
    round the other way, one BF16 ulp (2^-8 relative) that spreads through the later
    layers' 5x5x3 neighbourhoods (max measured up to 1.6e-4 on a 12-layer lattice)."""
    d = np.abs(got - want).reshape(-1)
    scale = np.abs(want).max()
    return float(d.max() / scale), float(np.median(d) / scale)


# film walk (every layer in one item) for 1, 2, 4 and 8 layers; the per-layer walk
# for 3 and 5
@pytest.mark.parametrize("dims", [(32, 4, 1), (32, 8, 4), (64, 6, 3), (64, 5, 8), (32, 7, 5), (96, 3, 2)])
def test_forward_matches_oracle(dims):
    sur, Ws, dm, T, (s_in, s_out, _) = setup(dims, 1)
    got = aos(sur.forward(), dims)
    want, _ = N.forward(Ws, dm, s_in, s_out, bf=True)
    err, bulk = _forward_errors(got, want)
    print(f"surrogate forward {dims}: rel err max {err:.2e}, median {bulk:.2e}", flush=True)
    assert bulk <= 2e-7 and err <= 1e-3      # measured: median 1.6e-8, max 1.6e-4 (a BF16 rounding tie)


# zero padding in x and y (open lattices) runs the per-layer walk at every depth
@pytest.mark.parametrize("dims", [(32, 4, 1), (32, 8, 4), (64, 6, 3), (96, 5, 2), (32, 9, 12)])
def test_open_forward_matches_oracle(dims):
    sur, Ws, dm, T, (s_in, s_out, _) = setup(dims, 11, periodic=False)
    got = aos(sur.forward(), dims)
    want, _ = N.forward(Ws, dm, s_in, s_out, bf=True, periodic=False)
    err, bulk = _forward_errors(got, want)
    print(f"open surrogate forward {dims}: rel err max {err:.2e}, median {bulk:.2e}", flush=True)
    assert bulk <= 2e-7 and err <= 1e-3      # measured: median 1.6e-8, max 1.6e-4 (a BF16 rounding tie)
    # the periodic network differs near the x and y faces: the test sees the padding
    wrap, _ = N.forward(Ws, dm, s_in, s_out, bf=True)
    assert np.abs(wrap - want).max() > 10 * np.abs(got - want).max()


@pytest.mark.parametrize("dims", [(32, 8, 4), (32, 6, 3)])
def test_open_gradients_match_oracle(dims):
    sur, Ws, dm, T, sc = setup(dims, 13, periodic=False)
    loss, grads = sur.lossGrads()
    L, want = N.loss_and_grads(Ws, dm, T, *sc, bf=True, periodic=False)
    print(f"open surrogate loss {dims}: gpu {loss:.6e} oracle {L:.6e}", flush=True)
    assert abs(loss - L) <= 1e-3 * L
    for l in range(3):
        rel = np.linalg.norm(grads[l] - want[l]) / np.linalg.norm(want[l])
        print(f"  layer {l}: grad rel err {rel:.2e}", flush=True)
        assert rel <= 1e-2, l


def test_weights_round_trip():
    sur, Ws, *_ = setup((32, 4, 2), 2)
    for l in range(3):
        assert np.allclose(sur.weights(l), np.asarray(Ws[l], dtype=np.float32), rtol=0, atol=0)


@pytest.mark.parametrize("dims", [(32, 8, 4), (64, 4, 2), (32, 5, 8), (32, 6, 3)])
def test_gradients_match_oracle(dims):
    sur, Ws, dm, T, sc = setup(dims, 3)
    loss, grads = sur.lossGrads()
    L, want = N.loss_and_grads(Ws, dm, T, *sc, bf=True)
    print(f"surrogate loss {dims}: gpu {loss:.6e} oracle {L:.6e}", flush=True)
    assert abs(loss - L) <= 1e-3 * L
    for l in range(3):
        rel = np.linalg.norm(grads[l] - want[l]) / np.linalg.norm(want[l])
        print(f"  layer {l}: grad rel err {rel:.2e}", flush=True)
        assert rel <= 1e-2, l


def test_adam_step_matches_oracle():
    sur, Ws, dm, T, sc = setup((32, 4, 2), 4)
    _, g = sur.lossGrads()
    w0 = [sur.weights(l) for l in range(3)]
    sur.adam(1e-3)
    for l in range(3):
        w = w0[l].copy()
        N.adam(w, g[l], np.zeros_like(w), np.zeros_like(w), 1, lr=1e-3)
        assert np.abs(sur.weights(l) - w).max() <= 1e-6


def test_training_fits_a_teacher_network():
    """A target made by a known network of the same form is learned: the loss falls 20x."""
    dims = (32, 8, 2)
    nx, ny, nz = dims
    rng = np.random.default_rng(5)
    teacher = N.init_weights(50)
    dm = rng.standard_normal((nz, ny, nx, 3)) * 0.01
    T, _ = N.forward(teacher, dm, 100.0, 0.05)
    sur = M.Surrogate(*dims)
    for l, Wl in enumerate(N.init_weights(7)):
        sur.setWeights(l, Wl)
    sur.setScales(100.0, 0.05, float(np.mean(T * T)))
    sur.setInput(soa(dm))
    sur.setTarget(soa(T))
    first = sur.trainStep(2e-3)
    for _ in range(300):
        last = sur.trainStep(2e-3)
    print(f"teacher fit: loss {first:.3e} -> {last:.3e}", flush=True)
    assert last <= first / 20


def test_exact_target_is_the_dipole_field_of_the_difference():
    spec = W.FilmSpec(32, 32, 2, g=0.1, K=0.02)
    g = W.build(M, spec)
    x, y, z = O.random_unit_spins(3, spec.n)
    g["ss"].setSpins(x, y, z)
    ref = g["ss"].copy()
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    heun.step(5)
    sur = M.Surrogate(32, 32, 2)
    sur.setTargetDipole(g["dip"], g["ss"], ref)
    # the target read back through a zero network: loss = mean(T^2) / sigma2
    sur.setScales(1.0, 1.0, 1.0)
    sur.setInput(np.zeros((3, spec.n)))
    loss, _ = sur.lossGrads()
    o = W.build(O, spec)
    a, b = g["ss"].spins(), ref.spins()
    o["ss"].setSpins(*(np.asarray(a[c]) - np.asarray(b[c]) for c in range(3)))
    o["dip"].apply(o["ss"])
    T = np.stack([np.asarray(v).reshape(-1) for v in o["ss"].field("Dipole")])
    assert abs(loss - np.mean(T * T)) <= 1e-3 * np.mean(T * T)


def film(seed, nx=64, ny=64, nz=4, T=0.2):
    spec = W.FilmSpec(nx, ny, nz, J=1.0, K=0.02, g=0.1, T=T, alpha=0.1, dt=0.005, seed=seed)
    g = W.build(M, spec)
    g["ss"].setSpins(*O.random_unit_spins(seed, spec.n))
    if g["temp"] is not None:
        g["temp"].setSeed(seed)
    return spec, g


def test_trained_surrogate_beats_the_frozen_field():
    """After training on one trajectory, on another the network's field K=10 steps after a
    refresh is closer to the exact field than the refreshed field held fixed."""
    spec, g = film(11)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    heun.step(20)
    sur = M.Surrogate(spec.nx, spec.ny, spec.nz)
    log = S.train_on_trajectory(sur, g["ss"], heun, g["dip"], steps=400, K=10, lr=2e-3, seed=1)
    print(f"training loss {log[0]:.3e} -> {np.mean(log[-20:]):.3e}", flush=True)
    spec2, h = film(12)
    heun2 = M.Heun(h["ss"], h["ex"], h["ani"], h["dip"], h["temp"], h["llg"])
    heun2.step(20)
    ref = h["ss"].copy()
    heun2.step(10)
    rel, frozen = sur.fieldError(h["dip"], h["ss"], ref)
    print(f"field error 10 steps after refresh: surrogate {rel:.3e}, frozen {frozen:.3e}", flush=True)
    assert rel < 0.5 * frozen


def test_learned_driver_tracks_the_exact_trajectory_and_falls_back():
    spec, g = film(21)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    heun.step(20)
    sur = M.Surrogate(spec.nx, spec.ny, spec.nz)
    S.train_on_trajectory(sur, g["ss"], heun, g["dip"], steps=300, K=10, lr=2e-3, seed=2)
    # two identical states, one driven exactly, one with the surrogate
    a = g["ss"].copy()
    b = g["ss"].copy()
    ea = M.Heun(a, g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    tb = W.build(M, spec)["temp"]
    tb.setSeed(spec.seed)
    ta = g["temp"]
    for t in (ta, tb):
        t.setSeed(21, draw=ta.draw())
    eb = M.Heun(b, g["ex"], g["ani"], g["dip"], tb, g["llg"])
    eb.setSurrogate(sur, K=10, tol=1e-2)
    ea.step(40)
    eb.step(40)
    st = eb.surrogateStats()
    dmax = float(np.abs(np.stack(a.spins()[:3]) - np.stack(b.spins()[:3])).max())
    print(f"learned vs exact after 40 steps: max|dm| {dmax:.3e}; {st}", flush=True)
    assert st["windows"] == 4
    assert dmax <= 1e-3
    assert abs(a.time() - b.time()) < 1e-12
    # an impossible tolerance: every window, the first included, fails its check and
    # runs again as exact steps from its restored state and noise
    eb.setSurrogate(sur, K=10, tol=1e-12)
    eb.step(30)
    st = eb.surrogateStats()
    assert st["windows"] == 3 and st["fallbacks"] == 3 and st["exact_windows"] == 3


@pytest.mark.parametrize("quat", [False, True])
def test_learned_driver_with_zero_tolerance_is_the_exact_driver(quat):
    """tol = 0: every window is checked, restored and run exactly, so the trajectory
    equals the exact driver's bit for bit (same state, noise and draw counter), and the
    time and draw counter advance as one exact run's do; with LLG.Cartesian and with
    LLG.Quaternion."""
    spec, g = film(23, T=0.3)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], g["dip"], g["temp"], g["llg"])
    heun.step(10)
    sur = M.Surrogate(spec.nx, spec.ny, spec.nz)
    S.train_on_trajectory(sur, g["ss"], heun, g["dip"], steps=40, K=5, lr=2e-3, seed=4)
    a, b = g["ss"].copy(), g["ss"].copy()
    ta, tb = W.build(M, spec)["temp"], W.build(M, spec)["temp"]
    for t in (ta, tb):
        t.setSeed(77, draw=3)
    llg = M.LLG.Quaternion() if quat else g["llg"]
    ea = M.Heun(a, g["ex"], g["ani"], g["dip"], ta, llg)
    eb = M.Heun(b, g["ex"], g["ani"], g["dip"], tb, llg)
    eb.setSurrogate(sur, K=7, tol=0.0)
    ea.step(30)
    eb.step(30)
    st = eb.surrogateStats()
    print(f"tol=0 learned driver: {st}", flush=True)
    assert st["windows"] == 5 and st["exact_windows"] == 5
    assert np.array_equal(np.stack(a.spins()), np.stack(b.spins()))
    assert a.time() == b.time() and ta.draw() == tb.draw() == 33


def open_cube(seed, n=32, T=0.3):
    spec = W.FilmSpec(n, n, n, J=1.0, K=0.05, g=0.0, pbc=(0, 0, 0), alpha=0.1, dt=0.002, T=T, seed=seed)
    g = W.build(M, spec)
    g["ss"].setSpins(*O.random_unit_spins(seed, spec.n))
    g["temp"].setSeed(seed)
    lr = M.LongRange3D(n, n, n, open=True)
    lr.setStrength(0.05)
    return spec, g, lr


def test_learned_driver_on_an_open_longrange3d_cube():
    """The surrogate over the open 3D convolution, zero-padded in x and y: trained on a
    trajectory it beats the frozen field, the learned driver tracks the exact trajectory
    from the same state and noise, and an impossible tolerance falls back to exact steps
    that continue the exact trajectory."""
    spec, g, lr = open_cube(31)
    heun = M.Heun(g["ss"], g["ex"], g["ani"], lr, g["temp"], g["llg"])
    heun.step(20)
    sur = M.Surrogate(spec.nx, spec.ny, spec.nz, periodic=False)
    log = S.train_on_trajectory(sur, g["ss"], heun, lr, steps=200, K=8, lr=2e-3, seed=3)
    print(f"open cube training loss {log[0]:.3e} -> {np.mean(log[-20:]):.3e}", flush=True)
    ref = g["ss"].copy()
    heun.step(8)
    rel, frozen = sur.fieldError(lr, g["ss"], ref)
    print(f"open cube field error 8 steps after refresh: surrogate {rel:.3e}, frozen {frozen:.3e}", flush=True)
    assert rel < 0.5 * frozen
    a, b = g["ss"].copy(), g["ss"].copy()
    ta, tb = W.build(M, spec)["temp"], W.build(M, spec)["temp"]
    ta.setSeed(5)
    tb.setSeed(5)
    ea = M.Heun(a, g["ex"], g["ani"], lr, ta, g["llg"])
    eb = M.Heun(b, g["ex"], g["ani"], lr, tb, g["llg"])
    eb.setSurrogate(sur, K=8, tol=1e-1)
    ea.step(32)
    eb.step(32)
    st = eb.surrogateStats()
    dmax = float(np.abs(np.stack(a.spins()[:3]) - np.stack(b.spins()[:3])).max())
    print(f"open cube learned vs exact after 32 steps: max|dm| {dmax:.3e}; {st}", flush=True)
    assert st["windows"] == 4 and st["exact_windows"] == 0
    assert dmax <= 1e-3
    # an impossible tolerance: every window fails its check and runs again exactly from
    # its restored state, continuing exactly as an exact driver does
    eb.setSurrogate(sur, K=8, tol=1e-12)
    eb.step(8)
    c = b.copy()
    tc = W.build(M, spec)["temp"]
    tc.setSeed(5, draw=tb.draw())
    ec = M.Heun(c, g["ex"], g["ani"], lr, tc, g["llg"])
    eb.step(16)
    ec.step(16)
    st = eb.surrogateStats()
    assert st["windows"] == 3 and st["exact_windows"] == 3
    assert np.array_equal(np.stack(b.spins()), np.stack(c.spins()))
