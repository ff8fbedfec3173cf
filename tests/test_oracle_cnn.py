"""Pins the FP64 surrogate-network oracle (oracle/cnn_oracle.py): the layer
against a brute-force loop, the transposed layer as its adjoint, weight
gradients against central finite differences, and the Adam step against its
closed form on a quadratic."""
import numpy as np
import pytest

from oracle import cnn_oracle as N


def brute_layer(X, W, periodic=True):
    nz, ny, nx, _ = X.shape
    Y = np.zeros((nz, ny, nx, W.shape[3]))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                for dz in range(-1, 2):
                    zz = z + dz
                    if zz < 0 or zz >= nz:
                        continue
                    for dy in range(-2, 3):
                        for dx in range(-2, 3):
                            yy, xx = y + dy, x + dx
                            if not periodic and not (0 <= yy < ny and 0 <= xx < nx):
                                continue
                            Y[z, y, x] += W[dz + 1, dy + 2, dx + 2] @ X[zz, yy % ny, xx % nx]
    return Y


def test_layer_matches_brute_force():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3, 4, 6, 5))
    W = rng.standard_normal((3, 5, 5, 2, 5))
    assert np.allclose(N.layer(X, W), brute_layer(X, W), atol=1e-12)
    assert np.allclose(N.layer(X, W, periodic=False), brute_layer(X, W, periodic=False), atol=1e-12)


def test_open_layer_has_no_wrap():
    """periodic=False: a tap reaching past the x or y edge reads zero, not the far side."""
    rng = np.random.default_rng(6)
    X = rng.standard_normal((2, 5, 7, 1))
    W = np.zeros((3, 5, 5, 1, 1))
    W[1, 2, 0] = 1.0           # dz=0, dy=0, dx=-2
    Y = N.layer(X, W, periodic=False)
    assert np.allclose(Y[:, :, :2], 0.0)
    assert np.allclose(Y[:, :, 2:], X[:, :, :5])
    W[:] = 0.0
    W[1, 4, 2] = 1.0           # dy=+2
    Y = N.layer(X, W, periodic=False)
    assert np.allclose(Y[:, 3:], 0.0)
    assert np.allclose(Y[:, :3], X[:, 2:])


def test_single_tap_is_a_shift():
    """A weight that is 1 at tap (dz,dy,dx) only gives Y(s) = X(s + d) (catches a sign flip)."""
    rng = np.random.default_rng(2)
    X = rng.standard_normal((4, 5, 7, 1))
    W = np.zeros((3, 5, 5, 1, 1))
    W[2, 4, 1] = 1.0           # dz=+1, dy=+2, dx=-1
    Y = N.layer(X, W)
    assert np.allclose(Y[1, 1, 3, 0], X[2, 3, 2, 0])
    assert np.allclose(Y[3], 0.0)                      # z+1 outside the lattice
    assert np.allclose(Y[0, 4, 0, 0], X[1, 1, 6, 0])   # y and x wrap


@pytest.mark.parametrize("periodic", [True, False])
def test_transpose_is_the_adjoint(periodic):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((4, 6, 8, 3))
    G = rng.standard_normal((4, 6, 8, 5))
    W = rng.standard_normal((3, 5, 5, 5, 3))
    lhs = np.sum(N.layer(X, W, periodic) * G)
    rhs = np.sum(X * N.layer_t(G, W, periodic))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


@pytest.mark.parametrize("periodic", [True, False])
def test_weight_gradients_match_finite_differences(periodic):
    rng = np.random.default_rng(4)
    nz, ny, nx = 3, 6, 8
    Ws = N.init_weights(5)
    dm = rng.standard_normal((nz, ny, nx, 3)) * 0.01
    T = rng.standard_normal((nz, ny, nx, 3)) * 0.02
    args = (dm, T, 80.0, 0.05, 3e-4)
    L, grads = N.loss_and_grads(Ws, *args, periodic=periodic)
    h = 1e-6
    for l in range(3):
        for _ in range(6):
            idx = tuple(int(rng.integers(0, s)) for s in Ws[l].shape)
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (N.loss_and_grads(Wp, *args, periodic=periodic)[0] -
                  N.loss_and_grads(Wm, *args, periodic=periodic)[0]) / (2 * h)
            assert abs(fd - grads[l][idx]) <= 1e-6 * max(1.0, abs(fd)), (l, idx, fd, grads[l][idx])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5e-3, 3.0e38])
    r = N.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie -> even
    assert r[2] == 1.0 + 2 ** -6         # tie -> even (up)
    assert abs(r[3] - -2.5e-3) <= 2.5e-3 * 2 ** -8
    assert np.isfinite(r[4])


def test_adam_converges_on_a_quadratic():
    W = np.array([3.0, -2.0])
    m = np.zeros(2)
    v = np.zeros(2)
    for t in range(1, 2001):
        N.adam(W, 2 * W, m, v, t, lr=0.05)
    assert np.abs(W).max() < 1e-3
    # first step moves every weight by lr against the gradient sign
    W = np.array([3.0, -2.0])
    N.adam(W, 2 * W, np.zeros(2), np.zeros(2), 1, lr=0.1)
    assert np.allclose(W, [2.9, -1.9], atol=1e-6)
