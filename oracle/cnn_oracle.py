"""FP64 numpy restatement of the dipolar-field surrogate network (test
infrastructure only: imported by tests/ as the checker of the GPU kernels).

The network has no counterpart in the reference; it is the learned part the
build adds beside the exact layered convolution (LongRange::apply,
cpu/longrange/spinoperationlongrange.cpp:354-430), predicting the change of that
field between exact refreshes.  Conventions shared with the GPU kernels:

* lattice arrays are [nz][ny][nx][channels] (site x + nx (y + ny z), channels last);
* a layer is Y[z,y,x,o] = sum_{dz,dy,dx,i} W[dz+1,dy+2,dx+2,o,i] X[z+dz,(y+dy)%ny,(x+dx)%nx,i],
  dz in -1..1, dy and dx in -2..2: periodic in x and y, zero outside 0 <= z < nz
  (periodic=False: zero outside the lattice in x and y as well);
* the network is X0 = s_in dm (3 channels), Y1 = L1(X0), Y2 = L2(Y1), Y3 = L3(Y2)
  (16, 16, 3 channels, no bias, identity activation: the target map is linear),
  dH = s_out Y3;
* the loss is mean((s_out Y3 - T)^2) / sigma2 over sites and the 3 components.

With bf16=True the activations X0, Y1, Y2, the weights and the back-propagated
dY1, dY2, dY3 are rounded to bfloat16 (round to nearest even) at the same places
as the GPU kernels, which accumulate in FP32; everything else is FP64.

How this file is pinned (tests/test_oracle_cnn.py): the layer against a
brute-force loop over sites and taps, the transposed layer as its exact adjoint
(<L X, Y> = <X, L^T Y>), and every weight gradient of the loss against central
finite differences in FP64.
"""
import numpy as np

KZ, KY, KX = 3, 5, 5
TAPS = KZ * KY * KX
CHANNELS = (3, 16, 16, 3)


def bf16(a):
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(X, dz, dy, dx, periodic=True):
    """S[z,y,x] = X[z+dz, (y+dy)%ny, (x+dx)%nx], zero where z+dz is outside the lattice
    (periodic=False: zero where y+dy or x+dx is outside too)."""
    nz, ny, nx = X.shape[:3]
    if periodic:
        S = np.roll(X, shift=(-dy, -dx), axis=(1, 2))
    else:
        S = np.zeros_like(X)
        y0, y1 = max(0, -dy), min(ny, ny - dy)
        x0, x1 = max(0, -dx), min(nx, nx - dx)
        if y1 > y0 and x1 > x0:
            S[:, y0:y1, x0:x1] = X[:, y0 + dy:y1 + dy, x0 + dx:x1 + dx]
    out = np.zeros_like(S)
    z0, z1 = max(0, -dz), min(nz, nz - dz)
    if z1 > z0:
        out[z0:z1] = S[z0 + dz:z1 + dz]
    return out


def taps():
    for dz in range(-1, 2):
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                yield dz, dy, dx


def layer(X, W, periodic=True):
    """Y = L_W(X); W is [KZ,KY,KX,cout,cin]."""
    Y = np.zeros(X.shape[:3] + (W.shape[3],))
    for dz, dy, dx in taps():
        Y += shift(X, dz, dy, dx, periodic) @ W[dz + 1, dy + 2, dx + 2].T
    return Y


def layer_t(G, W, periodic=True):
    """Adjoint of layer(., W): X-gradient from the output gradient G."""
    X = np.zeros(G.shape[:3] + (W.shape[4],))
    for dz, dy, dx in taps():
        X += shift(G @ W[dz + 1, dy + 2, dx + 2], -dz, -dy, -dx, periodic)
    return X


def layer_w(X, G, periodic=True):
    """dW[t,o,i] = sum_sites G[s,o] X[s+t,i] (the weight gradient of layer)."""
    dW = np.zeros((KZ, KY, KX, G.shape[3], X.shape[3]))
    g = G.reshape(-1, G.shape[3])
    for dz, dy, dx in taps():
        dW[dz + 1, dy + 2, dx + 2] = g.T @ shift(X, dz, dy, dx, periodic).reshape(-1, X.shape[3])
    return dW


def forward(Ws, dm, s_in, s_out, bf=False, periodic=True):
    """dm [nz,ny,nx,3] -> (dH, cache)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    Wq = [r(W) for W in Ws]
    X0 = r(s_in * np.asarray(dm, dtype=np.float64))
    Y1 = layer(X0, Wq[0], periodic)
    X1 = r(Y1)
    Y2 = layer(X1, Wq[1], periodic)
    X2 = r(Y2)
    Y3 = layer(X2, Wq[2], periodic)
    return s_out * Y3, dict(Wq=Wq, X=(X0, X1, X2), Y3=Y3)


def loss_and_grads(Ws, dm, target, s_in, s_out, sigma2, bf=False, periodic=True):
    """Loss and its gradients with respect to the three weight tensors."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    dH, c = forward(Ws, dm, s_in, s_out, bf, periodic)
    R = dH - target
    count = R.size
    L = float(np.sum(R * R) / (count * sigma2))
    G3 = r(2.0 * s_out * R / (count * sigma2))
    X0, X1, X2 = c["X"]
    Wq = c["Wq"]
    G2 = r(layer_t(G3, Wq[2], periodic))
    G1 = r(layer_t(G2, Wq[1], periodic))
    grads = [layer_w(X0, G1, periodic), layer_w(X1, G2, periodic), layer_w(X2, G3, periodic)]
    return L, grads


def init_weights(seed, gain=1.0):
    """Seeded weights, variance gain / fan_in per layer."""
    rng = np.random.default_rng(seed)
    Ws = []
    for l in range(3):
        ci, co = CHANNELS[l], CHANNELS[l + 1]
        Ws.append(rng.standard_normal((KZ, KY, KX, co, ci)) * np.sqrt(gain / (TAPS * ci)))
    return Ws


def adam(W, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) of W in place; t counts from 1."""
    m *= b1
    m += (1 - b1) * g
    v *= b2
    v += (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    W -= lr * mh / (np.sqrt(vh) + eps)
