"""Training workflow of the dipolar-field surrogate (maglua.Surrogate).

Pairs come from a trajectory of the exact method: within windows of K steps that
start at a reference state m_ref, every step gives the input m - m_ref and the
target H_dip(m) - H_dip(m_ref), which is the exact field of m - m_ref (the field
is linear in m) from the layered Dipole or the 3D LongRange3D operator.  Every pair is one Adam step of the
device-side forward/backward kernels; nothing leaves the GPU but the loss.
"""
import numpy as np

CHANNELS = (3, 16, 16, 3)


def init_weights(seed, gain=1.0):
    """Seeded Gaussian weights [3,5,5,cout,cin] with variance gain / fan_in per layer."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(3):
        ci, co = CHANNELS[l], CHANNELS[l + 1]
        out.append(rng.standard_normal((3, 5, 5, co, ci)) * np.sqrt(gain / (75 * ci)))
    return out


def fit_scales(sur, ss, ref, dip):
    """Set s_in = 1/rms(m - m_ref), s_out = rms(target), sigma2 = s_out^2 from one pair,
    so the network sees O(1) values and the loss is the relative mean-square error."""
    a, b = ss.spins(), ref.spins()
    dm = np.stack([np.asarray(a[c]) - np.asarray(b[c]) for c in range(3)])
    rms_in = float(np.sqrt(np.mean(dm * dm)))
    sur.setScales(1.0, 1.0, 1.0)
    sur.setTargetDipole(dip, ss, ref)
    sur.setInput(np.zeros((3, dm[0].size)))
    mean_t2, _ = sur.lossGrads()          # zero network: loss = mean(T^2)
    rms_out = float(np.sqrt(mean_t2))
    if rms_in <= 0 or rms_out <= 0:
        raise ValueError("fit_scales: the trajectory did not move")
    sur.setScales(1.0 / rms_in, rms_out, rms_out * rms_out)
    return 1.0 / rms_in, rms_out


def train_on_trajectory(sur, ss, heun, dip, steps, K=10, lr=2e-3, seed=0, weights=None, log_every=0, resume=False):
    """Advance `heun` (an exact Dipole or LongRange3D step driver on `ss`) `steps` steps, training `sur`
    on every step's (m - m_ref, H(m) - H(m_ref)) pair, m_ref the state at the start of
    the current K-step window.  resume=True continues from the network's current weights,
    scales and optimizer state (another trajectory of the same training run); otherwise
    the weights start from `weights` or init_weights(seed) and the scales are fitted
    on the first pair.  Returns the list of losses."""
    if not resume:
        for l, Wl in enumerate(weights if weights is not None else init_weights(seed)):
            sur.setWeights(l, Wl)
        sur.resetOptimizer()
    ref = ss.copy()
    losses = []
    scaled = resume
    for k in range(steps):
        if k % K == 0:
            ss.copySpinsTo(ref)
        heun.step(1)
        if not scaled:
            fit_scales(sur, ss, ref, dip)
            scaled = True
        sur.setTargetDipole(dip, ss, ref)
        sur.setInputSystems(ss, ref)
        losses.append(sur.trainStep(lr))
        if log_every and (k + 1) % log_every == 0:
            print(f"# surrogate step {k + 1}: loss {np.mean(losses[-log_every:]):.3e}", flush=True)
    return losses
