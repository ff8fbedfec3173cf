"""Writes film32_heun.npz: the FP64 oracle trajectory of the 32x32x1 film
(J=1, K=0.1 along z, layered dipole g=0.01 at truncation 1000, alpha=0.1,
gamma=1, Heun dt=0.005, T=0) from the seed-0 sphere-uniform state, at steps
1, 10, 100 and 1000.  Run: python tests/golden/make_film32_heun.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from maglua_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

spec = W.FILM_32
o = W.build(O, spec)
x0, y0, z0 = O.random_unit_spins(spec.seed, spec.n)
o["ss"].setSpins(x0, y0, z0)
sim = O.Sim(o["ss"], o["ex"], o["ani"], o["dip"], o["temp"], o["llg"])
steps = [1, 10, 100, 1000]
states, done = [], 0
for s in steps:
    sim.heun(s - done)
    done = s
    states.append(np.stack(o["ss"].spins()[:3]))
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "film32_heun.npz"),
                    x0=x0, y0=y0, z0=z0, steps=np.array(steps), states=np.array(states))
print("wrote film32_heun.npz", [float(np.abs(s).max()) for s in states])
