"""Host logic of make_adapt_step_function (common/scripts/AdaptiveTimeStep.lua) on a
scalar stand-in for a SpinSystem and an Euler step of dy/dt = -y, whose full and
half steps have a known difference y dt^2 / 4.  No GPU calls."""
import maglua_amd as M


class Scalar:
    def __init__(self, y, dt, t=0.0):
        self.y, self.dt, self.t = y, dt, t

    def copy(self):
        return Scalar(self.y, self.dt, self.t)

    def copySpinsTo(self, o):
        o.y = self.y

    def setTime(self, t):
        self.t = t

    def time(self):
        return self.t

    def setTimeStep(self, dt):
        self.dt = dt

    def timeStep(self):
        return self.dt

    def diff(self, o):
        d = abs(self.y - o.y)
        return (d, 0.0, 0.0, d)


def euler(s, skip_temperature=False):
    s.y *= 1.0 - s.dt
    s.t += s.dt


def test_accepted_step_keeps_two_half_steps_and_lengthens():
    s = Scalar(1.0, 0.1)
    step = M.make_adapt_step_function(s, euler, 1e-2)      # arguments in the Lua's order
    assert step() is True
    assert abs(s.y - 0.95 ** 2) < 1e-15 and abs(s.time() - 0.1) < 1e-15
    d = 0.1 ** 2 / 4                                       # |(1 - dt) - (1 - dt/2)^2|
    ee = min(max(((1e-2 - d) / 1e-2) ** 2, 0.1), 0.5)
    assert abs(s.timeStep() - 0.1 * (1 + ee - 0.1)) < 1e-15


def test_rejected_step_keeps_the_state_and_halves():
    s = Scalar(1.0, 0.1, t=3.0)
    step = M.make_adapt_step_function(1e-4, euler, s)      # any order: sorted by kind
    assert step() is False
    assert s.y == 1.0 and s.time() == 3.0 and s.timeStep() == 0.05


def test_adaptive_run_reaches_the_tolerance_region():
    """Repeated calls shrink the step until it is accepted, then integrate with the
    local error held under the tolerance: y(2) = e^-2 to about the tolerance."""
    import math
    s = Scalar(1.0, 0.5)
    step = M.make_adapt_step_function(s, euler, 1e-4)
    n = 0
    while s.time() < 2.0 - 1e-12 and n < 10000:
        if s.time() + s.timeStep() > 2.0:
            s.setTimeStep(2.0 - s.time())
        step()
        n += 1
    assert abs(s.time() - 2.0) < 1e-9
    assert abs(s.y - math.exp(-2.0)) < 5e-3
