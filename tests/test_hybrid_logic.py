"""Host logic of the hybrid integrator (activemodelbplus.surrogate.HybridIntegrator) on
the CPU, with stand-ins for the surrogate and the two integrators that script the audit
outcomes: interval doubling and reset, fallback to Euler after a failed audit, a
non-finite audit error, bounded speculation, and the step accounting."""
import math

import numpy as np
import pytest

from activemodelbplus import surrogate as SG
from activemodelbplus.integrator import Model, Stencil


class FakeIntegrator:
    def __init__(self):
        self.euler_steps = 0

    def run(self, n):
        self.euler_steps += n


class FakeSurrogate:
    """audit_end returns the scripted errors in turn (then a passing one)."""

    def __init__(self, errors, tol):
        self.errors, self.tol = list(errors), tol
        self.pending, self.advanced, self.log = False, 0, []

    def audit_begin(self, main, judge, params3, delta):
        assert not self.pending
        self.pending = True

    def advance(self, main, k, params3, delta):
        assert k > 0
        self.advanced += k
        self.log.append(("advance", k))

    def audit_end(self, tol):
        assert self.pending
        self.pending = False
        err = self.errors.pop(0) if self.errors else 0.001
        self.log.append(("audit", err))
        return err, bool(err <= tol)  # False for NaN, as the device audit


def hybrid(errors, audit_every=10, max_audit_every=200, tol=0.02):
    sur = FakeSurrogate(errors, tol)
    main, judge = FakeIntegrator(), FakeIntegrator()
    h = SG.HybridIntegrator(sur, np.zeros((4, 4)), Model(-0.25, 0, 0.25, 1, 0.5, -1, 0), Stencil(5e-3, 1, 1),
                            audit_every=audit_every, max_audit_every=max_audit_every, tol=tol, main=main, judge=judge)
    return h, sur, main


def test_intervals_double_to_the_cap_when_audits_pass():
    h, sur, main = hybrid([])
    h.run_coarse(1000)
    st = h.stats
    # segments 10, 20, 40, 80, 160, 200, 200, 200, 90
    assert st["audits"] == 9 and st["audit_failures"] == 0
    assert st["coarse_steps"] == 1000 and st["surrogate_steps"] == 1000 and st["fallback_steps"] == 0
    assert sur.advanced == 1000 - 9 and main.euler_steps == 0


def test_failed_audit_falls_back_for_the_minimum_interval():
    h, sur, main = hybrid([0.001, 0.001, 0.05])  # third audit (segment of 40) fails
    h.run_coarse(200)
    st = h.stats
    assert st["audit_failures"] == 1
    assert st["discarded_steps"] == 16           # speculation bound, not the 39 of the segment
    assert st["fallback_steps"] == 10            # the minimum interval, on the Euler path
    assert main.euler_steps == 9 * 100           # audited step comes from the judge
    assert st["coarse_steps"] == 200 and st["surrogate_steps"] + st["fallback_steps"] == 200
    # segments 10, 20, 40 (fails: 10 on Euler), then 10, 20, 40, 80 and the last 10
    assert st["audits"] == 8


def test_mid_band_error_keeps_the_interval():
    h, _, _ = hybrid([0.015, 0.015])  # within tol but above tol / 2
    h.run_coarse(20)
    assert h.stats["audits"] == 2 and h.interval == 10


def test_nonfinite_audit_error_is_a_failure():
    h, _, main = hybrid([math.nan])
    h.run_coarse(10)
    st = h.stats
    assert st["audit_failures"] == 1 and st["fallback_steps"] == 10 and main.euler_steps == 900
    assert st["max_audit_error"] == math.inf


def test_speculation_runs_before_the_audit_is_settled():
    h, sur, _ = hybrid([], audit_every=50, max_audit_every=50)
    h.run_coarse(50)
    assert sur.log[:3] == [("advance", 16), ("audit", 0.001), ("advance", 33)]


def test_fixed_interval_by_default():
    h, _, _ = hybrid([], audit_every=20, max_audit_every=None)
    h.run_coarse(200)
    assert h.stats["audits"] == 10
