"""Host bookkeeping of cmhip.EnergyTracker against the semantics of the
engine's EnergyTracker (engine/src/profile/energy_tracker.cpp:65-262):
snapshot fields derived from one energy report, drifts against the first
recorded snapshot, the strict threshold comparisons of check_diagnostics and
the 300-entry rolling window.  A stand-in backend returns chosen reports, so
every expected value below is a closed form of those inputs."""
import math

import cmhip


class _Reports:
    """Backend stand-in whose energy() returns the next queued report."""

    def __init__(self, reports):
        self.reports = list(reports)

    def energy(self, softening, with_potential=1):
        return self.reports.pop(0)


def _report(ke, pe, p=(0.0, 0.0, 0.0), lv=(0.0, 0.0, 0.0), m=2.0, com=(0.0, 0.0, 0.0)):
    return {"kinetic": ke, "potential": pe, "momentum": p, "angular_momentum": lv,
            "angular_momentum_mag": math.sqrt(sum(x * x for x in lv)), "total_mass": m,
            "active_count": 3, "center_of_mass": com}


def test_snapshot_fields():
    be = _Reports([_report(0.5, -1.0, p=(3.0, 4.0, 0.0), lv=(0.0, 6.0, 8.0), m=2.0,
                           com=(1.0, 2.0, 3.0)),
                   _report(0.25, 0.0, m=0.0, com=(9.0, 9.0, 9.0))])
    s = cmhip.EnergyTracker.compute(be, 1e-4, step_number=7)
    assert s["total_energy"] == -0.5
    assert s["momentum_magnitude"] == 5.0
    assert s["angular_momentum_magnitude"] == 10.0
    assert s["com"] == (1.0, 2.0, 3.0)
    assert s["com_velocity"] == (1.5, 2.0, 0.0)      # momentum / total mass
    assert s["virial_ratio"] == 1.0                   # 2 KE / |PE|
    assert s["step_number"] == 7
    z = cmhip.EnergyTracker.compute(be, 1e-4)
    # no mass: centre of mass and its velocity stay 0; no potential: virial 0
    assert z["com"] == (0.0, 0.0, 0.0) and z["com_velocity"] == (0.0, 0.0, 0.0)
    assert z["virial_ratio"] == 0.0


def test_drifts_and_thresholds():
    t = cmhip.EnergyTracker()
    assert not t.has_initial and t.check_diagnostics() and t.energy_drift() == 0.0
    be = _Reports([_report(1.0, -2.0, p=(1.0, 0.0, 0.0), lv=(0.0, 0.0, 1.0)),
                   _report(1.0009765625, -2.0, p=(1.0, 0.0, 0.0), lv=(0.0, 0.0, 1.0)),
                   _report(1.0, -2.0, p=(1.0, 3e-8, 4e-8), lv=(0.0, 0.0, 1.0)),
                   _report(1.0, -2.0, p=(1.0, 0.0, 0.0), lv=(0.0, 0.0, 1.0 + 2.0 ** -26))])
    t.record(cmhip.EnergyTracker.compute(be, 1e-4))
    assert t.energy_drift() == 0.0 and t.check_diagnostics()
    t.record(cmhip.EnergyTracker.compute(be, 1e-4))
    # E0 = -1, E = -0.9990234375: drift 2^-10 = 9.765625e-4 <= 1e-3 passes
    assert t.energy_drift() == 2.0 ** -10
    assert t.check_diagnostics()
    assert not t.check_diagnostics(cmhip.DiagnosticThresholds(max_energy_drift=9e-4))
    t.record(cmhip.EnergyTracker.compute(be, 1e-4))
    # |dp| = 5e-8 > 1e-8; the velocity of the centre of mass moves by 2.5e-8
    assert abs(t.momentum_drift() - 5e-8) < 1e-22
    assert abs(t.com_velocity_drift() - 2.5e-8) < 1e-22
    assert not t.check_diagnostics()
    assert t.check_diagnostics(cmhip.DiagnosticThresholds(max_momentum_drift=1e-7,
                                                          max_com_velocity_drift=1e-7))
    t.record(cmhip.EnergyTracker.compute(be, 1e-4))
    assert t.angular_momentum_drift() == 2.0 ** -26      # 1.49e-8 > 1e-8
    assert not t.check_diagnostics(cmhip.DiagnosticThresholds(max_angular_momentum_drift=1e-8))
    assert t.accumulated_error == 2.0 ** -10


def test_rolling_window_and_reset():
    t = cmhip.EnergyTracker()
    assert t.rolling_avg_energy() == 0.0 and t.rolling_avg_drift() == 0.0
    be = _Reports([_report(float(k), -1.0, m=1.0) for k in range(351)])
    for _ in range(351):
        t.record(cmhip.EnergyTracker.compute(be, 1e-4))
    # energies k - 1 for k = 0..350, drifts |k - 1 - (-1)| = k; the window
    # holds the last 300 (k = 51..350)
    assert t.rolling_avg_energy() == sum(k - 1.0 for k in range(51, 351)) / 300
    assert t.rolling_avg_drift() == sum(float(k) for k in range(51, 351)) / 300
    assert t.accumulated_error == sum(float(k) for k in range(351))
    t.reset()
    assert not t.has_initial and t.accumulated_error == 0.0 and t.rolling_avg_drift() == 0.0
