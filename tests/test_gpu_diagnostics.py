"""NBodySolver.Step's SimulationState (Solvers/NBodySolver.cs:307-379,
Types/SimulationState.cs:9-25) from the device-resident step loop
(cm_nbody_step_diag): every field equals what the same steps taken one by one
with cm_nbody_step and cm_energy give (bit for bit: the same kernels, only the
bookkeeping differs), and the energies agree with the oracle's exact sums."""
import numpy as np
import pytest

import cmhip
import oracle as O
from conftest import soa_to_state

pytestmark = pytest.mark.gpu

EPS, DT = 1e-4, 1e-3


def _reference_states(be, n_steps, collisions, with_pe):
    """SimulationState by hand, as NBodySolver.Step forms it, from single
    steps and cm_energy."""
    out = []
    e = be.energy(EPS, with_pe)
    e0 = None
    t = 0.0
    for _ in range(n_steps):
        before = e
        k = be.nbody_step(EPS, DT, 1, collisions=collisions)
        e = be.energy(EPS, with_pe)
        t += DT
        tot = e["kinetic"] + e["potential"]
        if e0 is None:
            e0 = tot
        dp = np.array(before["momentum"]) - np.array(e["momentum"])
        out.append({"time": t, "active_body_count": e["active_count"],
                    "kinetic_energy": e["kinetic"], "potential_energy": e["potential"],
                    "total_energy": tot, "total_momentum": tuple(e["momentum"]),
                    "energy_drift": (tot - e0) / abs(e0) if abs(e0) > 1e-15 else 0.0,
                    "collision_count": k, "current_dt": DT,
                    "step_energy_delta": (before["kinetic"] + before["potential"]) - tot,
                    "step_momentum_delta": float(np.sqrt((dp * dp).sum())),
                    "step_angular_momentum_delta": abs(before["angular_momentum_mag"]
                                                       - e["angular_momentum_mag"])})
    return out


@pytest.mark.parametrize("kind,prec,with_pe", [
    (cmhip.BACKEND_DIRECT_PAIRWISE, cmhip.PREC_FP64, 1),
    (cmhip.BACKEND_BARNES_HUT, cmhip.PREC_FP32_SPLIT, 2),
    (cmhip.BACKEND_BARNES_HUT, cmhip.PREC_FP64, 0),
])
def test_simulation_state_matches_single_steps(plummer, kind, prec, with_pe):
    n = 3000
    soa = plummer(n, cut=0.999)
    soa.is_active[:n][::50] = 0
    runs = []
    for diag in (True, False):
        be = cmhip.HipPhysicsBackend(n, kind=kind, precision=prec, theta=0.5)
        be.upload(soa)
        runs.append(be.nbody_step_diag(EPS, DT, 6, collisions=False, with_potential=with_pe)
                    if diag else _reference_states(be, 6, False, with_pe))
    for got, ref in zip(*runs):
        for k, v in ref.items():
            assert got[k] == v, (k, got[k], v)
        assert got["body_count"] == n
    assert runs[0][0]["energy_drift"] == 0.0
    if with_pe == 1:
        # the final energies against the oracle's exact sums of the same state
        be = cmhip.HipPhysicsBackend(n, kind=kind, precision=prec)
        be.upload(soa)
        be.nbody_step(EPS, DT, 6, collisions=False)
        out = cmhip.BodySoA(n)
        be.download(out)
        x = O.energy_exact(soa_to_state(out), EPS, True, nthreads=8)
        last = runs[0][-1]
        assert abs(last["kinetic_energy"] - x["ke"]) <= 1e-13 * abs(x["ke"])
        assert abs(last["potential_energy"] - x["pe"]) <= 1e-13 * abs(x["pe"])


def test_simulation_state_with_collisions():
    """A crowded disk merges bodies every step: CollisionCount and the active
    count follow the single-step path."""
    soa = cmhip.generate_disk(6000, seed=5, a_min=0.5, a_max=0.7, sigma_i=0.002)
    runs = []
    for diag in (True, False):
        be = cmhip.HipPhysicsBackend(soa.count, precision=cmhip.PREC_FP64, theta=0.5)
        be.upload(soa)
        runs.append(be.nbody_step_diag(EPS, DT, 4, collisions=True, with_potential=1)
                    if diag else _reference_states(be, 4, True, 1))
    assert sum(r["collision_count"] for r in runs[0]) > 0
    for got, ref in zip(*runs):
        for k, v in ref.items():
            assert got[k] == v, (k, got[k], v)


def test_energy_tracker_over_device_steps(plummer):
    """EnergyTracker (engine/src/profile/energy_tracker.cpp) over device steps:
    each snapshot's energies, momentum and centre of mass against the oracle's
    exact sums of the downloaded state, and the tracker's drift equal to the
    SimulationState's of the same steps (both relative to the first step)."""
    n = 3000
    soa = plummer(n, cut=0.999)
    soa.is_active[:n][::50] = 0
    be = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5)
    be.upload(soa)
    tr = cmhip.EnergyTracker()
    for k in range(5):
        be.nbody_step(EPS, DT, 1, collisions=False)
        tr.record(cmhip.EnergyTracker.compute(be, EPS, with_potential=1, step_number=k + 1))
        out = cmhip.BodySoA(n)
        be.download(out)
        s = soa_to_state(out)
        x = O.energy_exact(s, EPS, True, nthreads=8)
        cur = tr.current
        assert abs(cur["kinetic_energy"] - x["ke"]) <= 1e-13 * abs(x["ke"])
        assert abs(cur["potential_energy"] - x["pe"]) <= 1e-13 * abs(x["pe"])
        assert abs(cur["virial_ratio"] - 2 * x["ke"] / abs(x["pe"])) <= 1e-12
        cx, cy, cz, m = O.center_of_mass(s)
        assert abs(cur["total_mass"] - m) <= 1e-14 * m
        for got, want in zip(cur["com"], (cx, cy, cz)):
            assert abs(got - want) <= 1e-13
        p = (x["px"], x["py"], x["pz"])
        for got, want in zip(cur["momentum"], p):
            assert abs(got - want) <= 1e-14
        for got, want in zip(cur["com_velocity"], p):
            assert abs(got - want / m) <= 1e-14
        if k == 0:
            p0 = p
    dp = np.array(p) - np.array(p0)
    assert abs(tr.momentum_drift() - float(np.sqrt((dp * dp).sum()))) <= 1e-14
    be2 = cmhip.HipPhysicsBackend(n, precision=cmhip.PREC_FP64, theta=0.5)
    be2.upload(soa)
    states = be2.nbody_step_diag(EPS, DT, 5, collisions=False, with_potential=1)
    assert tr.energy_drift() == states[-1]["energy_drift"]
    assert tr.initial["step_number"] == 1 and tr.current["step_number"] == 5
    loose = cmhip.DiagnosticThresholds(max_energy_drift=1e-3, max_momentum_drift=1e-5,
                                       max_angular_momentum_drift=1e-5,
                                       max_com_velocity_drift=1e-5)
    assert tr.energy_drift() != 0.0 and tr.check_diagnostics(loose)
    loose.max_energy_drift = abs(tr.energy_drift()) / 2
    assert not tr.check_diagnostics(loose)
