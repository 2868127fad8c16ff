"""Pins of the oracle's population control, trial-frequency bound and
steady-state test (CPU only).

Population control (nonParallelCollisionTasks, BoltzmannMC.C:1282-1408) is
checked against what its own rules fix exactly, independently of how the
oracle is written: attached slots take the ejected electrons last-in first-out
with the time of the parent's slot; surplus attachments take clones drawn
uniformly from the electrons that are not (or no longer) attached; surplus
ejected electrons and the ensemble lose a uniform subset, so every member of
the pool of N + k survives with probability N / (N + k) and every pair with
N (N - 1) / ((N + k)(N + k - 1)); energyGrowth equals the energy of the
clone sources minus the energy of the removed electrons, to the bit.

The energy bound of maximizationAccelerationEnergy (BoltzmannMC.C:765-802) is
checked against flights of accelerateElectron: no electron starting at or
below the bound's initial energy gets above it within the interval, for every
field case, and in DC the bound is reached exactly by an electron moving along
the acceleration for the whole interval.

checkSteadyState (BoltzmannMC.C:1787-1893) is checked on constructed series:
the comparison of the [t/2, 3t/4] and (3t/4, t] means, the absolute (not
relative) 0.01 eV limit on the standard error of the second part, the closed
interval of the first part and the forced steady state at
maxCollisionsBeforeSteadyState.
"""
import itertools

import numpy as np
import pytest

import ekmc
import oracle_py as O
from helpers import condition, mixture

ION, ATT = 1, 2


def _model(n_e=8, **kw):
    e = np.logspace(-2, 3, 60)
    eth = e[e >= 10.0]
    procs = [dict(type=0, elastic=True, energy=e, sigma=1e-20 * np.ones_like(e)),
             dict(type=ION, threshold=10.0, energy=eth, sigma=1e-21 * np.ones_like(eth)),
             dict(type=ATT, energy=e, sigma=1e-22 * np.ones_like(e))]
    num = ekmc.numerics_default()
    for k, v in kw.items():
        setattr(num, k, v)
    return O.Model(mixture(procs), condition(n_e=n_e), num)


def _ensemble(n):
    """Slot i carries the label i (in x), energy i + 1 and time 10 + i; the
    electron ejected by slot i carries the label 100 + i and energy 101 + i."""
    pos = np.zeros((n, 3))
    pos[:, 0] = np.arange(n)
    ej = np.zeros((n, 3))
    ej[:, 0] = 100 + np.arange(n)
    return pos, np.arange(n) + 1.0, 10.0 + np.arange(n), ej, 101.0 + np.arange(n)


def _run(m, pid, seed):
    n = len(pid)
    pos, en, t, ejp, eje = _ensemble(n)
    dE = np.where(np.asarray(pid) >= 0, -1.0, 0.0)
    p, e, tt, g, cnt = m.population_control(pid, dE, pos, en, t, ejp, eje, seed)
    return p[:, 0].astype(int), e, tt, g, cnt


def test_attached_slots_take_ejected_electrons_last_in_first_out():
    m = _model(6)
    pid = [ION, 0, ION, ATT, ekmc.NULL_COLLISION, ATT]
    lab, e, t, g, cnt = _run(m, pid, 1)
    # attachment order 3 then 5; ejected list [0, 2]: slot 3 <- ejected of 2,
    # slot 5 <- ejected of 0, each with the time of its parent's slot
    assert lab.tolist() == [0, 1, 2, 102, 4, 100]
    assert e.tolist() == [1, 2, 3, 103, 5, 101]
    assert t.tolist() == [10, 11, 12, 12, 14, 10]
    assert g == 0.0
    assert cnt.tolist() == [1, 2, 2]


def test_surplus_attachments_clone_uniform_non_attached_electrons():
    """Slot 2 takes the only ejected electron and is a clone source again;
    slots 5 and 6 take independent uniform clones of {0, 1, 2, 3, 4, 7}."""
    m = _model(8)
    pid = [0, ION, ATT, 0, 0, ATT, ATT, ekmc.PARTIAL_FLIGHT]
    sources = [0, 1, 101, 3, 4, 7]
    energy = {0: 1, 1: 2, 101: 102, 3: 4, 4: 5, 7: 8}
    time = {0: 10, 1: 11, 101: 11, 3: 13, 4: 14, 7: 17}
    runs = 7200
    joint = np.zeros((6, 6))
    for s in range(runs):
        lab, e, t, g, _ = _run(m, pid, 1000 + s)
        assert lab[2] == 101 and t[2] == 11
        a, b = sources.index(lab[5]), sources.index(lab[6])
        joint[a, b] += 1
        for slot in (5, 6):
            assert e[slot] == energy[lab[slot]] and t[slot] == time[lab[slot]]
        assert g == energy[lab[5]] + energy[lab[6]]
        keep = [0, 1, 3, 4, 7]
        assert lab[keep].tolist() == keep
    exp = runs / 36.0
    chi2 = ((joint - exp) ** 2 / exp).sum()
    assert chi2 < 72.0, joint  # 35 degrees of freedom, p ~ 2e-4


def test_surplus_ejected_electrons_remove_a_uniform_subset():
    """Three ionizations and one attachment in N = 5: slot 3 takes the ejected
    electron of slot 2; the pool {0, 1, 2, 102, 4, 100, 101} loses a uniform
    pair."""
    m = _model(5)
    pid = [ION, ION, ION, ATT, 0]
    pool = [0, 1, 2, 102, 4, 100, 101]
    energy = {0: 1, 1: 2, 2: 3, 102: 103, 4: 5, 100: 101, 101: 102}
    runs = 4200
    removed = {pr: 0 for pr in itertools.combinations(pool, 2)}
    alive = {x: 0 for x in pool}
    for s in range(runs):
        lab, e, t, g, _ = _run(m, pid, 5000 + s)
        assert len(set(lab.tolist())) == 5 and set(lab.tolist()) <= set(pool)
        gone = tuple(x for x in pool if x not in lab)
        removed[gone] += 1
        for x in lab:
            alive[x] += 1
        assert [energy[x] for x in lab] == e.tolist()
        assert g == -sum(energy[x] for x in gone)
    exp = runs / 21.0
    chi2 = sum((v - exp) ** 2 / exp for v in removed.values())
    assert chi2 < 50.0, removed  # 20 degrees of freedom, p ~ 2e-4
    for x, k in alive.items():
        p = 5.0 / 7.0
        assert abs(k / runs - p) < 4.0 * np.sqrt(p * (1 - p) / runs), (x, k)


def test_growth_identity_and_constant_n_on_random_passes():
    """Random pass outcomes: N is constant, no attached electron survives in
    its slot unless refilled, and energyGrowth = final energy - energy of the
    pool (non-attached slots + all ejected electrons) exactly."""
    m = _model(64)
    rng = np.random.default_rng(7)
    for s in range(300):
        n = 64
        pid = rng.choice([ekmc.NON_DEF, ekmc.NULL_COLLISION, ekmc.PARTIAL_FLIGHT, 0, ION, ATT], n,
                         p=[0.1, 0.3, 0.2, 0.2, 0.1, 0.1]).astype(np.int32)
        lab, e, t, g, cnt = _run(m, pid, s)
        en0 = np.arange(n) + 1.0
        ej = 101.0 + np.arange(n)
        pool = en0[pid != ATT].sum() + ej[pid == ION].sum()
        assert len(e) == n
        assert e.sum() - pool == g
        for p in (0, ION, ATT):
            assert cnt[p] == (pid == p).sum()
        att = np.flatnonzero(pid == ATT)
        assert not np.any(lab[att] == att)


CASES = {
    "dc": dict(en=300.0, angle=60.0),
    "ac": dict(en=300.0, angle=60.0, freq=3e9),
    "dcb": dict(en=300.0, angle=90.0, bn=900.0),
    "acb": dict(en=300.0, angle=45.0, freq=2e9, bn=900.0),
    "ecr": dict(en=300.0, angle=45.0, bn=900.0),
}


def _field_model(case):
    kw = dict(CASES[case])
    N = 3.22e22
    if case == "ecr":
        kw["freq"] = ekmc.CHARGE * kw["bn"] * 1e-27 * N / ekmc.EMASS / (2 * np.pi)
    cond = condition(**kw)
    cond.gas_density = N
    e = np.logspace(-2, 3, 60)
    mix = mixture([dict(type=0, elastic=True, energy=e, sigma=1e-20 * np.ones_like(e))])
    return O.Model(mix, cond, ekmc.numerics_default())


@pytest.mark.parametrize("case", sorted(CASES))
def test_trial_frequency_energy_bound_holds_for_every_field_case(case):
    m = _field_model(case)
    rng = np.random.default_rng(3)
    for e0, nu in ((0.5, 2e12), (8.0, 5e12), (40.0, 1e13)):
        Dt = 10.0 / nu
        bound = m.max_accel_energy(e0, Dt)
        n = 40000
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        e = e0 * rng.uniform(0, 1, n) ** 0.25
        e[: n // 4] = e0
        f = m.field()
        a = np.array(f.accel)
        if np.linalg.norm(a) > 0:
            d[: n // 8] = a / np.linalg.norm(a)  # along the acceleration
        v = d * np.sqrt(2 * e * ekmc.CHARGE / ekmc.EMASS)[:, None]
        t0 = rng.uniform(0, 1e-8, n)
        dt = Dt * rng.uniform(0, 1, n)
        dt[: n // 2] = Dt
        _, _, eo, _ = m.accelerate(t0, dt, np.zeros((n, 3)), v)
        assert eo.max() <= bound * (1 + 1e-12), (case, e0, eo.max(), bound)
        assert bound > e0


def test_dc_energy_bound_is_reached_along_the_acceleration():
    m = _field_model("dc")
    f = m.field()
    a = np.array(f.accel)
    for e0, Dt in ((0.3, 1e-12), (5.0, 4e-12), (30.0, 1e-11)):
        v = a / np.linalg.norm(a) * np.sqrt(2 * e0 * ekmc.CHARGE / ekmc.EMASS)
        _, _, eo, _ = m.accelerate(np.zeros(1), np.array([Dt]), np.zeros((1, 3)), v[None, :])
        bound = m.max_accel_energy(e0, Dt)
        assert abs(eo[0] - bound) <= 1e-12 * bound
        # against the closed form e0 + E v0 Dt + e E^2 Dt^2 / (2 m_e)
        E = np.linalg.norm(f.elec_field)
        v0 = np.sqrt(2 * e0 * ekmc.CHARGE / ekmc.EMASS)
        closed = e0 + E * v0 * Dt + 0.5 * ekmc.CHARGE / ekmc.EMASS * E ** 2 * Dt ** 2
        assert abs(bound - closed) <= 1e-13 * closed


def _series(first, second_level, delta, n_first=16, n_second=16):
    """Sampling times 0..T: points in [0, T/2) at `first`, in [T/2, 3T/4] at
    first, in (3T/4, T] alternating second_level +/- delta."""
    n = 4 * n_first
    t = np.linspace(0.0, 1.0, n + 1)
    e = np.full(n + 1, first, float)
    late = t > 0.75
    k = np.flatnonzero(late)
    e[k] = second_level + delta * np.where(np.arange(len(k)) % 2 == 0, 1.0, -1.0)
    return t, e


def test_steady_state_needs_a_non_increasing_mean_energy():
    m = _model(100, max_collisions_before_ss=1e100)
    t = np.linspace(0, 1e-6, 401)
    assert m.check_steady_state(t, 5.0 + 3.0 * np.exp(-t / 1e-7), t[-1], 10)
    assert not m.check_steady_state(t, 5.0 - 3.0 * np.exp(-t / 1e-7), t[-1], 10)


@pytest.mark.parametrize("level", [0.1, 1000.0])
def test_steady_state_standard_error_limit_is_absolute_in_ev(level):
    """The second part's standard error must be below 0.01 (eV, not relative):
    16 points alternating +/- delta have standard error delta / 4."""
    m = _model(100, max_collisions_before_ss=1e100)
    t, e = _series(level + 1.0, level, 0.039)
    assert (t > 0.75).sum() == 16
    assert m.check_steady_state(t, e, t[-1], 10)
    t, e = _series(level + 1.0, level, 0.041)
    assert not m.check_steady_state(t, e, t[-1], 10)


def test_steady_state_first_part_is_a_closed_interval():
    """A point exactly at 3T/4 belongs to the first part: a large value there
    raises the first mean (steady state), not the second part's spread."""
    m = _model(100, max_collisions_before_ss=1e100)
    t, e = _series(5.0, 5.0, 0.0)
    i = np.flatnonzero(t == 0.75)[0]
    e[i] = 50.0
    assert m.check_steady_state(t, e, t[-1], 10)
    e[i] = 5.0
    e[np.flatnonzero(t == 0.5)[0]] = -50.0  # the first part starts at T/2, inclusive
    assert not m.check_steady_state(t, e, t[-1], 10)


def test_steady_state_forced_by_max_collisions_before_ss():
    m = _model(100, max_collisions_before_ss=30)
    t = np.linspace(0, 1e-6, 401)
    rising = 5.0 - 3.0 * np.exp(-t / 1e-7)
    assert not m.check_steady_state(t, rising, t[-1], 2999)
    assert m.check_steady_state(t, rising, t[-1], 3000)
