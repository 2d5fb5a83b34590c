"""The CPU oracle against the golden vectors of the real OpenPhase code (bit for bit)."""
import numpy as np
import pytest

from _golden import CASES, GOLDEN, Case
import oracle as O


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_bitwise(name):
    case = Case(name)
    orc = case.setup(O.Oracle(case.cfg))
    # InterfaceProperties::ReportMaximumTimeStep before any IP.Set: the maxima are still 0
    assert orc.max_time_step() == case.max_dt_init == np.finfo(np.float64).max
    for s, got, vol, vr in case.run(orc):
        exp, evol, evr = case.expect[s]
        assert np.array_equal(got.count, exp.count), f"{name} step {s}: entry counts differ"
        assert np.array_equal(got.index, exp.index), f"{name} step {s}: active index lists differ"
        assert np.array_equal(got.value, exp.value), (
            f"{name} step {s}: values differ, max {np.abs(got.value - exp.value).max()}")
        assert np.array_equal(vol, evol), f"{name} step {s}: grain volumes differ"
        assert np.array_equal(vr, evr), f"{name} step {s}: volume ratios differ"
        if s > 0:
            # maxima of energy and mobility the step's IP.Set saw (CUBIC: orientation dependent)
            assert orc.max_time_step() == case.max_dt[s - 1], f"{name} step {s}: ReportMaximumTimeStep differs"
            # DoubleObstacle::AverageEnergyDensity with that IP.Set's energies (one thread: the
            # reference's summation order)
            assert orc.interface_energy("previous") / case.ncells == case.energy[s - 1], f"{name} step {s}: energy"
    assert orc.limiter_warnings() == case.limiter_warnings
    assert orc.interface_energy("current") / case.ncells == case.energy_ip_current
    assert orc.interface_energy("none") == 0.0


def test_max_time_step_closed_forms():
    """dt_max = c dx^2 / (mu sigma RegularizationFactor), c = 0.25 (2D), 0.16 (3D), for the
    benchmark parameters dx = 1e-6 m, mu = 4e-9, sigma = 0.24: 2.6042e-4 s and 1.6667e-4 s."""
    assert Case("gg2d_32").max_dt[-1] == pytest.approx(0.25e-12 / (4e-9 * 0.24), rel=1e-15)
    assert Case("gg3d_16").max_dt[-1] == pytest.approx(0.16e-12 / (4e-9 * 0.24), rel=1e-15)
    assert Case("gg2d_32").max_dt[-1] == pytest.approx(2.6042e-4, rel=1e-4)
    assert Case("gg3d_16").max_dt[-1] == pytest.approx(1.6667e-4, rel=1e-4)
    # two phases with different tables: the largest energy and mobility present (0.30, 4e-9)
    assert Case("twophase2d_32").max_dt[-1] == pytest.approx(0.25e-12 / (4e-9 * 0.30), rel=1e-15)


def _limiter_runs():
    z = np.load(GOLDEN / "limiter_cases.npz", allow_pickle=False)
    for r in range(int(z["runs"])):
        cnt = z[f"r{r}_count"]
        start = np.concatenate([[0], np.cumsum(cnt)])
        pin = {k: z[f"r{r}_in_{k}"] for k in "nabv"}
        pout = {k: z[f"r{r}_out_{k}"] for k in "nabv"}
        ps_in = np.concatenate([[0], np.cumsum(pin["n"])])
        ps_out = np.concatenate([[0], np.cumsum(pout["n"])])
        for c in range(cnt.size):
            own = slice(start[c], start[c + 1])
            pi = slice(ps_in[c], ps_in[c + 1])
            po = slice(ps_out[c], ps_out[c + 1])
            yield (z[f"r{r}_index"][own], z[f"r{r}_value"][own],
                   list(zip(pin["a"][pi], pin["b"][pi], pin["v"][pi])),
                   list(zip(pout["a"][po], pout["b"][po], pout["v"][po])),
                   int(z[f"r{r}_warnings"][c]), float(z["dt"]))


def test_limiter_matches_reference_on_injected_increments():
    """PhaseField::NormalizeIncrementsSR of the reference on pair increments set directly
    (tests/golden/make_limiter_cases.py): the oracle keeps the same pairs in the same order with
    bit-identical limited values and counts the same plausibility warnings. A warning means the
    limiter stopped at its 25-pass cap before reaching [0, 1], so the fixture covers the cap."""
    nodes = warned = 0
    for idx, val, pairs_in, pairs_out, warn, dt in _limiter_runs():
        got, w = O.normalize_pairs(idx, val, pairs_in, dt)
        assert [(a, b) for a, b, _ in got] == [(int(a), int(b)) for a, b, _ in pairs_out]
        assert np.array_equal(np.array([v for *_, v in got]), np.array([v for *_, v in pairs_out], np.float64))
        assert w == warn
        nodes += 1
        warned += w > 0
    assert nodes == 6 * 36 and warned >= 10


def test_limiter_fixture_separates_a_changed_cap():
    """The fixture is sensitive to the pass cap: stopping after 24 or 26 passes changes the limited
    values of some node (a Python restatement of the loop of src/PhaseField.cpp:1451-1519 run
    with the cap moved, on the same inputs)."""
    def limit(idx, val, pairs, dt, cap):
        order = []
        vals = {}
        for a, b, v in pairs:  # NodeAB::add_asym1
            if (a, b) in vals:
                vals[(a, b)] += v
            elif (b, a) in vals:
                vals[(b, a)] += v * (-1.0)
            else:
                vals[(a, b)] = v
                order.append((a, b))
        phi = dict(zip((int(i) for i in idx), val))
        it, need = 0, True
        while need:
            it += 1
            need = False
            pos = {i: 0.0 for i in phi}
            neg = {i: 0.0 for i in phi}
            for k in order:
                a, b = k
                if vals[k] < 0.0:
                    neg[a] += vals[k]; pos[b] += -vals[k]
                if vals[k] > 0.0:
                    pos[a] += vals[k]; neg[b] += -vals[k]
            l1, l2 = {}, {}
            for i, p in phi.items():
                npf = p + (pos[i] + neg[i]) * dt
                l1[i] = l2[i] = 1.0
                if npf < 0.0:
                    l2[i] = min(1.0, -(p + pos[i] * dt) / (neg[i] * dt))
                if npf > 1.0:
                    l1[i] = min(1.0, (1.0 - (p + neg[i] * dt)) / (pos[i] * dt))
            for k in order:
                a, b = k
                if vals[k] < 0.0:
                    t = min(l2[a], l1[b]); vals[k] *= t; need |= t < 1.0
                if vals[k] > 0.0:
                    t = min(l1[a], l2[b]); vals[k] *= t; need |= t < 1.0
            if it > cap - 1:
                need = False
        return [vals[k] for k in order]

    differs = {24: False, 26: False}
    for idx, val, pairs_in, pairs_out, warn, dt in _limiter_runs():
        if warn == 0:
            continue
        ref25 = limit(idx, val, pairs_in, dt, 25)
        for cap in differs:
            differs[cap] |= limit(idx, val, pairs_in, dt, cap) != ref25
    assert all(differs.values())
