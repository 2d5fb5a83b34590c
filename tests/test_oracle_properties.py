"""Checks of the CPU oracle against what the mathematics fixes (independent of the golden files):
stencil identities, the stationary flat interface, the curvature-driven shrinking law, sum-to-one
and bounds, and plain-Python brute force of one step on tiny grids."""
import math

import numpy as np
import pytest

from openphase_amd.engine import StepConfig
from openphase_amd.microstructure import Lists, voronoi_grains
import oracle as O


def _run(cfg, lists, ngrains, steps, dt=None):
    orc = O.Oracle(cfg)
    orc.set_grains(np.zeros(ngrains, np.int32))
    orc.upload(lists)
    orc.finalize_initialization()
    if steps:
        orc.step(cfg.max_time_step() if dt is None else dt, steps)
    return orc


def _dense(lists, ng):
    out = np.zeros((ng, lists.ncells))
    cell = np.repeat(np.arange(lists.ncells), lists.count)
    out[lists.index, cell] = lists.value
    return out


def _profile_lists(nx, values_a):
    """1-cell-thick slab in x: grain 0 with the given values, grain 1 = 1 - value."""
    count, index, value = [], [], []
    for v in values_a:
        if v >= 1.0:
            count.append(1); index.append(0); value.append(1.0)
        elif v <= 0.0:
            count.append(1); index.append(1); value.append(1.0)
        else:
            count.append(2); index += [0, 1]; value += [v, 1.0 - v]
    return Lists((nx, 1, 1), np.array(count, np.int32), np.array(index, np.int32), np.array(value))


def test_flat_double_obstacle_profile_is_stationary():
    """phi = 1/2 - 1/2 sin(pi (x - x0)/eta) inside the interface is the flat equilibrium of the
    double-obstacle model (theory.tex:46): the profile must not drift (2D, interface along y)."""
    nx, ny = 64, 8
    cfg = StepConfig(nx, ny, 0, iwidth=5.0)
    x0 = 32.0
    eta = cfg.iwidth

    def phi(x):
        s = (x - x0) / eta
        if s <= -0.5:
            return 1.0
        if s >= 0.5:
            return 0.0
        return 0.5 - 0.5 * math.sin(math.pi * s)

    # periodic in x: two interfaces (x0 and its image at x0 - nx/2 with inverted orientation)
    xs = np.arange(nx) + 0.5
    vals = [phi(x) if x > 16 else phi(64 - x) for x in xs]
    base = _profile_lists(nx, vals)
    count = np.repeat(base.count, ny)
    off = base.offsets()
    index = np.concatenate([np.tile(base.index[off[i]:off[i + 1]], ny) for i in range(nx)])
    value = np.concatenate([np.tile(base.value[off[i]:off[i + 1]], ny) for i in range(nx)])
    lists = Lists((nx, ny, 1), count, index, value)
    orc = _run(cfg, lists, 2, 200)
    got = _dense(orc.download(), 2)[0].reshape(nx, ny)[:, 0]
    ref = _dense(lists, 2)[0].reshape(nx, ny)[:, 0]
    # the interface position (integral of phi) stays put to well below a cell
    assert abs(got.sum() - ref.sum()) < 0.05
    assert np.abs(got - ref).max() < 0.05


@pytest.mark.parametrize("dim", [2, 3])
def test_single_grain_shrinks_with_curvature_law(dim):
    """A grain of radius R0 in a matrix shrinks as R^2 = R0^2 - 2(d-1) mu sigma t
    (SingleGrain benchmark, example_singlegrain.tex:10-28); within a few percent."""
    n = 48 if dim == 2 else 32
    r0 = 14.0 if dim == 2 else 11.0
    cfg = StepConfig(n, n, 0 if dim == 2 else n)
    shape = cfg.shape
    grid = np.indices(shape).reshape(3, -1).T.astype(float) + 0.5
    c = np.array([n / 2, n / 2, 0.5 if dim == 2 else n / 2])
    r = np.sqrt(((grid - c)[:, :dim] ** 2).sum(axis=1))
    inside = np.clip(0.5 - (r - r0) / cfg.iwidth, 0.0, 1.0)
    count, index, value = [], [], []
    for v in inside:
        if v >= 1.0:
            count.append(1); index.append(1); value.append(1.0)
        elif v <= 0.0:
            count.append(1); index.append(0); value.append(1.0)
        else:
            count.append(2); index += [0, 1]; value += [1.0 - v, v]
    lists = Lists(shape, np.array(count, np.int32), np.array(index, np.int32), np.array(value))
    dt = cfg.max_time_step()
    orc = _run(cfg, lists, 2, 0)
    vol0 = orc.grains()[0][1]
    steps = 400 if dim == 2 else 150
    orc.step(dt, steps)
    vol1 = orc.grains()[0][1]
    # equivalent radii from the grain volume
    if dim == 2:
        ra, rb = math.sqrt(vol0 / math.pi), math.sqrt(vol1 / math.pi)
    else:
        ra, rb = (3 * vol0 / (4 * math.pi)) ** (1 / 3), (3 * vol1 / (4 * math.pi)) ** (1 / 3)
    rate = (ra * ra - rb * rb) * cfg.dx * cfg.dx / (steps * dt)
    expected = 2 * (dim - 1) * 4e-9 * 0.24
    assert rate == pytest.approx(expected, rel=0.08)


def test_sum_to_one_bounds_and_volume_conservation():
    cfg = StepConfig(48, 48, 0)
    lists = voronoi_grains(48, 48, 0, 14, seed=9)
    orc = _run(cfg, lists, 14, 60)
    got = orc.download()
    assert got.value.min() > 0.0 and got.value.max() <= 1.0
    sums = np.add.reduceat(got.value, np.concatenate([[0], np.cumsum(got.count)[:-1]]))
    assert np.abs(sums - 1.0).max() < 1e-13
    vol, _ = orc.grains()
    assert abs(vol.sum() - got.ncells) < 1e-9 * got.ncells


# ------------------------------------------------------------------------------------------
# brute force: one step restated with Python dicts, straight from the reference formulas
def _brute_step(lists, shape, cfg, vr, dt):
    nx, ny, nz = shape
    dn = [int(cfg.nx > 0), int(cfg.ny > 0), int(cfg.nz > 0)]
    off = lists.offsets()
    cells = [[(int(lists.index[k]), float(lists.value[k])) for k in range(off[c], off[c + 1])] for c in range(lists.ncells)]

    def cid(i, j, k):
        return ((i % nx) * ny + (j % ny)) * nz + (k % nz)

    active = sum(dn)
    lw = {2: {0: -10 / 3, 1: 2 / 3, 2: 1 / 6}, 3: {0: -64 / 15, 1: 7 / 15, 2: 1 / 10, 3: 1 / 30}}[active]
    flag2 = [len(l) > 1 for l in cells]
    P = math.pi ** 2 / (cfg.iwidth * cfg.dx) ** 2
    sigma, mu = float(np.asarray(cfg.sigma).ravel()[0]), float(np.asarray(cfg.mu).ravel()[0])
    out = []
    for c in range(lists.ncells):
        i, j, k = c // (ny * nz), (c // nz) % ny, c % nz
        box = [(a, b, d) for a in range(-dn[0], dn[0] + 1) for b in range(-dn[1], dn[1] + 1) for d in range(-dn[2], dn[2] + 1)]
        wide = flag2[c] or any(flag2[cid(i + a, j + b, k + d)] for a, b, d in box)
        own = list(cells[c])
        if not wide:
            out.append(own)
            continue
        ids = [g for g, _ in own]
        val = {g: v for g, v in own}
        lap = {g: 0.0 for g in ids}
        for a, b, d in box:
            w = lw[(a != 0) + (b != 0) + (d != 0)] / (cfg.dx * cfg.dx)
            for g, v in cells[cid(i + a, j + b, k + d)]:
                if v == 0.0:
                    continue
                if g not in lap:
                    ids.append(g); val[g] = 0.0; lap[g] = 0.0
                lap[g] += w * v
        n = len(ids)
        norm_1 = 1.0 / sum(vr[g] for g in ids)
        psi = {}
        for x in range(n):
            for y in range(x + 1, n):
                ga, gb = ids[x], ids[y]
                scale = math.sqrt(vr[ga] * vr[gb])
                cf = 0.5 * (1.0 + scale) * P
                d = sigma * ((lap[ga] + cf * val[ga]) - (lap[gb] + cf * val[gb]))
                psi[(ga, gb)] = d * (mu * norm_1 * scale)
        if n == 2:
            (ga, gb), v = next(iter(psi.items()))
            va, old = v * dt, val[ga]
            if (old == 0.0 and va < 0.0) or (old == 1.0 and va > 0.0):
                psi = {}
            else:
                nv, nrm = old + va, 1.0
                if nv < 0.0:
                    nrm *= -old / va
                elif nv > 1.0:
                    nrm *= (1.0 - old) / va
                psi = {(ga, gb): v * nrm} if nrm > np.finfo(float).eps else {}
        elif n > 2:
            return None  # the brute force covers 1-2 entry lists only
        for (ga, gb), v in psi.items():
            value = v * dt
            if abs(value) >= np.finfo(float).eps:
                val[ga] += value
                val[gb] += -value
        keep = []
        total = 0.0
        for g in ids:
            v = val[g]
            if v >= 1.0 - np.finfo(float).eps:
                v = 1.0
            if v <= np.finfo(float).eps:
                v = 0.0
            total += v
            if v != 0.0:
                keep.append((g, v))
        if len(keep) > 1:
            keep = [(g, v * (1.0 / total)) for g, v in keep]
        else:
            keep = [(keep[0][0], 1.0)]
        out.append(keep)
    return out


def test_one_step_matches_python_brute_force():
    """Two grains (lists of at most two entries): the oracle equals a direct dict-based restatement."""
    nx, ny = 24, 6
    vals = [1.0 if 6 <= x < 14 else (0.0 if x >= 16 or x < 4 else 0.5) for x in range(nx)]
    base = _profile_lists(nx, vals)
    count = np.repeat(base.count, ny)
    off = base.offsets()
    index = np.concatenate([np.tile(base.index[off[i]:off[i + 1]], ny) for i in range(nx)])
    value = np.concatenate([np.tile(base.value[off[i]:off[i + 1]], ny) for i in range(nx)])
    lists = Lists((nx, ny, 1), count, index, value)
    cfg = StepConfig(nx, ny, 0)
    dt = cfg.max_time_step()
    orc = _run(cfg, lists, 2, 0)
    start = orc.download()
    vol, vr = orc.grains()
    orc.step(dt, 1)
    got = orc.download()
    exp = _brute_step(start, (nx, ny, 1), cfg, {0: vr[0], 1: vr[1]}, dt)
    assert exp is not None
    off = got.offsets()
    for c in range(got.ncells):
        mine = [(int(got.index[k]), float(got.value[k])) for k in range(off[c], off[c + 1])]
        assert [g for g, _ in mine] == [g for g, _ in exp[c]], c
        for (_, a), (_, b) in zip(mine, exp[c]):
            assert a == pytest.approx(b, abs=1e-15), c


@pytest.mark.parametrize("name", ["avg_dg2d_32", "avg_counter2d_24"])
def test_golden_pins_driving_force_averaging(name):
    """The averaging fixtures separate DrivingForce::Average from SkipAverage: the oracle without
    averaging misses them (so a dropped or broken averaging pass cannot pass the golden test)."""
    import dataclasses

    from _golden import Case

    case = Case(name)
    orc = case.setup(O.Oracle(dataclasses.replace(case.cfg, dg_average=False)))
    differs = False
    for s, got, _, _ in case.run(orc):
        exp = case.expect[s][0]
        differs = differs or not (np.array_equal(got.count, exp.count) and np.array_equal(got.index, exp.index)
                                  and np.array_equal(got.value, exp.value))
    orc.close()
    assert differs


def test_averaged_constant_driving_force_is_the_constant():
    """Closed form: averaging a constant driving force returns the constant up to rounding wherever
    weights are positive. With the threshold near zero every pair takes part in both passes, so one
    step equals the unaveraged step to rounding."""
    import dataclasses

    from _golden import Case

    case = Case("avg_dg2d_32")
    b = case.setup(O.Oracle(dataclasses.replace(case.cfg, dg_threshold=1e-9)))
    c = case.setup(O.Oracle(dataclasses.replace(case.cfg, dg_average=False)))
    for o in (b, c):
        o.step(case.dt, 1)
    lb, lc = b.download(), c.download()
    # a threshold near zero lets every pair average itself: the result equals SkipAverage to rounding
    # (a sign slip in the pair orientation or a wrong normalisation would move it by O(1))
    assert np.array_equal(lb.count, lc.count) and np.array_equal(lb.index, lc.index)
    assert np.abs(lb.value - lc.value).max() < 1e-13
    for o in (b, c):
        o.close()


@pytest.mark.parametrize("name,change", [
    ("reg2d_32", {"regularization": False}), ("reg2d_32", {"reg_average": False}), ("reg2d_32", {"reg_range": 2}),
    ("reg3d_16", {"reg_threshold": 0.0}), ("reg_noavg2d_24", {"reg_average": True}),
    ("reg_unit2d_24", {"regularization": False}), ("reg_dg2d_32", {"reg_average": False}),
    ("reg_cubic2d_32", {"regularization": False}),
])
def test_golden_pins_interface_regularization(name, change):
    """The regularization fixtures separate the curvature overload from the plain step, averaged
    from unaveraged curvature, and the Range / Threshold settings; with both factors at 1 they still
    hold the overload's own rounding (loc_increment * mobility * norm_1 * scale, left to right)."""
    import dataclasses

    from _golden import Case

    case = Case(name)
    orc = case.setup(O.Oracle(dataclasses.replace(case.cfg, **change)))
    differs = False
    for s, got, _, _ in case.run(orc):
        exp = case.expect[s][0]
        differs = differs or not (np.array_equal(got.count, exp.count) and np.array_equal(got.index, exp.index)
                                  and np.array_equal(got.value, exp.value))
    orc.close()
    assert differs


def test_regularization_with_equal_factors_doubles_the_mobility():
    """Closed form: with CurvatureFactor = RegularizationFactor = 2 the curvature vanishes
    ((CF - RF) = 0) and the increments are twice the plain ones, i.e. the plain step with twice the
    mobility, to rounding (a factor applied twice, or in the wrong place, moves values by O(dt))."""
    import dataclasses

    from _golden import Case

    case = Case("reg2d_32")
    a = case.setup(O.Oracle(dataclasses.replace(case.cfg, regularization_factor=2.0, curvature_factor=2.0)))
    b = case.setup(O.Oracle(dataclasses.replace(case.cfg, regularization=False, mu=2.0 * np.asarray(case.cfg.mu))))
    dt = 0.5 * case.dt
    for o in (a, b):
        o.step(dt, 5)
    la, lb = a.download(), b.download()
    assert np.array_equal(la.count, lb.count) and np.array_equal(la.index, lb.index)
    assert np.abs(la.value - lb.value).max() < 1e-13
    for o in (a, b):
        o.close()
