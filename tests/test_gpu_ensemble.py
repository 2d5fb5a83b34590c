"""Ensembles of uneven domains (one captured HIP graph per two steps of all members) and the
launch-shape autotuner: both must leave the results of every domain bit-identical to stepping it
alone with the default launch shape."""
import numpy as np
import pytest

from openphase_amd.engine import Engine, Ensemble, StepConfig
from openphase_amd.microstructure import voronoi_grains

pytestmark = pytest.mark.gpu


def _domains():
    """Three 2D grain-growth domains of different sizes, grain counts and tables (one in FP32)."""
    specs = [
        (48, 48, 9, 0.24, 4e-9, 0, 11),
        (64, 40, 14, 0.35, 2e-9, 0, 12),
        (56, 56, 12, 0.18, 7e-9, 1, 13),
    ]
    out = []
    for nx, ny, ng, sig, mu, prec, seed in specs:
        cfg = StepConfig(nx, ny, 0, sigma=np.array([[sig]]), mu=np.array([[mu]]), precision=prec, slots=8)
        out.append((cfg, voronoi_grains(nx, ny, 0, ng, seed=seed), ng, 0.9 * cfg.max_time_step()))
    return out


def _dg_domain():
    """Two phases, constant driving force averaged over the interface."""
    cfg = StepConfig(40, 52, 0, nphases=2, sigma=np.full((2, 2), 0.24), mu=np.full((2, 2), 4e-9),
                     dg_mode=np.array([[0, 1], [1, 0]], np.int32), dg_value=np.array([[0.0, 1e5], [-1e5, 0.0]]),
                     dg_average=True, slots=8)
    return cfg, voronoi_grains(40, 52, 0, 10, seed=14), 10, 0.9 * cfg.max_time_step()


def _make(cfg, lists, ng):
    e = Engine(cfg)
    e.set_grains((np.arange(ng) % cfg.nphases).astype(np.int32))
    e.upload(lists)
    e.finalize_initialization()
    return e


def _same(a, b):
    x, y = a.download(), b.download()
    assert np.array_equal(x.count, y.count)
    assert np.array_equal(x.index, y.index)
    assert np.array_equal(x.value, y.value)
    assert np.array_equal(a.grains()[0], b.grains()[0])


def test_ensemble_matches_members_stepped_alone():
    doms = _domains() + [_dg_domain()]
    solo = [_make(c, l, g) for c, l, g, _ in doms]
    team = [_make(c, l, g) for c, l, g, _ in doms]
    dts = [d for *_, d in doms]
    ens = Ensemble(team)
    # odd count (graph replays + one direct step), then a second call that recaptures
    for n in (7, 4, 6):
        ens.step(dts, n)
        for e, dt in zip(solo, dts):
            e.step(dt, n)
        for a, b in zip(solo, team):
            _same(a, b)
    ens.close()
    for e in solo + team:
        e.close()


def test_autotune_keeps_state_and_results():
    cfg, lists, ng, dt = _domains()[1]
    a, b = _make(cfg, lists, ng), _make(cfg, lists, ng)
    a.step(dt, 3)
    b.step(dt, 3)
    x, g = a.autotune(dt, steps=2)
    assert x in (0, 1) and g in (256, 1024, 4096, 16384)
    _same(a, b)                       # the autotuner restored the state
    a.step(dt, 5)
    b.step(dt, 5)
    _same(a, b)                       # and the shape does not change results
    for xo, gr in ((0, 256), (1, 1024)):
        b.set_launch_shape(xo, gr)
        a.step(dt, 2)
        b.step(dt, 2)
        _same(a, b)
    a.close()
    b.close()
