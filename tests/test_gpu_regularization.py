"""Interface regularization on the GPU (DoubleObstacle's curvature overload,
src/DoubleObstacle.cpp:686-770, with InterfaceRegularization::Average and
MergePhaseFieldIncrements, src/InterfaceRegularization.cpp:156-456) against the oracle on the same
seeded inputs, beyond the golden cases (which test_gpu_parity runs for every golden fixture), and
the combinations the library refuses."""
import numpy as np
import pytest

from openphase_amd import _capi
from openphase_amd._capi import MpfError
from openphase_amd.engine import Engine, Ensemble, StepConfig
from openphase_amd.microstructure import voronoi_grains
import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12

CASES = {
    # 2D, default Range (3) and Threshold, RegularizationFactor 1.5, CurvatureFactor 0.7
    "2d": (StepConfig(64, 64, 0, slots=16, regularization=True, regularization_factor=1.5, curvature_factor=0.7),
           24, 60),
    # 3D, two phases with TJ VALUE, Range 2, Threshold 0.05
    "3d": (StepConfig(32, 32, 32, nphases=2, sigma=np.array([[0.24, 0.30], [0.30, 0.20]]),
                      mu=np.array([[4e-9, 2e-9], [2e-9, 3e-9]]), tj_model=1, tj_energy=0.05, slots=16,
                      regularization=True, regularization_factor=2.0, curvature_factor=1.0, reg_range=2,
                      reg_threshold=0.05), 16, 12),
    # the 256^3 benchmark's grain size on 96^3 (53 grains), 20 steps
    "3d-96": (StepConfig(96, 96, 96, slots=16, regularization=True, regularization_factor=1.5, curvature_factor=0.7),
              53, 20),
    # a reduced y axis, curvature not averaged (the merge adds nothing, the increments are scaled)
    "xz-noavg": (StepConfig(40, 0, 36, slots=16, regularization=True, regularization_factor=0.8,
                            curvature_factor=0.5, reg_average=False), 12, 40),
}


@pytest.mark.parametrize("kernel", ["tile", "cell"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_regularization_matches_oracle(name, kernel):
    cfg, ngrains, steps = CASES[name]
    lists = voronoi_grains(cfg.nx, cfg.ny, cfg.nz, ngrains, seed=300 + ngrains)
    phase = (np.arange(ngrains) % cfg.nphases).astype(np.int32)
    dt = 0.9 * cfg.max_time_step()
    print(f"regularization {name} {cfg.shape} grains {ngrains} steps {steps} kernel {kernel}", flush=True)
    eng = Engine(cfg)
    if kernel == "cell":
        eng.set_tile_shape(-1)
    orc = O.Oracle(cfg)
    for e in (eng, orc):
        e.set_grains(phase)
        e.upload(lists)
        e.finalize_initialization()
    done = 0
    for s in (1, steps // 3, steps):
        eng.step(dt, s - done)
        orc.step(dt, s - done)
        done = s
        got, exp = eng.download(), orc.download()
        assert np.array_equal(got.count, exp.count), f"step {s}: entry counts differ"
        assert np.array_equal(got.index, exp.index), f"step {s}: index lists differ"
        err = np.abs(got.value - exp.value).max()
        assert err <= FP64_TOL, f"step {s}: max |dphi| {err}"
        _, orr = orc.grains()
        if np.all(orr == 1.0):
            # VolumeRatio 1 everywhere: the same operations in the same order, bit for bit
            assert np.array_equal(got.value, exp.value), f"step {s}: not bit-exact ({err})"
    assert eng.limiter_warnings() == orc.limiter_warnings()
    assert eng.overflow_cell() == -1
    eng.close()
    orc.close()


def test_regularization_refusals():
    """FP32 storage, ensembles, the adjoint and the allocation-free step do not take the
    regularization; each says so with MPF_ERR_UNSUPPORTED instead of stepping without it."""
    base = dict(slots=8, regularization=True, regularization_factor=1.5, curvature_factor=0.7)
    with pytest.raises(MpfError) as ei:
        Engine(StepConfig(32, 32, 0, precision=_capi.MPF_FP32, **base))
    assert ei.value.status == _capi.MPF_ERR_UNSUPPORTED
    eng = Engine(StepConfig(32, 32, 0, **base))
    eng.set_grains(np.zeros(6, np.int32))
    eng.upload(voronoi_grains(32, 32, 0, 6, seed=7))
    eng.finalize_initialization()
    with pytest.raises(MpfError) as ei:
        Ensemble([eng])
    assert ei.value.status == _capi.MPF_ERR_UNSUPPORTED
    eng.close()
