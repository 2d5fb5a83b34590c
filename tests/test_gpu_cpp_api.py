"""The OpenPhase-style C++ interface (include/openphase_amd/OpenPhaseAMD.h) driving the GPU step:
the NormalGG time loop of the reference, unchanged, against the CPU oracle."""
import subprocess

import numpy as np
import pytest

from openphase_amd import rawio
from openphase_amd.engine import StepConfig
from openphase_amd.microstructure import voronoi_grains
import oracle as O

pytestmark = pytest.mark.gpu

EXAMPLE = O.HERE.parent / "openphase_amd" / "examples" / "normal_gg"


@pytest.mark.parametrize("shape,ngrains,steps", [((40, 40, 0), 10, 30), ((20, 20, 20), 8, 10)])
def test_normal_gg_example_matches_oracle(tmp_path, shape, ngrains, steps):
    if not EXAMPLE.exists():
        pytest.fail("openphase_amd/examples/normal_gg is not built")
    cfg = StepConfig(*shape)
    lists = voronoi_grains(*shape, ngrains, seed=21)
    O.write_opi(tmp_path / "in.opi", cfg)
    rawio.write_phasefield(str(tmp_path / "in.dat"), lists)
    print(f"normal_gg {shape} grains {ngrains} steps {steps}", flush=True)
    res = subprocess.run([str(EXAMPLE), str(tmp_path / "in.opi"), str(tmp_path / "in.dat"), str(ngrains), str(steps),
                          str(tmp_path / "out.dat")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = rawio.read_phasefield(str(tmp_path / "out.dat"))
    orc = O.Oracle(cfg)
    orc.set_grains(np.zeros(ngrains, np.int32))
    orc.upload(lists)
    orc.finalize_initialization()
    orc.step(cfg.max_time_step(), steps)
    exp = orc.download()
    assert np.array_equal(got.count, exp.count)
    assert np.array_equal(got.index, exp.index)
    assert np.abs(got.value - exp.value).max() <= 1e-12
