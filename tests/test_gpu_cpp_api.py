"""The OpenPhase-style C++ interface (include/openphase_amd/OpenPhaseAMD.h) driving the GPU step:
the NormalGG time loop of the reference, unchanged, against the CPU oracle; its PhaseField::Write,
WriteVTK and GrainsProperties::Write output against the output-format writers fed with the oracle's
state (those writers match the reference's own files byte for byte)."""
import subprocess

import numpy as np
import pytest

from _formats import assert_files_close, write_files
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
    O.write_opi(tmp_path / "in.opi", cfg, dt=cfg.max_time_step(), nsteps=steps)
    rawio.write_phasefield(str(tmp_path / "in.dat"), lists)
    print(f"normal_gg {shape} grains {ngrains} steps {steps}", flush=True)
    (tmp_path / "files" / "VTK").mkdir(parents=True)
    res = subprocess.run([str(EXAMPLE), str(tmp_path / "in.opi"), str(tmp_path / "in.dat"), str(ngrains), str(steps),
                          str(tmp_path / "out.dat"), str(tmp_path / "files")], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    full = rawio.read_phasefield(str(tmp_path / "out.dat"))
    orc = O.Oracle(cfg)
    orc.set_grains(np.zeros(ngrains, np.int32))
    orc.upload(lists)
    orc.finalize_initialization()
    orc.step(cfg.max_time_step(), steps)
    exp, flags = orc.download(with_flags=True)
    # IP.ReportMaximumTimeStep after the loop: from the maxima of the last IP.Set, as the oracle's
    reported = float(res.stdout.split("ReportMaximumTimeStep")[1].split()[0])
    assert reported == orc.max_time_step()
    # PhaseField::Write keeps the zero-valued entries the derivative pass appended
    keep = full.value != 0.0
    cell = np.repeat(np.arange(full.ncells), full.count)
    got_count = np.bincount(cell[keep], minlength=full.ncells)
    assert full.count.sum() > exp.count.sum()
    assert np.array_equal(got_count, exp.count)
    assert np.array_equal(full.index[keep], exp.index)
    assert np.abs(full.value[keep] - exp.value).max() <= 1e-12
    # the same state through the output-format writers: the same files up to the last bits of the
    # values, identical when the GPU values equal the oracle's
    vol, _ = orc.grains()
    ref = write_files(tmp_path, cfg, exp, flags, np.zeros(ngrains, np.int32), None, orc.grain_records(), vol)
    got = {"dat": (tmp_path / "out.dat").read_bytes(),
           "vtk": (tmp_path / "files" / "VTK" / f"PhaseField_{steps:08d}.vts").read_bytes(),
           "grains": (tmp_path / "files" / "GrainsProperties.dat").read_bytes()}
    assert_files_close(got, ref, 1e-12)
    if np.array_equal(full.value[keep], exp.value):
        assert got["dat"] == ref["dat"] and got["vtk"] == ref["vtk"]
