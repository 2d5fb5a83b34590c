"""The output formats of include/openphase_amd/io.h against the bytes of the reference's own
writers (PhaseField::Write with the zero-valued appended entries, GrainsProperties::Write,
PhaseField::WriteVTK) for the same finalized state: from the CPU oracle (pinned to the reference
bit for bit) here, and from the GPU step in the gpu-marked test."""
from pathlib import Path

import numpy as np
import pytest

from _formats import assert_files_close, write_files
from _golden import GOLDEN, IO_CASES
from openphase_amd import rawio
from openphase_amd.engine import StepConfig
from openphase_amd.microstructure import Lists, quaternion_to_matrix, storage_shape
import oracle as O


class IOCase:
    def __init__(self, name):
        z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
        self.cfg = StepConfig(int(z["nx"]), int(z["ny"]), int(z["nz"]), nphases=int(z["nphases"]), sigma=z["sigma"],
                              mu=z["mu"], energy_model=z["energy_model"] if z["energy_model"].any() else None,
                              epsilon=z["epsilon"] if z["energy_model"].any() else None, slots=16)
        self.dt, self.step = float(z["dt"]), int(z["step"])
        self.phase = z["grain_phase"].astype(np.int32)
        q = z["quaternions"]
        self.quat = q if np.any(q) else None
        self.init = Lists(storage_shape(self.cfg.nx, self.cfg.ny, self.cfg.nz), z["init_count"], z["init_index"],
                          z["init_value"])
        self.expect = {k: z[k].tobytes() for k in ("dat", "grains", "vtk")}

    def run(self, engine):
        engine.set_grains(self.phase, None if self.quat is None else quaternion_to_matrix(self.quat))
        engine.upload(self.init)
        engine.finalize_initialization()
        engine.step(self.dt, self.step)
        lists, flags = engine.download(with_flags=True)
        vol, _ = engine.grains()
        return lists, flags, engine.grain_records(), vol


def vtk_fractions(vtk: bytes, cfg):
    """PhaseFraction_n arrays of a VTK file as the printed strings, [nphases, cells] in cell order
    (x outer, z inner; the file lists points x fastest)."""
    nx, ny, nz = storage_shape(cfg.nx, cfg.ny, cfg.nz)
    lines = vtk.decode().splitlines()
    out = []
    for n in range(cfg.nphases):
        start = next(i for i, ln in enumerate(lines) if f'Name = "PhaseFraction_{n}"' in ln) + 1
        vals = np.array(lines[start:start + nx * ny * nz], dtype=object).reshape(nz, ny, nx)
        out.append(vals.transpose(2, 1, 0).reshape(-1))
    return np.array(out)


def _printed(x):
    """As the VTK writer prints a Float64 (std::defaultfloat, precision 16)."""
    return np.array(["%.16g" % v for v in x.reshape(-1)], dtype=object).reshape(x.shape)


@pytest.mark.parametrize("name", IO_CASES)
def test_oracle_fractions_match_reference_vtk(name):
    """PhaseField::CalculateFractions of the oracle against the PhaseFraction_n fields of the
    reference's own VTK file (printed digits, exactly)."""
    case = IOCase(name)
    orc = O.Oracle(case.cfg)
    case.run(orc)
    frac = orc.phase_fractions()
    orc.close()
    assert np.array_equal(_printed(frac), vtk_fractions(case.expect["vtk"], case.cfg))
    # two phases: a planted mistake (the grain index instead of its phase) changes the digits
    if case.cfg.nphases > 1:
        assert not np.array_equal(_printed(frac[::-1]), vtk_fractions(case.expect["vtk"], case.cfg))


def _check(got, exp, what):
    for k in ("dat", "grains", "vtk"):
        assert got[k] == exp[k], f"{what}: {k} differs ({len(got[k])} vs {len(exp[k])} bytes)"


@pytest.mark.parametrize("name", IO_CASES)
def test_writers_match_reference_bytes(name, tmp_path):
    case = IOCase(name)
    orc = O.Oracle(case.cfg)
    lists, flags, rec, vol = case.run(orc)
    orc.close()
    _check(write_files(tmp_path, case.cfg, lists, flags, case.phase, case.quat, rec, vol), case.expect, name)


def test_full_lists_read_back(tmp_path):
    """The reference-format file with zero-valued entries reads back into the same finalized state
    (FinalizeInitialization erases the zeros; its renormalisation moves values by rounding only)."""
    case = IOCase(IO_CASES[0])
    orc = O.Oracle(case.cfg)
    lists, flags, rec, vol = case.run(orc)
    write_files(tmp_path, case.cfg, lists, flags, case.phase, case.quat, rec, vol)
    full = rawio.read_phasefield(str(tmp_path / "out_full.dat"))
    assert full.count.sum() > lists.count.sum()
    again = O.Oracle(case.cfg)
    again.set_grains(case.phase)
    again.upload(full)
    again.finalize_initialization()
    back = again.download()
    assert np.array_equal(back.count, lists.count) and np.array_equal(back.index, lists.index)
    assert np.abs(back.value - lists.value).max() < 1e-15
    orc.close()
    again.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", IO_CASES)
def test_gpu_state_writes_reference_bytes(name, tmp_path):
    from openphase_amd.engine import Engine

    case = IOCase(name)
    print(f"output formats from the GPU state: {name}", flush=True)
    eng = Engine(case.cfg)
    lists, flags, rec, vol = case.run(eng)
    eng.close()
    got = write_files(tmp_path, case.cfg, lists, flags, case.phase, case.quat, rec, vol)
    # same structure (lists incl. appended zeros, flags, junctions, grain records) and values within
    # the FP64 bar (grain volumes are summed in another order; cubic energies round differently)
    assert_files_close(got, case.expect, 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("name", IO_CASES)
def test_gpu_fractions_match_oracle(name):
    """Per-cell phase fractions computed on the device (PhaseField::CalculateFractions) against
    the oracle's (bit for bit when the states are) and the reference VTK's printed digits."""
    from openphase_amd.engine import Engine

    case = IOCase(name)
    print(f"device fractions: {name}", flush=True)
    eng = Engine(case.cfg)
    lists, _, _, _ = case.run(eng)
    got = eng.phase_fractions()
    eng.close()
    orc = O.Oracle(case.cfg)
    exp_lists, _, _, _ = case.run(orc)
    exp = orc.phase_fractions()
    orc.close()
    assert got.shape == exp.shape
    assert np.abs(got - exp).max() <= 1e-12
    if np.array_equal(lists.value, exp_lists.value):
        assert np.array_equal(got, exp)
        assert np.array_equal(_printed(got), vtk_fractions(case.expect["vtk"], case.cfg))
