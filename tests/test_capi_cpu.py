"""The C ABI library loads on a machine without a GPU and exports every entry point of
include/mpf.h (no compute calls here)."""
import re
from pathlib import Path

import pytest

from openphase_amd import _capi

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "mpf.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpf_[a-z_]+)\s*\(", text)))


def test_header_declarations_are_known():
    assert set(_declared()) == set(_capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not _capi.LIB_PATH.exists():
        pytest.skip("libmpf_hip.so not built")
    lib = _capi.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_cpp_interface_compiles():
    """The OpenPhase-style header and the NormalGG example compile with the host compiler."""
    import shutil
    import subprocess

    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    res = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
                          str(ROOT / "openphase_amd" / "examples" / "normal_gg.cpp")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_cpp_interface_reads_regularization_and_refuses_extrapolation(tmp_path):
    """The C++ classes read @InterfaceRegularization and CurvatureFactor from the input file, and
    refuse ExtrapolationMode FIRSTORDER (the reference rewrites halo energies inside IP.Set) instead
    of stepping without it; ZEROTHORDER, a no-op in the reference, is accepted. Input parsing only:
    no device context is created."""
    import shutil
    import subprocess

    import oracle as O
    from openphase_amd.engine import StepConfig

    gxx = shutil.which("g++")
    if not gxx or not _capi.LIB_PATH.exists():
        pytest.skip("no g++ or libmpf_hip.so not built")
    src = tmp_path / "probe.cpp"
    src.write_text(r"""
#include <cstdio>
#include "openphase_amd/OpenPhaseAMD.h"
using namespace openphase_amd;
int main(int, char* argv[])
{
    try {
        Settings S; S.ReadInput(argv[1]);
        InterfaceProperties IP(S, argv[1]);
        InterfaceRegularization K(S, argv[1]);
        std::printf("ok %.17g %.17g %d %d %.17g\n", IP.RegularizationFactor, IP.CurvatureFactor, int(K.Averaging),
                    K.Range, K.PhiThreshold);
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
    }
    return 0;
}
""")
    exe = tmp_path / "probe"
    res = subprocess.run([gxx, "-std=c++17", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(src),
                          f"-L{_capi.LIB_PATH.parent}", "-lmpf_hip", f"-Wl,-rpath,{_capi.LIB_PATH.parent}", "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    cfg = StepConfig(32, 32, 0, regularization=True, regularization_factor=1.5, curvature_factor=0.7, reg_range=2,
                     reg_threshold=0.05)
    opi = tmp_path / "in.opi"
    O.write_opi(opi, cfg)
    out = subprocess.run([str(exe), str(opi)], capture_output=True, text=True, timeout=60).stdout.split()
    assert out[0] == "ok" and float(out[1]) == 1.5 and float(out[2]) == 0.7
    assert out[3:5] == ["1", "2"] and float(out[5]) == 0.05
    for mode, ok in (("ZEROTHORDER", True), ("FIRSTORDER", False)):
        text = opi.read_text().replace("@InterfaceProperties\n", f"@InterfaceProperties\n$ExtrapolationMode : {mode}\n")
        (tmp_path / "ex.opi").write_text(text)
        got = subprocess.run([str(exe), str(tmp_path / "ex.opi")], capture_output=True, text=True, timeout=60).stdout
        assert got.startswith("ok") == ok, got


def test_invalid_arguments_fail_without_a_gpu():
    """Argument checks run before any device call: a null grid is MPF_ERR_ARG."""
    if not _capi.LIB_PATH.exists():
        pytest.skip("libmpf_hip.so not built")
    import ctypes as C

    lib = _capi.load()
    h = C.c_void_p()
    assert lib.mpf_create(None, None, C.byref(h)) == 1
    assert lib.mpf_step(None, 1e-3, 1) == 1
    assert lib.mpf_error_message(None) == b"null context"


def test_ensemble_rollout_rejects_mismatched_arguments():
    """Argument checks of ensemble_rollout run before any device work."""
    import types

    import pytest
    import torch

    from openphase_amd.adjoint import ensemble_rollout

    def member(nph):
        eng = types.SimpleNamespace(cfg=types.SimpleNamespace(nphases=nph))
        return types.SimpleNamespace(engine=eng)

    a, b = member(1), member(1)
    ens = types.SimpleNamespace(engines=[a.engine, b.engine])
    p = torch.zeros(2, 3, 1, 1, dtype=torch.float64)
    with pytest.raises(ValueError, match="ensemble's order"):
        ensemble_rollout(ens, [b, a], p, [None, None], 2, [1.0, 1.0])
    with pytest.raises(ValueError, match=r"\[members, 3, P, P\]"):
        ensemble_rollout(ens, [a, b], torch.zeros(2, 3, 2, 2, dtype=torch.float64), [None, None], 2, [1.0, 1.0])
    with pytest.raises(ValueError, match="one initial state"):
        ensemble_rollout(ens, [a, b], p, [None], 2, [1.0, 1.0])
    c = member(2)
    with pytest.raises(ValueError, match="phase count"):
        ensemble_rollout(types.SimpleNamespace(engines=[a.engine, c.engine]), [a, c], p, [None, None], 2, [1.0, 1.0])
