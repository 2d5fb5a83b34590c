"""mpf_step_into (one step without allocation, scratch in a static device array) against mpf_step
through the C ABI: bit-identical states and grain volumes on two calls in a row, the second on a
smaller grid; sizes beyond the scratch limit are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from openphase_amd import _capi
from openphase_amd._capi import MpfError
from openphase_amd.engine import Engine, StepConfig
from openphase_amd.microstructure import Lists, voronoi_grains

pytestmark = pytest.mark.gpu


def _pair(cfg, ngrains, seed):
    lists = voronoi_grains(cfg.nx, cfg.ny, cfg.nz, ngrains, seed=seed)
    engines = []
    for _ in range(2):
        e = Engine(cfg)
        e.set_grains(np.zeros(ngrains, np.int32))
        e.upload(lists)
        e.finalize_initialization()
        engines.append(e)
    return engines


def _step_into(eng, dt):
    """One mpf_step_into from the engine's current device state into fresh caller buffers;
    returns the output state as Lists."""
    cfg = eng.cfg
    n = int(np.prod(cfg.shape))
    vt = torch.float32 if cfg.precision == _capi.MPF_FP32 else torch.float64
    cnt = torch.zeros(n, dtype=torch.uint8, device="cuda")
    ids = torch.zeros(cfg.slots * n, dtype=torch.int16, device="cuda")
    val = torch.zeros(cfg.slots * n, dtype=vt, device="cuda")
    c_in, i_in, v_in, _ = eng.device_state()
    torch.cuda.synchronize()
    st = eng.lib.mpf_step_into(eng.h, float(dt), C.c_void_p(c_in), C.c_void_p(i_in), C.c_void_p(v_in),
                               C.c_void_p(cnt.data_ptr()), C.c_void_p(ids.data_ptr()), C.c_void_p(val.data_ptr()))
    eng._check(st, "mpf_step_into")
    count = cnt.cpu().numpy().astype(np.int32)
    if cfg.precision == _capi.MPF_FP32:
        raise AssertionError("FP64 states only in this test")
    ii = ids.cpu().numpy().view(np.uint16).astype(np.int32).reshape(cfg.slots, n)
    vv = val.cpu().numpy().reshape(cfg.slots, n)
    return Lists.from_slots(cfg.shape, count, ii, vv)


@pytest.mark.parametrize("first,second", [((48, 48, 48, 20), (64, 64, 0, 12))])
def test_step_into_matches_step(first, second):
    for nx, ny, nz, ng in (first, second):  # the second call is on a smaller grid
        cfg = StepConfig(nx, ny, nz, slots=8 if nz == 0 else 16)
        print(f"mpf_step_into {cfg.shape} grains {ng}", flush=True)
        a, b = _pair(cfg, ng, seed=7 + ng)
        dt = cfg.max_time_step()
        b.step(dt, 1)
        got = _step_into(a, dt)
        exp = b.download()
        assert np.array_equal(got.count, exp.count)
        assert np.array_equal(got.index, exp.index)
        assert np.array_equal(got.value, exp.value)
        (va, ra), (vb, rb) = a.grains(), b.grains()
        assert np.array_equal(va, vb) and np.array_equal(ra, rb)
        # the context's own state is untouched: stepping it still gives the same first step
        a.step(dt, 1)
        again = a.download()
        assert np.array_equal(again.index, exp.index) and np.array_equal(again.value, exp.value)
        for e in (a, b):
            e.close()


def test_step_into_refuses_grids_beyond_scratch():
    cfg = StepConfig(81, 81, 81, slots=8)  # 531441 cells > 2^19
    e = Engine(cfg)
    lists = voronoi_grains(81, 81, 81, 8, seed=1)
    e.set_grains(np.zeros(8, np.int32))
    e.upload(lists)
    e.finalize_initialization()
    c_in, i_in, v_in, _ = e.device_state()
    st = e.lib.mpf_step_into(e.h, 1e-4, C.c_void_p(c_in), C.c_void_p(i_in), C.c_void_p(v_in),
                             C.c_void_p(c_in + 1), C.c_void_p(i_in + 2), C.c_void_p(v_in + 8))
    assert st == _capi.MPF_ERR_UNSUPPORTED
    assert b"scratch limit" in e.lib.mpf_error_message(e.h)
    with pytest.raises(MpfError):
        e._check(st, "mpf_step_into")
    e.close()
