"""The HIP step against the CPU oracle at the benchmark configurations' own sizes.

- 2D normal grain growth, 256 x 256 periodic, 64 Voronoi grains (seed 0x0FE1), 1000 steps at the
  stability limit: active sets and phase fractions bit for bit at steps 1, 100 and 1000.
- 3D normal grain growth, 256^3 periodic, 1000 Voronoi grains (seed 0x0FE2, one-cell overlap
  band as the reference's Voronoi set-up): bit for bit against the oracle (OpenMP, threads
  stated) for 20 steps. Between steps 4 and 23 a few cells hold 9 phase fields (the oracle's
  lists, unbounded like the reference's, show it), so the comparison runs with 16 slots and a
  context with the 8-slot cap reports the overflow at step 4 with the first such cell.
- The FP32 variant of the same 3D run against the FP64 GPU step (bit-equal to the oracle above):
  maximum / mean phase-fraction error and per-grain volume drift at steps 100, 500 and 1000.
  The maximum is set by discrete events, not by rounding drift: an entry a neighbour appends at
  value 0 enters a list when its pair increment is positive, and that sign comes from
  (L_a + P phi_a) - L_b, differences of terms near 1e11 that FP32 storage (relative 6e-8) only
  resolves to about 1e4. Once a tiny entry (1.8e-8 in the first such cell, at step 4) exists in
  one run and not in the other, LocalN of the neighbouring cells differs (3 vs 4 entries) and
  their increments scale by 3/4: 6.8e-3 in one step. The mean error stays near 1e-7 per 100 steps.
"""
import os
import time

import numpy as np
import pytest

from openphase_amd import _capi
from openphase_amd._capi import MpfError
from openphase_amd.engine import Engine, StepConfig
from openphase_amd.microstructure import voronoi_grains
import oracle as O

pytestmark = pytest.mark.gpu

# FP32 variant at 256^3 after 1000 steps (measured on MI355X: max 3.4e-2, mean 1.3e-4, largest
# grain volume drift 21 cells): bounds with a margin; the 1e-5 max-error target is not reached
FP32_MAX_1000 = 0.1
FP32_MEAN_1000 = 1e-3


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def fraction_errors(a, b, ngrains):
    """Max and mean |phi_a - phi_b| over every (cell, grain) entry present in either state."""
    def keys(x):
        cell = np.repeat(np.arange(x.ncells, dtype=np.int64), x.count)
        return cell * ngrains + x.index.astype(np.int64)
    k = np.concatenate([keys(a), keys(b)])
    v = np.concatenate([a.value, -b.value])
    _, inv = np.unique(k, return_inverse=True)
    d = np.abs(np.bincount(inv, weights=v))
    return float(d.max()), float(d.mean())


def _same(got, exp, what):
    assert np.array_equal(got.count, exp.count), f"{what}: entry counts differ"
    assert np.array_equal(got.index, exp.index), f"{what}: active index lists differ"
    assert np.array_equal(got.value, exp.value), f"{what}: values differ by {np.abs(got.value - exp.value).max()}"


def test_2d_256_64_grains_1000_steps_bit_exact():
    cfg = StepConfig(256, 256, 0, slots=8)
    lists = voronoi_grains(256, 256, 0, 64, seed=0x0FE1)
    phase = np.zeros(64, np.int32)
    dt = cfg.max_time_step()
    print(f"2D 256^2, 64 grains, dt {dt}", flush=True)
    eng, orc = Engine(cfg), O.Oracle(cfg, threads=_threads())
    for e in (eng, orc):
        e.set_grains(phase)
        e.upload(lists)
        e.finalize_initialization()
    assert eng.max_time_step() == orc.max_time_step() == np.finfo(np.float64).max
    done = 0
    for s in (1, 100, 1000):
        for e in (eng, orc):
            e.step(dt, s - done)
        done = s
        _same(eng.download(), orc.download(), f"step {s}")
        (gv, gr), (ov, orr) = eng.grains(), orc.grains()
        np.testing.assert_allclose(gv, ov, rtol=1e-13, atol=1e-9)
        assert np.array_equal(gr, orr)
        assert eng.max_time_step() == orc.max_time_step() == pytest.approx(2.6042e-4, rel=1e-4)
        print(f"step {s}: bit-exact, {int(eng.download().count.sum())} entries", flush=True)
    assert eng.limiter_warnings() == orc.limiter_warnings()
    assert eng.overflow_cell() == -1
    eng.close()
    orc.close()


def test_3d_256_1000_grains_bit_exact_20_steps():
    cfg = StepConfig(256, 256, 256, slots=16)
    lists = voronoi_grains(256, 256, 256, 1000, seed=0x0FE2, sides="plus")
    phase = np.zeros(1000, np.int32)
    dt = cfg.max_time_step()
    threads = _threads()
    print(f"3D 256^3, 1000 grains, oracle on {threads} threads", flush=True)
    eng, orc = Engine(cfg), O.Oracle(cfg, threads=threads)
    cap8 = Engine(StepConfig(256, 256, 256, slots=8))
    for e in (eng, orc, cap8):
        e.set_grains(phase)
        e.upload(lists)
        e.finalize_initialization()
    done = 0
    for s in (1, 3, 4, 5, 20):
        t0 = time.perf_counter()
        orc.step(dt, s - done)
        t1 = time.perf_counter()
        eng.step(dt, s - done)
        exp = orc.download()
        if s <= 4:
            if exp.count.max() <= 8:
                cap8.step(dt, s - done)
            else:
                with pytest.raises(MpfError) as err:
                    cap8.step(dt, s - done)
                assert err.value.status == _capi.MPF_SLOT_OVERFLOW
                assert cap8.overflow_cell() == int(np.nonzero(exp.count > 8)[0][0])
                print(f"step {s}: 8 slots overflow at cell {cap8.overflow_cell()}", flush=True)
        done = s
        _same(eng.download(), exp, f"step {s}")
        (gv, gr), (ov, orr) = eng.grains(), orc.grains()
        np.testing.assert_allclose(gv, ov, rtol=1e-13, atol=1e-9)
        assert np.array_equal(gr, orr)
        assert eng.max_time_step() == orc.max_time_step() == pytest.approx(1.6667e-4, rel=1e-4)
        print(f"step {s}: bit-exact (oracle {t1 - t0:.1f} s)", flush=True)
    assert eng.limiter_warnings() == orc.limiter_warnings()
    assert eng.overflow_cell() == -1
    for e in (eng, orc, cap8):
        e.close()


def test_3d_256_fp32_error_vs_fp64():
    """FP32 fields, stencils and increments with FP64 limiter / normalization / finalize and grain
    volumes against the FP64 GPU step (which equals the oracle bit for bit)."""
    lists = voronoi_grains(256, 256, 256, 1000, seed=0x0FE2, sides="plus")
    phase = np.zeros(1000, np.int32)
    c64 = StepConfig(256, 256, 256, slots=16)
    c32 = StepConfig(256, 256, 256, slots=15, precision=1)
    dt = c64.max_time_step()
    e64, e32 = Engine(c64), Engine(c32)
    for e in (e64, e32):
        e.set_grains(phase)
        e.upload(lists)
        e.finalize_initialization()
    done = 0
    report = {}
    for s in (100, 500, 1000):
        for e in (e64, e32):
            e.step(dt, s - done)
        done = s
        a, b = e64.download(), e32.download()
        emax, emean = fraction_errors(a, b, 1000)
        v64, _ = e64.grains()
        v32, _ = e32.grains()
        drift = np.abs(v32 - v64)
        big = v64 >= 10.0 * (4.0 / 3.0) * np.pi * 125.0  # ten reference nucleus volumes
        rel = float((drift[big] / v64[big]).max())
        report[s] = (emax, emean, float(drift.max()), rel)
        print(f"FP32 step {s}: max |dphi| {emax:.3e}, mean |dphi| {emean:.3e}, largest grain volume drift "
              f"{drift.max():.3f} cells ({rel:.2e} of the grain, grains >= 10 nucleus volumes)", flush=True)
        # every state sums to one per cell and stays in [0, 1]
        sums = np.add.reduceat(b.value, b.offsets()[:-1])
        assert np.abs(sums - 1.0).max() < 1e-12
        assert b.value.min() > 0.0 and b.value.max() <= 1.0
    assert report[1000][0] <= FP32_MAX_1000 and report[1000][1] <= FP32_MEAN_1000
    assert report[100][1] <= 1e-6
    assert e64.overflow_cell() == -1 and e32.overflow_cell() == -1
    e64.close()
    e32.close()
