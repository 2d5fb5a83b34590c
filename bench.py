#!/usr/bin/env python3
"""Benchmark of the MI355X multiphase-field step.

Workload (N=1): 3D normal grain growth on a 256^3 periodic grid, 1000 Voronoi grains
(uniform-random seeds, one-cell overlap band as the reference's Voronoi set-up), isotropic energy
0.24 J/m^2, mobility 4e-9 m^4/(J s), interface width 5 cells, 16 phase slots per cell (a few cells
hold 9 phase fields between steps 4 and 23, so 8 slots overflow), FP64 (the plain drop-in step),
dt at the stability limit. A "step" is one full MPF time step of the whole grid (IP.Set +
increments + normalize + merge + finalize + grain volumes). The same step in FP32
storage/stencil arithmetic is measured beside it, with its error against the FP64 step.

Prints one JSON line: value = cell-updates/s of the FP64 step (whole job, inputs resident in HBM),
plus `roofline` for the fused step kernel (HBM bound: algorithmic bytes / average kernel time from
HIP events on the library's stream) and `cpu_baseline` (the reference OpenPhase build, timed on
this host on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
NX = NY = NZ = 256
NGRAINS = 1000
SEED = 0x0FE2
# stored-list capacity: 8 overflows between steps 4 and 23 of this 3D case (9 grains meet in a few
# cells); only occupied slots are read or written, so the capacity does not change the traffic
SLOTS = 16
SLOTS_FP32 = 15   # the FP32 count byte holds the list length in 4 bits beside the implicit slot


def _device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def algorithmic_bytes(count_in: np.ndarray, count_out: np.ndarray, value_bytes: int) -> float:
    """Bytes one step must move at minimum: read every cell's list once and write the next one
    once: per cell 1 B count + n entries x (2 B index + value) on each side."""
    return float(count_in.size * 2 + (count_in.sum() + count_out.sum()) * (2 + value_bytes))


# Step-kernel tile (planes along x, then the y x z tile) and launch shape (XCD-contiguous block
# order, deferred-cell grid) for this workload on MI355X: the tile mpf_autotune picks, with the
# XCD-contiguous order (as fast as the other order here, with far fewer L2 misses); fixed so that
# runs and profiles use the same kernel shape
TUNED = {0: ((8, 16, 16), (1, 4096)), 1: ((4, 16, 16), (1, 4096))}


def run_gpu(precision: int, lists, steps: int, warmup: int):
    from openphase_amd.engine import Engine, StepConfig

    cfg = StepConfig(NX, NY, NZ, precision=precision, slots=SLOTS if precision == 0 else SLOTS_FP32)
    dt = cfg.max_time_step()
    eng = Engine(cfg)
    eng.set_grains(np.zeros(NGRAINS, np.int32))
    eng.upload(lists)
    eng.finalize_initialization()
    tile, launch = TUNED[precision]
    eng.set_tile_shape(*tile)
    eng.set_launch_shape(*launch)
    if warmup:
        eng.step(dt, warmup)
    c_in = eng.download().count.copy()
    eng.enable_timing(False)
    _device_sync()
    t0 = time.perf_counter()
    eng.step_async(dt, steps)
    eng.synchronize()
    _device_sync()
    wall = time.perf_counter() - t0
    t1 = time.perf_counter()
    st_out = eng.download()
    t_down = time.perf_counter() - t1
    c_out = st_out.count
    up = Engine(cfg)
    up.set_grains(np.zeros(NGRAINS, np.int32))
    _device_sync()
    t1 = time.perf_counter()
    up.upload(st_out)
    _device_sync()
    t_up = time.perf_counter() - t1
    up.close()
    # second pass with per-kernel HIP events on the library's stream (same number of steps)
    eng.enable_timing(True)
    eng.step(dt, steps)
    k_step_ms, k_rest_ms = eng.last_timing()
    d_mid, d_wide = eng.step_stats()
    total_steps = warmup + 2 * steps
    eng.enable_timing(False)
    copy_bytes = 1 << 30
    copy_ms = eng.copy_bandwidth(copy_bytes, 10)
    eng.close()
    vbytes = 8 if precision == 0 else 4
    b_alg = algorithmic_bytes(c_in, c_out, vbytes)
    ncell = NX * NY * NZ
    return {
        "cell_updates_per_s": ncell * steps / wall,
        "ms_per_step": 1e3 * wall / steps,
        "kernel_ms": k_step_ms,
        "tile_and_launch_shape": [list(tile), list(launch)],
        "deferred_and_grain_ms": k_rest_ms,
        "bytes_per_step": b_alg,
        "mean_entries_per_cell": float((c_in.sum() + c_out.sum()) / (2 * ncell)),
        "achieved_gbs": b_alg / (k_step_ms * 1e-3) / 1e9 if k_step_ms > 0 else None,
        "copy_gbs": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9,
        "deferred_cells_per_step": ((d_mid - d_wide) / total_steps, d_wide / total_steps),
        # host buffers cross PCIe only in upload / download (the state as lists, host-side packing
        # included); a caller that hands over host lists every step would see this rate
        "upload_ms": 1e3 * t_up,
        "download_ms": 1e3 * t_down,
        "pcie_inclusive_cell_updates_per_s": ncell / (t_up + wall / steps + t_down),
    }


PROFILE = ROOT / "profiles" / "mpf_step_256cube_1000grains.md"
MAIN_KERNEL = "tile_kernel<double, double, false, false, true>"


def profiled_traffic(kernel: str = MAIN_KERNEL):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary (FETCH_SIZE and
    WRITE_SIZE in KiB per dispatch, scaled by the factors the copy kernel gives: 2.00 and 1.00).
    None when the summary does not list the kernel."""
    path = PROFILE
    if not path.exists():
        return None
    fetch = write = None
    for line in path.read_text().splitlines():
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) >= 3 and cols[0] == f"void {kernel}":
            if cols[1] == "FETCH_SIZE":
                fetch = float(cols[2])
            elif cols[1] == "WRITE_SIZE":
                write = float(cols[2])
    if fetch is None or write is None:
        return None
    return fetch * 1024 * 2.0 + write * 1024 * 1.0


def fraction_errors(a, b, ngrains: int):
    """Max and mean |phi_a - phi_b| over every (cell, grain) entry present in either state."""
    def keys(x):
        cell = np.repeat(np.arange(x.ncells, dtype=np.int64), x.count)
        return cell * ngrains + x.index.astype(np.int64)
    k = np.concatenate([keys(a), keys(b)])
    v = np.concatenate([a.value, -b.value])
    _, inv = np.unique(k, return_inverse=True)
    d = np.abs(np.bincount(inv, weights=v))
    return float(d.max()), float(d.mean())


def run_fp32_accuracy(lists, checks=(100, 500, 1000)):
    """The FP32 variant against the FP64 step (bit-equal to the CPU oracle) on the benchmark
    workload: max / mean phase-fraction error over all (cell, grain) entries and the largest grain
    volume drift (cells, and relative for grains of at least ten nucleus volumes) at each check."""
    from openphase_amd.engine import Engine, StepConfig

    c64 = StepConfig(NX, NY, NZ, slots=SLOTS)
    c32 = StepConfig(NX, NY, NZ, precision=1, slots=SLOTS_FP32)
    dt = c64.max_time_step()
    e64, e32 = Engine(c64), Engine(c32)
    for e in (e64, e32):
        e.set_grains(np.zeros(NGRAINS, np.int32))
        e.upload(lists)
        e.finalize_initialization()
    out, done = {}, 0
    big_volume = 10.0 * (4.0 / 3.0) * np.pi * 125.0
    for s in checks:
        for e in (e64, e32):
            e.step(dt, s - done)
        done = s
        emax, emean = fraction_errors(e64.download(), e32.download(), NGRAINS)
        v64, _ = e64.grains()
        v32, _ = e32.grains()
        drift = np.abs(v32 - v64)
        big = v64 >= big_volume
        out[f"step_{s}"] = {"max_phase_fraction_error": emax, "mean_phase_fraction_error": emean,
                            "max_grain_volume_drift_cells": float(drift.max()),
                            "max_grain_volume_drift_rel": float((drift[big] / v64[big]).max())}
    e64.close()
    e32.close()
    return out


def run_adjoint(steps: int, segment: int, precision: int):
    """Differentiable rollout of the calibration workload: 2D 512^2 periodic, 200 grains of two
    phases (random half/half), uniform random cubic orientations, cubic anisotropy 0.05,
    sigma 0.24, mu 4e-9, constant driving force 1e5 J/m^3 between the phases, dt = 0.9 x the 2D
    limit, FP32 or FP64 states. Forward `steps` steps with a checkpoint every `segment` steps,
    mismatch loss against a target grown with tables perturbed by exp(0.1 N(0,1)), then the
    adjoint pass (segment recomputation included). Returns (forward + adjoint) steps / wall
    second, the forward and backward rates apart, and the adjoint step's own device time."""
    import torch
    from openphase_amd import _capi
    from openphase_amd.adjoint import DeviceStates, dense_fields, mismatch_loss, rollout
    from openphase_amd.engine import Engine, StepConfig
    from openphase_amd.microstructure import random_rotations, voronoi_grains

    n, ng = 512, 200
    cfg = StepConfig(n, n, 0, nphases=2, sigma=np.full((2, 2), 0.24), mu=np.full((2, 2), 4e-9),
                     energy_model=np.full((2, 2), _capi.MPF_ENERGY_CUBIC, np.int32), epsilon=np.full((2, 2), 0.05),
                     dg_mode=np.array([[0, 1], [1, 0]], np.int32), dg_value=np.array([[0.0, 1e5], [-1e5, 0.0]]),
                     slots=8, precision=precision)
    dt = 0.9 * cfg.max_time_step()
    rng = np.random.default_rng(0x0FE4)
    phase = np.zeros(ng, np.int32)
    phase[rng.permutation(ng)[: ng // 2]] = 1
    _, rot = random_rotations(ng, seed=0x0FE4)
    eng = Engine(cfg)
    eng.set_grains(phase, rot)
    eng.upload(voronoi_grains(n, n, 0, ng, seed=0x0FE4))
    eng.finalize_initialization()
    st = DeviceStates(eng)
    init = st.save()
    p0 = np.stack([cfg.sigma, cfg.mu, cfg.dg_value])
    f = np.exp(0.1 * np.random.default_rng(0x0FE5).standard_normal(p0.shape))
    f[0] = np.sqrt(f[0] * f[0].T)
    f[1] = np.sqrt(f[1] * f[1].T)
    f[2] = np.triu(f[2]) + np.triu(f[2], 1).T
    with torch.no_grad():
        v, c, i = rollout(st, torch.tensor(p0 * f), init, steps, dt, segment=steps)
        target = dense_fields(v, c, i, ng)
    params = torch.tensor(p0, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v, c, i = rollout(st, params, init, steps, dt, segment=segment)
    loss = mismatch_loss(v, c, i, target)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # device time of one forward step and one adjoint step at the rollout's final state
    fin = st.save()
    st.restore(fin)
    lam = torch.randn(st.slots, st.ncells, dtype=torch.float64, device="cuda")
    prev = torch.empty_like(lam)
    grad = torch.zeros(12, dtype=torch.float64, device="cuda")
    eng.step(dt, 1)
    nxt = st.save()
    nc, ni = st.list_pointers(nxt)
    st.restore(fin)
    reps = 20
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    eng.step(dt, reps)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    st.restore(fin)
    for _ in range(reps):
        eng.adjoint_step(dt, lam.data_ptr(), prev.data_ptr(), grad.data_ptr(), nc, ni)
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    eng.close()
    fwd_ms, adj_ms = 1e3 * (t4 - t3) / reps, 1e3 * (t5 - t4) / reps
    return {
        "steps_per_s": 2 * steps / (t2 - t0),
        "forward_steps_per_s": steps / (t1 - t0),
        "backward_steps_per_s": steps / (t2 - t1),
        "forward_step_ms": fwd_ms,
        "adjoint_step_ms": adj_ms,
        "adjoint_vs_forward_speed": fwd_ms / adj_ms,
        "steps": steps,
        "segment": segment,
        "precision": "fp32" if precision else "fp64",
        "loss": float(loss.detach()),
        "grad": [float(x) for x in params.grad.reshape(-1)],
        "workload": "2D 512^2 periodic, 200 grains, 2 phases, cubic anisotropy 0.05, driving force 1e5 J/m^3, "
                    "forward + adjoint with segment checkpointing (steps/s = (forward + adjoint steps) / wall; "
                    "the backward pass recomputes each segment's states from its checkpoint); step times at the "
                    "rollout's final state",
    }


def run_ensemble(steps: int):
    """Calibration ensemble: 64 periodic 2D domains, side round(128 * 4^U) with U ~ U[0,1] (128 to
    512), 20-300 grains (log-uniform), sigma ~ U[0.1, 0.5] J/m^2 and mu ~ logU[1e-9, 1e-8] per
    domain, FP32, dt = 0.9 x each domain's limit; seed 0x0FE6. Each member is autotuned, then the
    ensemble (one captured graph per two steps of all members) and the same domains stepped one
    after another are timed. Returns total cell-updates/s of both."""
    from openphase_amd.engine import Engine, Ensemble, StepConfig
    from openphase_amd.microstructure import calibration_domains, voronoi_grains

    engines, dts, cells = [], [], 0
    for d in calibration_domains(64, 0x0FE6):
        side, ng = d.side, d.ngrains
        cfg = StepConfig(side, side, 0, sigma=np.array([[d.sigma]]), mu=np.array([[d.mu]]), precision=1, slots=8)
        e = Engine(cfg)
        e.set_grains(np.zeros(ng, np.int32))
        e.upload(voronoi_grains(side, side, 0, ng, seed=d.seed))
        e.finalize_initialization()
        dt = 0.9 * cfg.max_time_step()
        e.step(dt, 2)
        e.autotune(dt, steps=2)
        engines.append(e)
        dts.append(dt)
        cells += side * side
    ens = Ensemble(engines)
    ens.step(dts, 2)
    _device_sync()
    t0 = time.perf_counter()
    ens.step(dts, steps)
    _device_sync()
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        for e, dt in zip(engines, dts):
            e.step(dt, 1)
    _device_sync()
    t_seq = time.perf_counter() - t0
    adj = run_ensemble_adjoint(ens, engines, dts, steps, cells)
    tiles = sorted({str(e.tile_shape()) for e in engines})
    ens.close()
    for e in engines:
        e.close()
    return {
        "cell_updates_per_s": cells * steps / t_graph,
        "sequential_cell_updates_per_s": cells * steps / t_seq,
        "domains": 64,
        "cells": cells,
        "steps": steps,
        "tiles": tiles,
        "adjoint": adj,
        "workload": "64 periodic 2D domains, sides 128-512, 20-300 grains, per-domain sigma/mu/dt, FP32, every "
                    "member autotuned (step-kernel tile, launch shape); one captured graph per two steps of all "
                    "members vs the domains stepped one by one",
    }


def run_ensemble_adjoint(ens, engines, dts, steps: int, cells: int, segment: int = 50):
    """Differentiable rollout of the whole ensemble (ensemble_rollout): `steps` steps of all
    members together with a checkpoint of every member each `segment` steps, the sum of the
    members' mismatch losses against their starting fields, and the reverse sweep of every member
    (recomputation included). Returns the wall times and rates."""
    import torch
    from openphase_amd.adjoint import DeviceStates, dense_fields, ensemble_rollout, mismatch_loss

    states = [DeviceStates(e) for e in engines]
    inits = [st.save() for st in states]
    targets = []
    for st, init, e in zip(states, inits, engines):
        c, i, v = st.views(init)
        targets.append(dense_fields(v.clone(), c.clone(), i.clone(), e.ngrains))
    p0 = np.stack([np.stack([e.pair_params["sigma"].reshape(1, 1), e.pair_params["mu"].reshape(1, 1), np.zeros((1, 1))])
                   for e in engines])
    params = torch.tensor(p0, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = ensemble_rollout(ens, states, params, inits, steps, dts, segment=segment)
    loss = sum(mismatch_loss(v, c, i, t) for (v, c, i), t in zip(outs, targets))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {
        "forward_s": t1 - t0,
        "backward_s": t2 - t1,
        "cell_updates_per_s": 2 * cells * steps / (t2 - t0),
        "backward_cell_steps_per_s": cells * steps / (t2 - t1),
        "segment": segment,
        "loss": float(loss.detach()),
        "grad_abs_max": float(params.grad.abs().max()),
        "workload": "forward with checkpoints (all members together) + reverse sweep per member (segment "
                    "recomputation included); cell_updates_per_s = (forward + adjoint cell-steps) / wall",
    }


def run_dg_average(steps: int):
    """Driving force averaged over the interface (DrivingForce::Average, PHASEFIELDS weights,
    Range 5, Threshold 1/3) on the calibration microstructure: 2D 512^2 periodic, 200 grains of two
    phases, isotropic sigma 0.24 / mu 4e-9, constant driving force 1e5 J/m^3, dt = 0.9 x the 2D
    limit. Cell-updates/s with and without averaging, FP64 and FP32 (wall time of `steps` steps)."""
    from openphase_amd.engine import Engine, StepConfig
    from openphase_amd.microstructure import voronoi_grains

    n, ng = 512, 200
    rng = np.random.default_rng(0x0FE4)
    phase = np.zeros(ng, np.int32)
    phase[rng.permutation(ng)[: ng // 2]] = 1
    lists = voronoi_grains(n, n, 0, ng, seed=0x0FE4)
    out = {"steps": steps,
           "workload": "2D 512^2 periodic, 200 grains, 2 phases, driving force 1e5 J/m^3, averaged over "
                       "the interface (Range 5, PHASEFIELDS weights) vs not averaged"}
    for prec, tag in ((0, "fp64"), (1, "fp32")):
        for avg in (False, True):
            cfg = StepConfig(n, n, 0, nphases=2, sigma=np.full((2, 2), 0.24), mu=np.full((2, 2), 4e-9),
                             dg_mode=np.array([[0, 1], [1, 0]], np.int32),
                             dg_value=np.array([[0.0, 1e5], [-1e5, 0.0]]), dg_average=avg, precision=prec, slots=8)
            dt = 0.9 * cfg.max_time_step()
            e = Engine(cfg)
            e.set_grains(phase)
            e.upload(lists)
            e.finalize_initialization()
            e.step(dt, 5)
            _device_sync()
            t0 = time.perf_counter()
            e.step(dt, steps)
            _device_sync()
            t = time.perf_counter() - t0
            e.close()
            out[f"{tag}_{'averaged' if avg else 'plain'}_cell_updates_per_s"] = n * n * steps / t
    return out


def run_regularization(steps: int, lists3d):
    """Interface regularization (DoubleObstacle's curvature overload + InterfaceRegularization
    Average / MergePhaseFieldIncrements, RegularizationFactor 1.5, CurvatureFactor 0.7, default
    Range 3 / Threshold) against the plain step, FP64: 2D 512^2 with 200 grains and the 256^3 /
    1000-grain microstructure of the main leg. Cell-updates/s and ms per step from the wall time of
    `steps` steps."""
    from openphase_amd.engine import Engine, StepConfig
    from openphase_amd.microstructure import voronoi_grains

    out = {"steps": steps,
           "workload": "regularized vs plain FP64 step (RegularizationFactor 1.5, CurvatureFactor 0.7, Range 3, "
                       "curvature averaged): 2D 512^2 / 200 grains and 3D 256^3 / 1000 grains, periodic, slots 16 "
                       "(the regularized 256^3 run outgrows 8 slots within 43 steps)"}
    for tag, (nx, ny, nz, ng, lists) in (("2d_512", (512, 512, 0, 200, None)), ("3d_256", (256, 256, 256, 1000, lists3d))):
        if lists is None:
            lists = voronoi_grains(nx, ny, nz, ng, seed=0x0FE5)
        ncells = nx * ny * max(nz, 1)
        for reg in (False, True):
            cfg = StepConfig(nx, ny, nz, slots=16, regularization=reg, regularization_factor=1.5 if reg else 1.0,
                             curvature_factor=0.7 if reg else 1.0)
            dt = 0.5 * cfg.max_time_step()
            e = Engine(cfg)
            e.set_grains(np.zeros(ng, np.int32))
            e.upload(lists)
            e.finalize_initialization()
            e.step(dt, 3)
            _device_sync()
            t0 = time.perf_counter()
            e.step(dt, steps)
            _device_sync()
            t = time.perf_counter() - t0
            e.close()
            out[f"{tag}_{'regularized' if reg else 'plain'}_cell_updates_per_s"] = ncells * steps / t
            out[f"{tag}_{'regularized' if reg else 'plain'}_ms_per_step"] = 1e3 * t / steps
    return out


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(lists, steps: int = 3):
    """Times the reference OpenPhase build (oracle/_ref/ref_driver, compiled from the reference
    sources) on this host: same 256^3 state, one warm-up step then `steps` timed steps with the
    reference's four per-call stamps (examples/NormalGG/NormalGG.cpp:76-84), on every core of this
    process's CPU share (at most 16, the share of one GPU on this machine)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from openphase_amd.engine import StepConfig

    if not O.REF_DRIVER.exists():
        return None
    avail = len(os.sched_getaffinity(0))
    threads = min(16, avail)
    cfg = StepConfig(NX, NY, NZ)
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        timing, _ = O.run_reference(td, cfg, lists, np.zeros(NGRAINS, np.int32), None, steps + 1, cfg.max_time_step(), [],
                                    threads=threads, timeout=600, untimed=1)
        wall = time.perf_counter() - t0
    per = {k: timing[k] / steps for k in ("IP.Set", "CalculatePhaseFieldIncrements", "NormalizeIncrements",
                                          "MergeIncrements")}
    return {
        "value": timing["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "reference",
        "nproc": os.cpu_count(),
        "cpu_share": avail,
        "cpu_model": _cpu_model(),
        "seconds_per_step": {k: round(v, 4) for k, v in per.items()},
        "sample": f"OpenPhase reference build, same 256^3 1000-grain state, 1 warm-up + {steps} timed steps "
                  f"(IP.Set / Increments / Normalize / Merge stamps), {threads} OpenMP threads; "
                  f"{timing['total_s'] / steps:.2f} s per step, {wall:.1f} s incl. file I/O, set-up and warm-up",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp32-accuracy", action="store_true", help="skip the FP32-vs-FP64 error leg")
    ap.add_argument("--only-main", action="store_true",
                    help="only the 256^3 FP64 / FP32 step legs (for profiling runs)")
    ap.add_argument("--only-adjoint", action="store_true",
                    help="only the adjoint leg, FP64 then FP32 (for profiling runs); prints its JSON")
    ap.add_argument("--adjoint-steps", type=int, default=500, help="rollout length of the adjoint leg")
    ap.add_argument("--ensemble-steps", type=int, default=200, help="steps of the ensemble leg")
    ap.add_argument("--dg-steps", type=int, default=200, help="steps of the driving-force averaging leg")
    ap.add_argument("--reg-steps", type=int, default=40, help="steps of the interface regularization leg")
    ap.add_argument("--only-regularization", action="store_true",
                    help="only the interface regularization leg (for profiling runs); prints its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the step on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented).", file=sys.stderr)
        sys.exit(2)

    from openphase_amd.microstructure import voronoi_grains

    if args.only_adjoint:
        seg = min(50, args.adjoint_steps)
        print(json.dumps({"fp64": run_adjoint(args.adjoint_steps, seg, 0),
                          "fp32": run_adjoint(args.adjoint_steps, seg, 1)}))
        return
    t0 = time.perf_counter()
    lists = voronoi_grains(NX, NY, NZ, NGRAINS, seed=SEED, sides="plus")
    print(f"[bench] microstructure {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    if args.only_regularization:
        print(json.dumps(run_regularization(args.reg_steps, lists)))
        return
    r64 = run_gpu(0, lists, args.steps, args.warmup)
    print(f"[bench] fp64 {r64}", file=sys.stderr, flush=True)
    r32 = run_gpu(1, lists, args.steps, args.warmup)
    print(f"[bench] fp32 {r32}", file=sys.stderr, flush=True)
    acc32 = None if args.no_fp32_accuracy else run_fp32_accuracy(lists)
    print(f"[bench] fp32 accuracy {acc32}", file=sys.stderr, flush=True)
    adj = ens = dga = cpu = reg = None
    if not args.only_main:
        adj = {"fp32": run_adjoint(args.adjoint_steps, 50, 1)}
        print(f"[bench] adjoint fp32 {adj['fp32']}", file=sys.stderr, flush=True)
        adj["fp64"] = run_adjoint(args.adjoint_steps, 50, 0)
        print(f"[bench] adjoint fp64 {adj['fp64']}", file=sys.stderr, flush=True)
        ens = run_ensemble(args.ensemble_steps)
        print(f"[bench] ensemble {ens}", file=sys.stderr, flush=True)
        dga = run_dg_average(args.dg_steps)
        print(f"[bench] driving-force averaging {dga}", file=sys.stderr, flush=True)
        reg = run_regularization(args.reg_steps, lists)
        print(f"[bench] interface regularization {reg}", file=sys.stderr, flush=True)
    if not (args.no_cpu or args.only_main):
        cpu = cpu_baseline(lists)
    print(f"[bench] cpu {cpu}", file=sys.stderr, flush=True)

    ach = r64["achieved_gbs"]
    traffic = profiled_traffic()
    out = {
        "metric": "cell-updates/s + HBM GB/s, 256³ MPF step FP32 vs FP64; adjoint steps/s",
        "value": r64["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded periodic Voronoi microstructure, 1000 grains)",
        "config": {
            "workload": "3D normal grain growth, 256^3 periodic, 1000 Voronoi grains, isotropic "
                        "sigma=0.24 J/m^2, mu=4e-9, iWidth=5, 16 slots, FP64 plain step",
            "step_kernel": MAIN_KERNEL,
            "cells": NX * NY * NZ,
            "grains": NGRAINS,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (ach / HBM_PEAK_GBS) if ach else None,
            "traffic": traffic,
            # HBM bytes the profiled launch actually moved, over this run's kernel time: how far the
            # kernel is from the bandwidth bound with its real traffic (latency-bound when low)
            "traffic_gbs": (traffic / (r64["kernel_ms"] * 1e-3) / 1e9) if traffic else None,
            "traffic_frac": (traffic / (r64["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "kernel": MAIN_KERNEL,
            "profile": str(PROFILE.relative_to(ROOT)),
        },
        "fp32": {
            "cell_updates_per_s": r32["cell_updates_per_s"],
            "ms_per_step": r32["ms_per_step"],
            "achieved_gbs": r32["achieved_gbs"],
            "frac": (r32["achieved_gbs"] / HBM_PEAK_GBS) if r32["achieved_gbs"] else None,
            "speedup_vs_fp64": r32["cell_updates_per_s"] / r64["cell_updates_per_s"],
            "error_vs_fp64": acc32,
        },
        "kernel": {
            "fp64_step_kernel_ms": r64["kernel_ms"],
            "fp64_deferred_and_grain_ms": r64["deferred_and_grain_ms"],
            "fp64_bytes_per_step": r64["bytes_per_step"],
            "fp64_mean_entries_per_cell": r64["mean_entries_per_cell"],
            "fp32_step_kernel_ms": r32["kernel_ms"],
            "fp32_bytes_per_step": r32["bytes_per_step"],
            "copy_kernel_gbs": r64["copy_gbs"],
        },
        "host_transfer": {
            "fp64_upload_ms": r64["upload_ms"],
            "fp64_download_ms": r64["download_ms"],
            "fp64_pcie_inclusive_cell_updates_per_s": r64["pcie_inclusive_cell_updates_per_s"],
            "note": "state handed over as host lists before and after every step (mpf_upload / mpf_download, "
                    "host-side packing included); value is device-resident",
        },
        "adjoint": adj,
        "ensemble": ens,
        "dg_average": dga,
        "regularization": reg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
