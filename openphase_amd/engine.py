"""Engine: one device-resident multiphase-field state behind the C ABI (include/mpf.h).

This is the numpy-level handle used by the OpenPhase-style classes (openphase_amd.openphase), the
benchmark and the tests. All numerics run in libmpf_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import MpfError, MpfGrid, MpfParams, ptr
from .microstructure import Lists, storage_shape


@dataclass
class StepConfig:
    """Grid, interface and driving-force parameters of one simulation (the .opi keys of
    GridParameters, InterfaceProperties, PhaseField, UserDrivingForce and DrivingForce)."""

    nx: int
    ny: int
    nz: int
    dx: float = 1e-6
    iwidth: float = 5.0
    nphases: int = 1
    sigma: np.ndarray = field(default_factory=lambda: np.array([[0.24]]))
    mu: np.ndarray = field(default_factory=lambda: np.array([[4e-9]]))
    energy_model: np.ndarray | None = None  # MPF_ENERGY_ISO / MPF_ENERGY_CUBIC per phase pair
    epsilon: np.ndarray | None = None
    tj_model: int = _capi.MPF_TJ_MODEL
    tj_factor: float = 0.0
    tj_energy: float = 0.0
    consider_nucleus_volume: bool = True
    nucleus_volume_factor: float = 1.0
    dg_mode: np.ndarray | None = None  # 1 = VALUE per phase pair
    dg_value: np.ndarray | None = None
    dg_limit: float = 0.95
    # DrivingForce Average / Range / Threshold / WeightsMode (0 PHASEFIELDS, 1 RANGE, 2 COUNTER);
    # Range <= 0 and Threshold <= 0 select the reference defaults (IWidth; 1/3 for IWidth >= 5)
    dg_average: bool = False
    dg_range: int = 0
    dg_threshold: float = 0.0
    dg_weights: int = 0
    regularization_factor: float = 1.0
    # InterfaceRegularization: the steps take DoubleObstacle's curvature overload plus the module's
    # Average and MergePhaseFieldIncrements; InterfaceProperties CurvatureFactor and the module's
    # Average / Range / Threshold (Range <= 0: (IWidth + 1)/2; Threshold <= 0: DBL_EPSILON)
    regularization: bool = False
    curvature_factor: float = 1.0
    reg_average: bool = True
    reg_range: int = 0
    reg_threshold: float = 0.0
    precision: int = _capi.MPF_FP64
    slots: int = 8

    @property
    def shape(self) -> tuple:
        return storage_shape(self.nx, self.ny, self.nz)

    @property
    def active(self) -> int:
        return int(self.nx > 0) + int(self.ny > 0) + int(self.nz > 0)

    def max_time_step(self) -> float:
        """InterfaceProperties::ReportMaximumTimeStep (src/InterfaceProperties.cpp:523-556) from the
        phase-pair maxima of the tables."""
        me = float(np.max(self._max_energy()))
        mm = float(np.max(np.asarray(self.mu, dtype=np.float64)))
        f = {1: 0.50, 2: 0.25, 3: 0.16}[self.active]
        return f * self.dx * self.dx / (mm * me * self.regularization_factor)

    def _max_energy(self) -> np.ndarray:
        sig = np.asarray(self.sigma, dtype=np.float64)
        if self.energy_model is None:
            return sig
        eps = np.zeros_like(sig) if self.epsilon is None else np.asarray(self.epsilon, dtype=np.float64)
        return np.where(np.asarray(self.energy_model) == _capi.MPF_ENERGY_CUBIC, sig * (1.0 + eps * 1.41), sig)


class Engine:
    """Owns an mpf_ctx. Methods raise MpfError on any non-OK status."""

    def __init__(self, cfg: StepConfig):
        self.lib = _capi.load()
        self.cfg = cfg
        g = MpfGrid()
        g.nx, g.ny, g.nz = cfg.nx, cfg.ny, cfg.nz
        g.dx, g.iwidth = cfg.dx, cfg.iwidth
        for i in range(6):
            g.bc[i] = 0
        g.precision = cfg.precision
        g.slots = cfg.slots
        np2 = cfg.nphases * cfg.nphases
        self._keep = []  # arrays referenced by the params struct during mpf_create

        def arr(a, dtype):
            if a is None:
                return None
            x = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(np2))
            self._keep.append(x)
            return x

        p = MpfParams()
        p.nphases = cfg.nphases
        p.sigma = ptr(arr(cfg.sigma, np.float64), C.c_double)
        p.mu = ptr(arr(cfg.mu, np.float64), C.c_double)
        p.energy_model = ptr(arr(cfg.energy_model, np.int32), C.c_int32)
        p.epsilon = ptr(arr(cfg.epsilon, np.float64), C.c_double)
        p.tj_model, p.tj_factor, p.tj_energy = cfg.tj_model, cfg.tj_factor, cfg.tj_energy
        p.consider_nucleus_volume = int(cfg.consider_nucleus_volume)
        p.nucleus_volume_factor = cfg.nucleus_volume_factor
        p.dg_mode = ptr(arr(cfg.dg_mode, np.int32), C.c_int32)
        p.dg_value = ptr(arr(cfg.dg_value, np.float64), C.c_double)
        p.dg_limit = cfg.dg_limit
        p.dg_average = int(cfg.dg_average)
        p.dg_range = int(cfg.dg_range)
        p.dg_threshold = float(cfg.dg_threshold)
        p.dg_weights = int(cfg.dg_weights)
        p.regularization_factor = cfg.regularization_factor
        p.regularization = int(cfg.regularization)
        p.curvature_factor = float(cfg.curvature_factor)
        p.reg_average = int(cfg.reg_average)
        p.reg_range = int(cfg.reg_range)
        p.reg_threshold = float(cfg.reg_threshold)
        h = C.c_void_p()
        st = self.lib.mpf_create(C.byref(g), C.byref(p), C.byref(h))
        if st != _capi.MPF_OK:
            raise MpfError(st, "mpf_create")
        self.h = h
        self.ngrains = 0
        # phase-pair tables the context currently holds (set_pair_params replaces them)
        np2 = cfg.nphases * cfg.nphases
        self.pair_params = {
            "sigma": np.asarray(cfg.sigma, np.float64).reshape(np2).copy(),
            "mu": np.asarray(cfg.mu, np.float64).reshape(np2).copy(),
            "dg_value": None if cfg.dg_value is None else np.asarray(cfg.dg_value, np.float64).reshape(np2).copy(),
        }
        self._ensembles = weakref.WeakSet()  # ensembles that hold this context

    # -- lifecycle --------------------------------------------------------------------------
    def close(self) -> None:
        """Destroys the context; refused while an open Ensemble holds it (close that first)."""
        if getattr(self, "h", None):
            if any(e.h for e in getattr(self, "_ensembles", ())):
                raise RuntimeError("Engine.close: an open Ensemble holds this context; close the Ensemble first")
            self.lib.mpf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int, where: str) -> None:
        if st != _capi.MPF_OK:
            msg = self.lib.mpf_error_message(self.h)
            raise MpfError(st, where, msg.decode() if msg else "")

    # -- state ------------------------------------------------------------------------------
    def set_grains(self, phase, rotation=None, solid=None) -> None:
        phase = np.ascontiguousarray(phase, dtype=np.int32)
        rot = None if rotation is None else np.ascontiguousarray(rotation, dtype=np.float64).reshape(-1)
        sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8)
        self._check(self.lib.mpf_set_grains(self.h, phase.shape[0], ptr(phase, C.c_int32), ptr(rot, C.c_double),
                                            ptr(sol, C.c_uint8)), "mpf_set_grains")
        self.ngrains = phase.shape[0]

    def upload(self, lists: Lists) -> None:
        if tuple(lists.shape) != self.cfg.shape:
            raise ValueError(f"state shape {lists.shape} != grid {self.cfg.shape}")
        cnt = np.ascontiguousarray(lists.count, dtype=np.int32)
        idx = np.ascontiguousarray(lists.index, dtype=np.int32)
        val = np.ascontiguousarray(lists.value, dtype=np.float64)
        self._check(self.lib.mpf_upload(self.h, ptr(cnt, C.c_int32), ptr(idx, C.c_int32), ptr(val, C.c_double)),
                    "mpf_upload")

    def finalize_initialization(self) -> None:
        self._check(self.lib.mpf_finalize_initialization(self.h), "mpf_finalize_initialization")

    def step(self, dt: float, nsteps: int = 1) -> None:
        self._check(self.lib.mpf_step(self.h, float(dt), int(nsteps)), "mpf_step")

    def step_async(self, dt: float, nsteps: int = 1) -> None:
        self._check(self.lib.mpf_step_async(self.h, float(dt), int(nsteps)), "mpf_step_async")

    def synchronize(self) -> None:
        self._check(self.lib.mpf_synchronize(self.h), "mpf_synchronize")

    def download(self, with_flags: bool = False):
        tot = C.c_int64()
        self._check(self.lib.mpf_total_entries(self.h, C.byref(tot)), "mpf_total_entries")
        n = int(np.prod(self.cfg.shape))
        cnt = np.zeros(n, np.int32)
        idx = np.zeros(tot.value, np.int32)
        val = np.zeros(tot.value, np.float64)
        flg = np.zeros(n, np.int8) if with_flags else None
        self._check(self.lib.mpf_download(self.h, ptr(cnt, C.c_int32), ptr(idx, C.c_int32), ptr(val, C.c_double),
                                          ptr(flg, C.c_int8)), "mpf_download")
        lists = Lists(self.cfg.shape, cnt, idx, val)
        return (lists, flg) if with_flags else lists

    def grains(self):
        vol = np.zeros(self.ngrains)
        vr = np.zeros(self.ngrains)
        self._check(self.lib.mpf_get_grains(self.h, ptr(vol, C.c_double), ptr(vr, C.c_double)), "mpf_get_grains")
        return vol, vr

    def grain_records(self):
        """(MAXVolume, RefVolume, Exist, Stage) per grain."""
        mv, rv = np.zeros(self.ngrains), np.zeros(self.ngrains)
        ex, st = np.zeros(self.ngrains, np.uint8), np.zeros(self.ngrains, np.int32)
        self._check(self.lib.mpf_get_grain_records(self.h, ptr(mv, C.c_double), ptr(rv, C.c_double),
                                                   ptr(ex, C.c_uint8), ptr(st, C.c_int32)), "mpf_get_grain_records")
        return mv, rv, ex, st

    def max_time_step(self) -> float:
        """InterfaceProperties::ReportMaximumTimeStep from the maxima the last step's IP.Set saw
        (DBL_MAX before the first step)."""
        v = C.c_double()
        self._check(self.lib.mpf_max_time_step(self.h, C.byref(v)), "mpf_max_time_step")
        return v.value

    IP_STATES = {"none": _capi.MPF_IP_NONE, "current": _capi.MPF_IP_CURRENT, "previous": _capi.MPF_IP_PREVIOUS}

    def interface_energy(self, ip: str = "previous") -> float:
        """Sum over cells of DoubleObstacle::PointEnergy with the interface energies of the last
        IP.Set, evaluated on no state ("none"), the current state ("current") or the state the
        last step started from ("previous"). / cells = AverageEnergyDensity."""
        v = C.c_double()
        self._check(self.lib.mpf_interface_energy(self.h, self.IP_STATES[ip], C.byref(v)), "mpf_interface_energy")
        return v.value

    def average_energy_density(self, ip: str = "previous") -> float:
        """DoubleObstacle::AverageEnergyDensity (src/DoubleObstacle.cpp:989-1005)."""
        return self.interface_energy(ip) / float(np.prod(self.cfg.shape))

    def phase_fractions(self) -> np.ndarray:
        """PhaseField::Fractions computed on the device: [nphases, cells]."""
        out = np.zeros((self.cfg.nphases, int(np.prod(self.cfg.shape))))
        self._check(self.lib.mpf_phase_fractions(self.h, ptr(out, C.c_double), 1), "mpf_phase_fractions")
        return out

    def set_step_terms(self, driving_force: bool = True, average: bool = True) -> None:
        """Driving force in the following steps or not (the reference's driver calls
        UserDrivingForce::SetDrivingForce / DrivingForce::Average / MergePhaseFieldIncrements),
        averaged or not (DrivingForce::SkipAverage) when the configuration averages."""
        self._check(self.lib.mpf_set_step_terms(self.h, int(bool(driving_force)), int(bool(average))),
                    "mpf_set_step_terms")

    def limiter_warnings(self) -> int:
        w = C.c_int64()
        self._check(self.lib.mpf_limiter_warnings(self.h, C.byref(w)), "mpf_limiter_warnings")
        return int(w.value)

    def overflow_cell(self) -> int:
        w = C.c_int64()
        self._check(self.lib.mpf_overflow_cell(self.h, C.byref(w)), "mpf_overflow_cell")
        return int(w.value)

    def enable_timing(self, on: bool = True) -> None:
        self._check(self.lib.mpf_enable_timing(self.h, int(on)), "mpf_enable_timing")

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.mpf_last_timing(self.h, C.byref(a), C.byref(b)), "mpf_last_timing")
        return a.value, b.value

    def step_stats(self):
        """Cumulative (mid-kernel, wide-kernel) cell-steps since creation."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.lib.mpf_step_stats(self.h, C.byref(a), C.byref(b)), "mpf_step_stats")
        return a.value, b.value

    def copy_bandwidth(self, nbytes: int, repeats: int = 20) -> float:
        """Average time (ms) of the library's streaming copy kernel over `nbytes`."""
        ms = C.c_double()
        self._check(self.lib.mpf_copy_bandwidth(self.h, int(nbytes), int(repeats), C.byref(ms)), "mpf_copy_bandwidth")
        return ms.value

    def device_state(self):
        cnt, idx, val, stream = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        self._check(self.lib.mpf_device_state(self.h, C.byref(cnt), C.byref(idx), C.byref(val), C.byref(stream)),
                    "mpf_device_state")
        return cnt.value, idx.value, val.value, stream.value

    # ---------------------------------------------------------------- calibration / adjoint
    def set_pair_params(self, sigma=None, mu=None, dg_value=None) -> None:
        """Replaces the phase-pair tables (nphases x nphases); None keeps a table."""
        np2 = self.cfg.nphases * self.cfg.nphases

        def tab(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(np2)
            return a
        s, m, g = tab(sigma), tab(mu), tab(dg_value)
        self._keep = (s, m, g)
        self._check(self.lib.mpf_set_pair_params(self.h, ptr(s, C.c_double), ptr(m, C.c_double), ptr(g, C.c_double)),
                    "mpf_set_pair_params")
        for key, a in (("sigma", s), ("mu", m), ("dg_value", g)):
            if a is not None:
                self.pair_params[key] = a.copy()

    def state_bytes(self) -> int:
        n = C.c_int64()
        self._check(self.lib.mpf_state_bytes(self.h, C.byref(n)), "mpf_state_bytes")
        return n.value

    def save_state(self, device_ptr: int) -> None:
        """Copies the current state into device memory of state_bytes() bytes."""
        self._check(self.lib.mpf_save_state(self.h, C.c_void_p(device_ptr)), "mpf_save_state")

    def restore_state(self, device_ptr: int) -> None:
        self._check(self.lib.mpf_restore_state(self.h, C.c_void_p(device_ptr)), "mpf_restore_state")

    def adjoint_step(self, dt: float, lam_next: int, lam_prev: int, grad: int, next_count: int | None = None,
                     next_index: int | None = None) -> None:
        """One adjoint step from the current state (device pointers, see mpf_adjoint_step); next_count /
        next_index: S_{t+1}'s count bytes and grain indices (required for FP32 states)."""
        self._check(self.lib.mpf_adjoint_step(self.h, float(dt), C.c_void_p(lam_next), C.c_void_p(lam_prev),
                                              C.c_void_p(grad), C.c_void_p(next_count), C.c_void_p(next_index)),
                    "mpf_adjoint_step")

    # ---------------------------------------------------------------- launch shape
    def autotune(self, dt: float, steps: int = 3):
        """Times the candidate step-kernel tiles, then the launch shapes, on this state (restored
        afterwards) and keeps the fastest (mpf_autotune); returns (xcd_order, deferred_grid), the
        tile by tile_shape()."""
        x, g = C.c_int32(), C.c_int32()
        self._check(self.lib.mpf_autotune(self.h, float(dt), int(steps), C.byref(x), C.byref(g)), "mpf_autotune")
        return x.value, g.value

    def set_tile_shape(self, tile_x: int, tile_y: int = 0, tile_z: int = 0) -> None:
        """LDS-tiled step kernel shape (mpf_set_tile_shape); tile_x < 0 selects the per-cell kernel."""
        self._check(self.lib.mpf_set_tile_shape(self.h, int(tile_x), int(tile_y), int(tile_z)), "mpf_set_tile_shape")

    def tile_shape(self):
        x, y, z = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self.lib.mpf_get_tile_shape(self.h, C.byref(x), C.byref(y), C.byref(z)), "mpf_get_tile_shape")
        return x.value, y.value, z.value

    def set_launch_shape(self, xcd_order: int, deferred_grid: int) -> None:
        self._check(self.lib.mpf_set_launch_shape(self.h, int(xcd_order), int(deferred_grid)), "mpf_set_launch_shape")


class Ensemble:
    """Independent domains advanced together: two steps of every member per captured HIP graph
    (include/mpf.h, mpf_ensemble_*)."""

    def __init__(self, engines):
        self.engines = list(engines)
        self.lib = _capi.load()
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        st = self.lib.mpf_ensemble_create(arr, len(self.engines), C.byref(h))
        if st != _capi.MPF_OK:
            raise MpfError(st, "mpf_ensemble_create", "")
        self.h = h
        for e in self.engines:
            e._ensembles.add(self)

    def step(self, dt, nsteps: int) -> None:
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(self.engines),)))
        st = self.lib.mpf_ensemble_step(self.h, ptr(d, C.c_double), int(nsteps))
        if st != _capi.MPF_OK:
            own = self.lib.mpf_ensemble_error_message(self.h)
            msgs = [own] + [e.lib.mpf_error_message(e.h) for e in self.engines]
            raise MpfError(st, "mpf_ensemble_step", "; ".join(m.decode() for m in msgs if m))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mpf_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
