"""Differentiable multiphase-field rollouts: a torch.autograd.Function over the step and its adjoint.

`rollout(engine, params, initial, nsteps, dt, segment)` advances the state saved in `initial`
(FP64 or FP32 storage, as the engine's) by `nsteps` steps with the phase-pair tables `params` =
[sigma, mu, driving force] (float64, shape [3, nphases, nphases]) and returns the final state's
values (float64, [slots, ncells], slot-major, on the GPU) with its counts and grain indices.
Gradients of any loss of those values flow back to `params` through the hand-written adjoint of
the step (mpf_adjoint_step): the forward pass keeps one checkpoint per `segment` steps; the
backward pass recomputes the states of each segment from its checkpoint and runs the adjoint
steps through it in reverse (segment checkpointing: memory ~ nsteps/segment + segment states).
FP32 rollouts are differentiated at their FP32 states with FP64 arithmetic.

The reference has no derivatives; the forward map is the reference's step (src/DoubleObstacle.cpp,
src/PhaseField.cpp). Gradients are those of the branches the forward steps take (clamps, erasures
and limiter factors as evaluated, VolumeRatios fixed).
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import Engine


def state_layout(ncells: int, slots: int, value_bytes: int, ngrains: int) -> dict:
    """Byte offsets of a saved state (mpf_save_state): count, index and value planes, then the
    grain arrays, each part 256-byte aligned."""
    def up(x):
        return (x + 255) & ~255
    lay = {"cnt": 0}
    lay["id"] = up(ncells)
    lay["val"] = up(lay["id"] + slots * ncells * 2)
    lay["volume"] = up(lay["val"] + slots * ncells * value_bytes)
    lay["maxvolume"] = up(lay["volume"] + ngrains * 8)
    lay["vr"] = up(lay["maxvolume"] + ngrains * 8)
    lay["exist"] = up(lay["vr"] + ngrains * 8)
    lay["stage"] = up(lay["exist"] + ngrains)
    lay["total"] = up(lay["stage"] + ngrains * 4)
    return lay


class DeviceStates:
    """Saved states of one Engine as torch uint8 tensors on the GPU."""

    def __init__(self, engine: Engine):
        self.engine = engine
        self.fp32 = engine.cfg.precision != 0
        self.nbytes = engine.state_bytes()
        n = int(np.prod(engine.cfg.shape))
        self.ncells, self.slots = n, engine.cfg.slots
        self.layout = state_layout(n, self.slots, 4 if self.fp32 else 8, engine.ngrains)
        if self.layout["total"] != self.nbytes:
            raise RuntimeError("saved-state layout mismatch")

    def save(self) -> torch.Tensor:
        blob = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        self.engine.save_state(blob.data_ptr())
        return blob

    def restore(self, blob: torch.Tensor) -> None:
        torch.cuda.synchronize()
        self.engine.restore_state(blob.data_ptr())

    def views(self, blob: torch.Tensor):
        """(counts uint8 [N], grain indices int32 [slots, N], values [slots, N]): views into an FP64
        blob; for FP32 storage decoded copies - counts without the implicit-slot bits, values in
        float64 with single-entry lists (nothing stored) at 1."""
        L, n, s = self.layout, self.ncells, self.slots
        cnt = blob[L["cnt"]:L["cnt"] + n]
        ids = blob[L["id"]:L["id"] + 2 * s * n].view(torch.int16).view(s, n).to(torch.int32) & 0xFFFF
        if not self.fp32:
            val = blob[L["val"]:L["val"] + 8 * s * n].view(torch.float64).view(s, n)
            return cnt, ids, val
        cnt = cnt & 15
        val = blob[L["val"]:L["val"] + 4 * s * n].view(torch.float32).view(s, n).to(torch.float64)
        val[0] = torch.where(cnt == 1, torch.ones_like(val[0]), val[0])
        return cnt, ids, val

    def list_pointers(self, blob: torch.Tensor):
        """Device addresses of a saved state's count bytes and grain indices (mpf_adjoint_step)."""
        return blob.data_ptr() + self.layout["cnt"], blob.data_ptr() + self.layout["id"]


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, states, initial, nsteps, dt, segment):
        eng = states.engine
        p = params.detach().to("cpu", torch.float64).numpy()
        eng.set_pair_params(p[0], p[1], p[2] if eng.cfg.dg_mode is not None else None)
        states.restore(initial)
        checkpoints = []
        done = 0
        while done < nsteps:
            checkpoints.append(states.save())
            n = min(segment, nsteps - done)
            eng.step(dt, n)
            done += n
        final = states.save()
        cnt, ids, val = states.views(final)
        val, cnt = val.clone(), cnt.clone()
        ctx.states, ctx.checkpoints, ctx.final = states, checkpoints, final
        ctx.params = p.copy()
        ctx.nsteps, ctx.dt, ctx.segment = nsteps, dt, segment
        ctx.nphases = eng.cfg.nphases
        ctx.mark_non_differentiable(cnt, ids)
        return val, cnt, ids

    @staticmethod
    def backward(ctx, g_val, g_cnt, g_ids):
        states, eng = ctx.states, ctx.states.engine
        # the recomputation and the adjoint steps run with this rollout's tables, whatever the
        # engine holds now (another rollout may have run since); the caller's tables come back after
        held = {k: (None if v is None else v.copy()) for k, v in eng.pair_params.items()}
        p = ctx.params
        eng.set_pair_params(p[0], p[1], p[2] if eng.cfg.dg_mode is not None else None)
        try:
            return _Rollout._backward(ctx, states, eng, g_val)
        finally:
            eng.set_pair_params(held["sigma"], held["mu"], held["dg_value"])

    @staticmethod
    def _backward(ctx, states, eng, g_val):
        npp = ctx.nphases * ctx.nphases
        grad = torch.zeros(3 * npp, dtype=torch.float64, device="cuda")
        lam = torch.zeros(states.slots, states.ncells, dtype=torch.float64, device="cuda")
        if g_val is not None:
            lam.copy_(g_val)
        nseg = len(ctx.checkpoints)
        prev = torch.empty_like(lam)
        for si in reversed(range(nseg)):
            t0 = si * ctx.segment
            n = min(ctx.segment, ctx.nsteps - t0)
            states.restore(ctx.checkpoints[si])
            seg_states = [ctx.checkpoints[si]]
            for _ in range(n - 1):
                eng.step(ctx.dt, 1)
                seg_states.append(states.save())
            seg_states.append(ctx.checkpoints[si + 1] if si + 1 < nseg else ctx.final)
            for i in reversed(range(n)):
                states.restore(seg_states[i])
                nc, ni = states.list_pointers(seg_states[i + 1])
                torch.cuda.synchronize()
                eng.adjoint_step(ctx.dt, lam.data_ptr(), prev.data_ptr(), grad.data_ptr(), nc, ni)
                lam, prev = prev, lam
        ctx.lam0 = lam
        g = grad.view(3, ctx.nphases, ctx.nphases).cpu()
        return g, None, None, None, None, None


def rollout(engine_states: DeviceStates, params: torch.Tensor, initial: torch.Tensor, nsteps: int,
            dt: float, segment: int = 50):
    """Differentiable rollout; returns (values [slots, N] float64, counts [N], grain ids [slots, N])."""
    return _Rollout.apply(params, engine_states, initial, int(nsteps), float(dt), int(segment))


def dense_fields(values: torch.Tensor, counts: torch.Tensor, ids: torch.Tensor, ngrains: int) -> torch.Tensor:
    """phi[g, c] from a slot-major state (differentiable in values)."""
    s, n = values.shape
    slot = torch.arange(s, device=values.device).view(s, 1)
    live = slot < counts.view(1, n).to(torch.int64)
    out = torch.zeros(ngrains, n, dtype=values.dtype, device=values.device)
    cell = torch.arange(n, device=values.device).view(1, n).expand(s, n)
    out.index_put_((ids[live].to(torch.int64), cell[live]), values[live], accumulate=True)
    return out


def mismatch_loss(values, counts, ids, target: torch.Tensor) -> torch.Tensor:
    """0.5 * sum over grains and cells of (phi - phi_target)^2 (target dense [ngrains, N])."""
    phi = dense_fields(values, counts, ids, target.shape[0])
    return 0.5 * ((phi - target) ** 2).sum()
