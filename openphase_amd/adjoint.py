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

from concurrent.futures import ThreadPoolExecutor

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
        # torch's stream (earlier users of the memory) before the engine's stream writes it; the
        # engine's copy has finished when save_state returns
        torch.cuda.current_stream().synchronize()
        self.engine.save_state(blob.data_ptr())
        return blob

    def restore(self, blob: torch.Tensor) -> None:
        torch.cuda.current_stream().synchronize()
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
        g, ctx.lam0 = _reverse(ctx.states, ctx.params, ctx.checkpoints, ctx.final, ctx.nsteps, ctx.dt, ctx.segment,
                               g_val)
        return g, None, None, None, None, None


def _reverse(states: DeviceStates, p: np.ndarray, checkpoints, final, nsteps: int, dt: float, segment: int, g_val):
    """Reverse sweep of one rollout of the engine behind `states` with tables p, from the
    cotangent g_val of its final values: each segment's states are recomputed from its checkpoint,
    then the adjoint steps run through it backwards. Returns (table cotangents [3, P, P] on the
    CPU, cotangent of the initial values). The recomputation and the adjoint steps run with the
    rollout's tables, whatever the engine holds now (another rollout may have run since); the
    engine's tables come back afterwards."""
    eng = states.engine
    held = {k: (None if v is None else v.copy()) for k, v in eng.pair_params.items()}
    eng.set_pair_params(p[0], p[1], p[2] if eng.cfg.dg_mode is not None else None)
    try:
        nph = eng.cfg.nphases
        grad = torch.zeros(3 * nph * nph, dtype=torch.float64, device="cuda")
        lam = torch.zeros(states.slots, states.ncells, dtype=torch.float64, device="cuda")
        if g_val is not None:
            lam.copy_(g_val)
        # the engine's stream reads lam, prev and grad from here on; every engine call below returns
        # after its own stream finished
        torch.cuda.current_stream().synchronize()
        nseg = len(checkpoints)
        prev = torch.empty_like(lam)
        for si in reversed(range(nseg)):
            t0 = si * segment
            n = min(segment, nsteps - t0)
            states.restore(checkpoints[si])
            seg_states = [checkpoints[si]]
            for _ in range(n - 1):
                eng.step(dt, 1)
                seg_states.append(states.save())
            seg_states.append(checkpoints[si + 1] if si + 1 < nseg else final)
            for i in reversed(range(n)):
                states.restore(seg_states[i])
                nc, ni = states.list_pointers(seg_states[i + 1])
                eng.adjoint_step(dt, lam.data_ptr(), prev.data_ptr(), grad.data_ptr(), nc, ni)
                lam, prev = prev, lam
        return grad.view(3, nph, nph).cpu(), lam
    finally:
        eng.set_pair_params(held["sigma"], held["mu"], held["dg_value"])


_REVERSE_THREADS = 8  # host threads driving the members' reverse sweeps of an ensemble rollout


class _EnsembleRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, ensemble, states, initials, nsteps, dts, segment):
        p = params.detach().to("cpu", torch.float64).numpy()
        for m, st in enumerate(states):
            eng = st.engine
            eng.set_pair_params(p[m][0], p[m][1], p[m][2] if eng.cfg.dg_mode is not None else None)
            st.restore(initials[m])
        checkpoints = [[] for _ in states]
        done = 0
        while done < nsteps:
            for m, st in enumerate(states):
                checkpoints[m].append(st.save())
            n = min(segment, nsteps - done)
            ensemble.step(list(dts), n)
            done += n
        finals = [st.save() for st in states]
        outs = []
        for st, fin in zip(states, finals):
            cnt, ids, val = st.views(fin)
            outs += [val.clone(), cnt.clone(), ids.clone()]
        ctx.states, ctx.checkpoints, ctx.finals = states, checkpoints, finals
        ctx.params = p.copy()
        ctx.nsteps, ctx.dts, ctx.segment = nsteps, dts, segment
        ctx.mark_non_differentiable(*outs[1::3], *outs[2::3])
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        # the members' reverse sweeps are independent and run on their engines' own streams: a
        # few host threads keep several in flight on the GPU (the library calls release the GIL)
        g = torch.zeros(ctx.params.shape, dtype=torch.float64)

        def one(m):
            gm, _ = _reverse(ctx.states[m], ctx.params[m], ctx.checkpoints[m], ctx.finals[m], ctx.nsteps, ctx.dts[m],
                             ctx.segment, grads[3 * m])
            return m, gm

        torch.cuda.current_stream().synchronize()
        with ThreadPoolExecutor(max_workers=min(_REVERSE_THREADS, len(ctx.states))) as pool:
            for m, gm in pool.map(one, range(len(ctx.states))):
                g[m] = gm
        return g, None, None, None, None, None, None


def ensemble_rollout(ensemble, states, params: torch.Tensor, initials, nsteps: int, dts, segment: int = 50):
    """Differentiable rollout of every member of an Ensemble (engine.Ensemble): the forward pass
    advances all members together (one captured graph per two steps of all members) and keeps a
    checkpoint of every member each `segment` steps; the backward pass runs each member's reverse
    sweep as `rollout` does. states: one DeviceStates per member, in the ensemble's order;
    params: float64 [members, 3, P, P] (the members share the phase count P); initials: a saved
    state per member; dts: a time step per member. Returns [(values, counts, grain ids)] per
    member."""
    engines = [st.engine for st in states]
    if len(engines) != len(ensemble.engines) or any(x is not y for x, y in zip(engines, ensemble.engines)):
        raise ValueError("states must hold the ensemble's members, in the ensemble's order")
    nph = {e.cfg.nphases for e in engines}
    if len(nph) != 1 or tuple(params.shape) != (len(states), 3, nph.pop(), engines[0].cfg.nphases):
        raise ValueError("params must be [members, 3, P, P] with the phase count P the members share")
    if len(initials) != len(states) or len(dts) != len(states):
        raise ValueError("one initial state and one time step per member")
    outs = _EnsembleRollout.apply(params, ensemble, tuple(states), tuple(initials), int(nsteps),
                                  tuple(float(d) for d in dts), int(segment))
    return [tuple(outs[3 * m:3 * m + 3]) for m in range(len(states))]


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
