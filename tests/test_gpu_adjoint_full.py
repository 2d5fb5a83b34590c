"""The adjoint at the calibration configuration: 512x512 periodic grid, 200 grains with random
uniform cubic orientations and cubic anisotropy (epsilon 0.05), two phase types with a constant
driving force of 1e5 J/m^3 between them, a 500-step rollout with a checkpoint every 50 steps.

1. Gradients of the L2 mismatch to a target grown with the pair tables perturbed by +-10%
   (lognormal) against central differences of the CPU oracle's FP64 rollout, for five table
   entries, within 1e-5 relative, over 20 steps with a checkpoint every 5. Over longer rollouts
   the loss is piecewise smooth at the scale of a difference step: entries that appear or vanish,
   values clamped at 0 or 1 and limiter switches make central differences with relative steps
   1e-6 and 1e-7 disagree with each other (by up to a factor 2.6 at 200 steps), so they pin
   nothing there; the 500-step gradient is covered by 2-4.
2. The dot-product test of one adjoint step at this size (state and table directions).
3. The 500-step gradient does not depend on where the checkpoints are (bit for bit).
4. The same gradient from an FP32 rollout: its deviation from the FP64 gradient is reported and
   bounded; gradients are identical from run to run.
"""
import dataclasses

import numpy as np
import pytest
import torch

import oracle as O
from openphase_amd import _capi
from openphase_amd.adjoint import DeviceStates, dense_fields, mismatch_loss, rollout
from openphase_amd.engine import Engine, StepConfig
from openphase_amd.microstructure import random_rotations, voronoi_grains

pytestmark = pytest.mark.gpu

N, NG, STEPS, SEGMENT = 512, 200, 500, 50
FD_STEPS, FD_SEGMENT = 20, 5


def _cfg(precision=_capi.MPF_FP64):
    return StepConfig(N, N, 0, nphases=2, sigma=np.full((2, 2), 0.24), mu=np.full((2, 2), 4e-9),
                      energy_model=np.full((2, 2), _capi.MPF_ENERGY_CUBIC, np.int32), epsilon=np.full((2, 2), 0.05),
                      dg_mode=np.array([[0, 1], [1, 0]], np.int32), dg_value=np.array([[0.0, 1e5], [-1e5, 0.0]]),
                      precision=precision)


PHASE = (np.arange(NG) % 2).astype(np.int32)
_, ROT = random_rotations(NG, 0x0FE4)
INIT = voronoi_grains(N, N, 0, NG, seed=0x0FE4)
DT = 0.9 * _cfg().max_time_step()


def _tables(cfg):
    return np.stack([np.asarray(cfg.sigma, np.float64), np.asarray(cfg.mu, np.float64), np.asarray(cfg.dg_value)])


def _engine(precision):
    eng = Engine(_cfg(precision))
    eng.set_grains(PHASE, ROT)
    eng.upload(INIT)
    eng.finalize_initialization()
    return eng


def _target(p0):
    """Microstructure grown with tables perturbed by exp(0.1 N(0,1)) (energy and mobility
    symmetric, driving force antisymmetric), as a calibration's measurement."""
    rng = np.random.default_rng(0x0FE5)
    f = np.exp(0.1 * rng.standard_normal(p0.shape))
    f[0] = np.sqrt(f[0] * f[0].T)
    f[1] = np.sqrt(f[1] * f[1].T)
    f[2] = np.triu(f[2]) + np.triu(f[2], 1).T
    return p0 * f


def _gradient(precision, p0, pt, steps=STEPS, segment=SEGMENT):
    eng = _engine(precision)
    st = DeviceStates(eng)
    init = st.save()
    with torch.no_grad():
        vt, ct, it = rollout(st, torch.tensor(pt), init, steps, DT, segment=segment)
        target = dense_fields(vt, ct, it, NG)
    params = torch.tensor(p0, requires_grad=True)
    vals, cnt, ids = rollout(st, params, init, steps, DT, segment=segment)
    loss = mismatch_loss(vals, cnt, ids, target)
    loss.backward()
    out = params.grad.numpy().copy(), float(loss.detach()), target.cpu().numpy()
    eng.close()
    return out


def _oracle_loss(tables, target, steps):
    cfg = dataclasses.replace(_cfg(), sigma=tables[0], mu=tables[1], dg_value=tables[2])
    orc = O.Oracle(cfg)
    orc.set_grains(PHASE, ROT)
    orc.upload(INIT)
    orc.finalize_initialization()
    orc.step(DT, steps)
    got = orc.download()
    orc.close()
    phi = np.zeros_like(target)
    cell = np.repeat(np.arange(got.ncells), got.count)
    phi[got.index, cell] = got.value
    return 0.5 * float(((phi - target) ** 2).sum())


def test_rollout_gradient_matches_oracle_differences_512():
    p0 = _tables(_cfg())
    pt = _target(p0)
    print(f"adjoint {N}^2 {NG} grains {FD_STEPS} steps segment {FD_SEGMENT}", flush=True)
    g, loss, target = _gradient(_capi.MPF_FP64, p0, pt, FD_STEPS, FD_SEGMENT)
    l0 = _oracle_loss(p0, target, FD_STEPS)
    print(f"loss GPU {loss:.12e} oracle {l0:.12e}", flush=True)
    assert abs(loss - l0) <= 1e-9 * l0
    checks = [
        ("sigma 0-1", 0, [(0, 1, 1.0), (1, 0, 1.0)]),
        ("sigma 0-0", 0, [(0, 0, 1.0)]),
        ("mu 0-0", 1, [(0, 0, 1.0)]),
        ("mu 0-1", 1, [(0, 1, 1.0), (1, 0, 1.0)]),
        ("dG 0-1", 2, [(0, 1, 1.0), (1, 0, -1.0)]),
    ]
    for name, t, entries in checks:
        h = 1e-6 * abs(p0[t][entries[0][0], entries[0][1]])
        dp = np.zeros_like(p0)
        for a, b, sgn in entries:
            dp[t][a, b] = sgn * h
        fd = (_oracle_loss(p0 + dp, target, FD_STEPS) - _oracle_loss(p0 - dp, target, FD_STEPS)) / (2 * h)
        ad = sum(sgn * g[t][a, b] for a, b, sgn in entries)
        print(f"{name}: adjoint {ad:.10e} oracle central difference {fd:.10e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), abs(ad)), name


def test_rollout_gradient_independent_of_checkpoints_512():
    """Each segment is replayed from its checkpoint before its reverse sweep; the replays are
    the forward steps again, so the gradient is the same bit for bit whatever the segment."""
    p0 = _tables(_cfg())
    pt = _target(p0)
    g50, l50, _ = _gradient(_capi.MPF_FP64, p0, pt, STEPS, 50)
    g125, l125, _ = _gradient(_capi.MPF_FP64, p0, pt, STEPS, 125)
    print("500-step gradient:", np.array2string(g50.reshape(-1), precision=6), flush=True)
    assert l50 == l125
    assert np.array_equal(g50, g125)


def test_adjoint_step_dot_product_512():
    """<w, J v> and <w, J_p u> by central differences of the GPU step against <J^T w, v> and
    <grad, u> from one adjoint step, after 100 steps of the rollout."""
    eng = _engine(_capi.MPF_FP64)
    eng.step(DT, 100)
    st = DeviceStates(eng)
    s0 = st.save()
    cnt0, _, val0 = st.views(s0)
    slot = torch.arange(st.slots, device="cuda").view(-1, 1)
    live0 = (slot < cnt0.view(1, -1).long()).double()
    g = torch.Generator(device="cuda").manual_seed(11)
    v = torch.randn(val0.shape, dtype=torch.float64, device="cuda", generator=g) * live0 * val0 * (1 - val0)
    p0 = _tables(_cfg())
    u = np.random.default_rng(5).standard_normal(p0.shape) * np.abs(p0)
    eps = 1e-7

    def forward(dv=0.0, dp=0.0):
        b = s0.clone()
        st.views(b)[2].add_(dv)
        p = p0 + dp
        eng.set_pair_params(p[0], p[1], p[2])
        st.restore(b)
        eng.step(DT, 1)
        c, i, x = st.views(st.save())
        return c.clone(), i.clone(), x.clone()

    runs = [forward(dv=eps * v), forward(dv=-eps * v), forward(dp=eps * u), forward(dp=-eps * u)]
    c1, i1, x1 = forward()
    s1 = st.save()
    live1b = slot < c1.view(1, -1).long()
    same = torch.ones_like(c1, dtype=torch.bool)
    for c, i, _ in runs:
        same &= (c == c1) & ((i == i1) | ~live1b).all(dim=0)
    print("cells with a flipping kept set:", int((~same).sum()), "of", same.numel(), flush=True)
    w = torch.randn(x1.shape, dtype=torch.float64, device="cuda", generator=g) * (live1b & same.view(1, -1)).double()
    jv = (runs[0][2] - runs[1][2]) / (2 * eps)
    jpu = (runs[2][2] - runs[3][2]) / (2 * eps)
    eng.set_pair_params(p0[0], p0[1], p0[2])
    st.restore(s0)
    lam = torch.empty_like(val0)
    grad = torch.zeros(p0.size, dtype=torch.float64, device="cuda")
    nc, ni = st.list_pointers(s1)
    torch.cuda.synchronize()
    eng.adjoint_step(DT, w.data_ptr(), lam.data_ptr(), grad.data_ptr(), nc, ni)
    lhs_v, rhs_v = float((w * jv).sum()), float((lam * v).sum())
    lhs_p, rhs_p = float((w * jpu).sum()), float((grad.cpu().numpy() * u.reshape(-1)).sum())
    print("state direction", lhs_v, rhs_v, "table direction", lhs_p, rhs_p, flush=True)
    assert abs(lhs_v - rhs_v) <= 1e-6 * max(abs(lhs_v), abs(rhs_v))
    assert abs(lhs_p - rhs_p) <= 1e-6 * max(abs(lhs_p), abs(rhs_p))
    eng.close()


def test_fp32_rollout_gradient_512():
    """The FP32 rollout's gradient (FP32 states, adjoint in FP64 at those states) against the FP64
    rollout's, and identical from run to run."""
    p0 = _tables(_cfg())
    pt = _target(p0)
    g64, _, _ = _gradient(_capi.MPF_FP64, p0, pt)
    g32, l32, _ = _gradient(_capi.MPF_FP32, p0, pt)
    g32b, _, _ = _gradient(_capi.MPF_FP32, p0, pt)
    assert np.array_equal(g32, g32b)
    used = np.abs(g64) > 0
    rel = np.abs(g32 - g64)[used] / np.abs(g64)[used]
    print("FP32 vs FP64 rollout gradient, relative deviation per table entry:", np.array2string(rel, precision=3),
          flush=True)
    assert rel.max() < 0.05
