"""Adjoint of the FP64 step (mpf_adjoint_step) and the differentiable rollout (openphase_amd.adjoint).

1. Dot-product test of one step: for random directions v (state values) and u (pair tables) and a
   random cotangent w, <w, J v> and <w, J_p u> from central differences of the GPU forward step
   must equal <J^T w, v> and <grad, u> from the adjoint.
2. Rollout gradients of a microstructure-mismatch loss with respect to the pair tables (energy,
   mobility, driving force; cubic anisotropy, two phases) against central differences of the CPU
   oracle's FP64 rollout, within 1e-4 relative.
Both on a 64x64 two-phase case with cubic anisotropy (epsilon 0.05), driving force between the
phases and grains far above the reference volume (VolumeRatio 1 throughout, the regime of the
calibration runs: the adjoint holds VolumeRatios fixed).
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


class _Case:
    """64x64, 8 grains of two phases, cubic anisotropy, driving force 1e5 J/m^3 between phases."""

    def __init__(self):
        cubic = np.full((2, 2), _capi.MPF_ENERGY_CUBIC, np.int32)
        dgm = np.array([[0, 1], [1, 0]], np.int32)
        dgv = np.array([[0.0, 1e5], [-1e5, 0.0]])
        self.cfg = StepConfig(64, 64, 0, nphases=2, sigma=np.array([[0.24, 0.30], [0.30, 0.20]]),
                              mu=np.array([[4e-9, 3e-9], [3e-9, 5e-9]]), energy_model=cubic,
                              epsilon=np.full((2, 2), 0.05), dg_mode=dgm, dg_value=dgv)
        self.dt = 0.9 * self.cfg.max_time_step()
        ng = 8
        self.phase = (np.arange(ng) % 2).astype(np.int32)
        _, self.rotation = random_rotations(ng, seed=0x0FE4)
        self.init = voronoi_grains(64, 64, 0, ng, seed=0x0FE4)

    def setup(self, engine):
        engine.set_grains(self.phase, self.rotation)
        engine.upload(self.init)
        engine.finalize_initialization()
        return engine


def _setup(case):
    eng = Engine(case.cfg)
    case.setup(eng)
    assert np.all(eng.grains()[1] == 1.0)
    return eng


def _tables(cfg):
    np_ = cfg.nphases
    dg = np.zeros((np_, np_)) if cfg.dg_value is None else np.asarray(cfg.dg_value, np.float64)
    return np.stack([np.asarray(cfg.sigma, np.float64), np.asarray(cfg.mu, np.float64), dg])


def test_adjoint_step_dot_product():
    case = _Case()
    eng = _setup(case)
    eng.step(case.dt, 5)
    st = DeviceStates(eng)
    s0 = st.save()
    cnt0, _, val0 = st.views(s0)
    slot = torch.arange(st.slots, device="cuda").view(-1, 1)
    live0 = (slot < cnt0.view(1, -1).long()).double()
    g = torch.Generator(device="cuda").manual_seed(7)
    # relative perturbation of the stored values: phi (1 - phi) keeps tiny entries tiny
    v = torch.randn(val0.shape, dtype=torch.float64, device="cuda", generator=g) * live0 * val0 * (1 - val0)
    p0 = _tables(case.cfg)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(p0.shape) * np.abs(p0)   # relative directions
    eps = 1e-7

    def forward(dv=0.0, dp=0.0):
        b = s0.clone()
        st.views(b)[2].add_(dv)
        p = p0 + dp
        eng.set_pair_params(p[0], p[1], p[2])
        st.restore(b)
        eng.step(case.dt, 1)
        out = st.save()
        c, i, x = st.views(out)
        return c.clone(), i.clone(), x.clone()

    cp, ip, xp = forward(dv=eps * v)
    cm, im, xm = forward(dv=-eps * v)
    cpp, ipp, xpp = forward(dp=eps * u)
    cpm, ipm, xpm = forward(dp=-eps * u)
    c1, i1, x1 = forward()
    # cells whose kept set flips under the +-eps runs (an entry within eps of a clamp) have no
    # derivative: they get no weight in w
    live1b = slot < c1.view(1, -1).long()
    same = torch.ones_like(c1, dtype=torch.bool)
    for c, i in ((cp, ip), (cm, im), (cpp, ipp), (cpm, ipm)):
        same &= (c == c1) & ((i == i1) | ~live1b).all(dim=0)
    print("cells with a flipping kept set:", int((~same).sum()), "of", same.numel())
    assert int((~same).sum()) < 0.01 * same.numel()
    live1 = (live1b & same.view(1, -1)).double()
    w = torch.randn(x1.shape, dtype=torch.float64, device="cuda", generator=g) * live1
    jv = (xp - xm) / (2 * eps)
    jpu = (xpp - xpm) / (2 * eps)

    eng.set_pair_params(p0[0], p0[1], p0[2])
    st.restore(s0)
    lam = torch.empty_like(val0)
    grad = torch.zeros(p0.size, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    eng.adjoint_step(case.dt, w.data_ptr(), lam.data_ptr(), grad.data_ptr())
    lhs_v, rhs_v = float((w * jv).sum()), float((lam * v).sum())
    lhs_p, rhs_p = float((w * jpu).sum()), float((grad.cpu().numpy() * u.reshape(-1)).sum())
    print("state direction", lhs_v, rhs_v, "table direction", lhs_p, rhs_p)
    assert abs(lhs_v - rhs_v) <= 1e-6 * max(abs(lhs_v), abs(rhs_v))
    assert abs(lhs_p - rhs_p) <= 1e-6 * max(abs(lhs_p), abs(rhs_p))
    assert float((lam * (1 - live0)).abs().sum()) == 0.0
    eng.close()


def _oracle_loss(case, tables, nsteps, target):
    cfg = dataclasses.replace(case.cfg, sigma=tables[0], mu=tables[1],
                              dg_value=tables[2] if case.cfg.dg_mode is not None else None)
    orc = O.Oracle(cfg)
    orc.set_grains(case.phase, case.rotation)
    orc.upload(case.init)
    orc.finalize_initialization()
    orc.step(case.dt, nsteps)
    got = orc.download()
    phi = np.zeros_like(target)
    cell = np.repeat(np.arange(got.ncells), got.count)
    phi[got.index, cell] = got.value
    orc.close()
    return 0.5 * float(((phi - target) ** 2).sum())


def test_rollout_gradient_matches_oracle_differences():
    case = _Case()
    nsteps = 20
    eng = _setup(case)
    st = DeviceStates(eng)
    init = st.save()
    p0 = _tables(case.cfg)
    ng = case.phase.shape[0]
    # target: the microstructure grown with perturbed tables (symmetric energy / mobility,
    # antisymmetric driving force), as in a calibration
    rng = np.random.default_rng(0x0FE5)
    f = np.exp(0.1 * rng.standard_normal(p0.shape))
    f[0] = np.sqrt(f[0] * f[0].T)
    f[1] = np.sqrt(f[1] * f[1].T)
    f[2] = np.triu(f[2]) + np.triu(f[2], 1).T
    pt = p0 * f
    with torch.no_grad():
        vt, ct, it = rollout(st, torch.tensor(pt), init, nsteps, case.dt, segment=20)
        target = dense_fields(vt, ct, it, ng)
    params = torch.tensor(p0, requires_grad=True)
    vals, cnt, ids = rollout(st, params, init, nsteps, case.dt, segment=6)
    assert np.all(eng.grains()[1] == 1.0)
    loss = mismatch_loss(vals, cnt, ids, target)
    loss.backward()
    g = params.grad.numpy()
    tgt = target.cpu().numpy()
    # the loss of the GPU forward equals the oracle's (bit-exact step)
    l0 = _oracle_loss(case, p0, nsteps, tgt)
    assert abs(float(loss.detach()) - l0) <= 1e-12 * max(1.0, l0)
    checks = [
        ("sigma 0-1", 0, [(0, 1, 1.0), (1, 0, 1.0)]),
        ("sigma 1-1", 0, [(1, 1, 1.0)]),
        ("mu 0-1", 1, [(0, 1, 1.0), (1, 0, 1.0)]),
        ("mu 0-0", 1, [(0, 0, 1.0)]),
        ("dG 0-1", 2, [(0, 1, 1.0), (1, 0, -1.0)]),
    ]
    for name, t, entries in checks:
        scale = max(abs(p0[t][entries[0][0], entries[0][1]]), 1e-30)
        h = 1e-5 * scale
        dp = np.zeros_like(p0)
        for a, b, sgn in entries:
            dp[t][a, b] = sgn * h
        fd = (_oracle_loss(case, p0 + dp, nsteps, tgt) - _oracle_loss(case, p0 - dp, nsteps, tgt)) / (2 * h)
        fd2 = (_oracle_loss(case, p0 + 10 * dp, nsteps, tgt) - _oracle_loss(case, p0 - 10 * dp, nsteps, tgt)) / (20 * h)
        ad = sum(sgn * g[t][a, b] for a, b, sgn in entries)
        print(f"{name}: adjoint {ad:.10e} oracle central difference {fd:.10e} (10x step {fd2:.10e})")
        assert abs(ad - fd) <= 1e-4 * max(abs(fd), abs(ad)), name
    eng.close()
