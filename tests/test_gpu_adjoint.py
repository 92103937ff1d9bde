"""Hand-written adjoint of the FP64 Yoshida-4 step (ladds_amd/csrc/adjoint.hip) through
torch.autograd, checked against the oracle's complex-step gradient, central finite
differences of the GPU forward pass, the straight-line closed form and the symplectic
invariant of the Jacobian.
"""
import numpy as np
import pytest
import torch

import ladds_amd as la
import oracle_lib as ol

pytestmark = pytest.mark.gpu

SMOOTH = la.KEP | la.J2 | la.C22 | la.S22 | la.SOL | la.LUN
DEV = "cuda:0"


def _adj():
    from ladds_amd import adjoint
    return adjoint


def _tensors(pop, requires_grad=True):
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)
    r0, v0, bc = t(pop.r()), t(pop.v()), t(pop.bc_inv)
    for x in (r0, v0, bc):
        x.requires_grad_(requires_grad)
    return r0, v0, bc, t(pop.aom), t(pop.cr)


def _gpu_vjp(pop, mask, steps, dt, g_r, g_v, every=64):
    adj = _adj()
    plan = adj.AdjointPlan(pop.n, steps, dt, components=mask, checkpoint_every=every)
    r0, v0, bc, aom, cr = _tensors(pop)
    rT, vT = adj.propagate(r0, v0, bc, plan, aom=aom, cr=cr)
    gr = torch.tensor(g_r, dtype=torch.float64, device=DEV)
    gv = torch.tensor(g_v, dtype=torch.float64, device=DEV)
    ((rT * gr).sum() + (vT * gv).sum()).backward()
    torch.cuda.synchronize()
    out = (r0.grad.cpu().numpy(), v0.grad.cpu().numpy(), bc.grad.cpu().numpy(),
           rT.detach().cpu().numpy(), vT.detach().cpu().numpy())
    plan.close()
    return out


def _normwise(a, b):
    return np.max(np.abs(a - b), 1) / np.max(np.abs(b), 1)


def test_forward_matches_engine_and_oracle():
    pop = la.generate_population(1, 0x1ADD5 + 4, 1024)
    mask, steps, dt = la.ALL_COMPONENTS, 100, 60.0
    *_, rT, vT = _gpu_vjp(pop, mask, steps, dt, np.zeros((pop.n, 3)), np.zeros((pop.n, 3)))
    with la.Engine(la.EngineConfig(dt=dt, components=mask, cutoff_km=10.0)) as eng:
        eng.set_objects(pop)
        eng.propagate(steps)
        r_eng, _ = eng.state()
    assert np.max(np.linalg.norm(rT - r_eng, axis=1) / np.linalg.norm(r_eng, axis=1)) < 1e-12
    ro, _ = ol.propagate(pop.r(), pop.v(), 0.0, dt, mask, steps, aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr)
    assert np.max(np.linalg.norm(rT - ro, axis=1) / np.linalg.norm(ro, axis=1)) < 1e-9 * steps


@pytest.mark.parametrize("mask", [la.KEP, la.KEP | la.J2, SMOOTH, la.ALL_COMPONENTS])
def test_gradient_matches_complex_step_oracle(mask):
    """4,096-object distributions at 256 objects, one day at dt = 60 s, checkpoints every 64
    steps: every object's (g_r0, g_v0) within 1e-8 normwise of the complex-step gradient,
    g_bc within 1e-6 relative where drag acts."""
    pop = la.generate_population(1, 0x1ADD5 + 5, 256)
    rng = np.random.default_rng(5)
    g_r, g_v = rng.normal(size=(pop.n, 3)), rng.normal(size=(pop.n, 3)) * 1e3
    gr, gv, gbc, _, _ = _gpu_vjp(pop, mask, 1440, 60.0, g_r, g_v)
    cr_, cv_, cbc = ol.vjp(pop.r(), pop.v(), 0.0, 60.0, mask, 1440, g_r, g_v, aom=pop.aom, bc_inv=pop.bc_inv,
                           cr=pop.cr)
    err = _normwise(np.concatenate([gr, gv], 1), np.concatenate([cr_, cv_], 1))
    assert err.max() < 1e-8, (err.max(), int(np.argmax(err)))
    if mask & la.DRAG:
        scale = np.max(np.abs(cbc))
        assert np.max(np.abs(gbc - cbc)) < 1e-6 * scale
    else:
        assert np.all(gbc == 0.0)


def test_straight_line_closed_form():
    pop = la.generate_population(1, 0x1ADD5 + 6, 128)
    rng = np.random.default_rng(6)
    g_r, g_v = rng.normal(size=(pop.n, 3)), rng.normal(size=(pop.n, 3))
    gr, gv, gbc, rT, _ = _gpu_vjp(pop, 0, 100, 10.0, g_r, g_v, every=7)
    np.testing.assert_allclose(gr, g_r, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(gv, g_v + 1000.0 * g_r, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(rT, pop.r() + 1000.0 * pop.v(), rtol=1e-12)


@pytest.mark.parametrize("every", [1, 7, 64, 1440])
def test_checkpoint_interval_does_not_change_the_gradient(every):
    """Recomputation from any checkpoint spacing reproduces the same trajectory bit for bit."""
    pop = la.generate_population(1, 0x1ADD5 + 7, 192)
    rng = np.random.default_rng(7)
    g_r, g_v = rng.normal(size=(pop.n, 3)), rng.normal(size=(pop.n, 3))
    ref = _gpu_vjp(pop, la.ALL_COMPONENTS, 300, 60.0, g_r, g_v, every=64)
    got = _gpu_vjp(pop, la.ALL_COMPONENTS, 300, 60.0, g_r, g_v, every=every)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_gradient_matches_central_differences_of_gpu_forward():
    """h = 1e-6 |x| central differences of <g, state_T> through the GPU forward pass."""
    adj = _adj()
    pop = la.generate_population(1, 0x1ADD5 + 8, 128)
    mask, steps, dt = SMOOTH, 1440, 60.0
    rng = np.random.default_rng(8)
    g_r, g_v = rng.normal(size=(pop.n, 3)), rng.normal(size=(pop.n, 3)) * 1e3
    gr, gv, _, _, _ = _gpu_vjp(pop, mask, steps, dt, g_r, g_v)
    plan = adj.AdjointPlan(pop.n, steps, dt, components=mask)

    def f(r, v):
        rt = torch.tensor(r, dtype=torch.float64, device=DEV)
        vt = torch.tensor(v, dtype=torch.float64, device=DEV)
        with torch.no_grad():
            R, V = adj.propagate(rt, vt, None, plan)
        return (g_r * R.cpu().numpy()).sum(1) + (g_v * V.cpu().numpy()).sum(1)

    r, v = pop.r(), pop.v()
    fd = np.zeros((pop.n, 6))
    for k in range(6):
        base = r if k < 3 else v
        h = 1e-6 * np.linalg.norm(base, axis=1)
        xp, xm = base.copy(), base.copy()
        xp[:, k % 3] += h
        xm[:, k % 3] -= h
        fd[:, k] = ((f(xp, v) - f(xm, v)) if k < 3 else (f(r, xp) - f(r, xm))) / (2 * h)
    plan.close()
    err = _normwise(fd, np.concatenate([gr, gv], 1))
    assert err.max() < 1e-7, err.max()


def test_symplectic_jacobian_of_adjoint():
    """M^T J M = J for the one-day map under every conservative component."""
    pop = la.generate_population(1, 0x1ADD5 + 9, 16)
    M = np.zeros((pop.n, 6, 6))
    for k in range(6):
        g = np.zeros((pop.n, 6))
        g[:, k] = 1.0
        gr, gv, _, _, _ = _gpu_vjp(pop, 0x7F, 1440, 60.0, g[:, :3], g[:, 3:])
        M[:, k, :3], M[:, k, 3:] = gr, gv
    J = np.block([[np.zeros((3, 3)), np.eye(3)], [-np.eye(3), np.zeros((3, 3))]])
    assert np.max(np.abs(np.einsum("nji,jk,nkl->nil", M, J, M) - J)) < 1e-6


def test_target_loss_gradient():
    """The differentiable-propagation workload at reduced size: loss = sum |r_T - target|^2
    with targets = nominal final state + N(0, 100 m); autograd gradient equals the
    oracle's complex-step VJP with g_r = 2 (r_T - target)."""
    adj = _adj()
    pop = la.generate_population(0, 0x1ADD5 + 4, 128)
    pop.bc_inv = np.random.default_rng(9).uniform(0.005, 0.05, pop.n)
    pop.cr = np.random.default_rng(10).uniform(1.0, 1.5, pop.n)
    steps, dt = 1440, 60.0
    ro, _ = ol.propagate(pop.r(), pop.v(), 0.0, dt, la.ALL_COMPONENTS, steps, aom=pop.aom, bc_inv=pop.bc_inv,
                         cr=pop.cr)
    target = ro + np.random.default_rng(11).normal(scale=0.1, size=ro.shape)
    plan = adj.AdjointPlan(pop.n, steps, dt)
    r0, v0, bc, aom, cr = _tensors(pop)
    rT, _ = adj.propagate(r0, v0, bc, plan, aom=aom, cr=cr)
    loss = ((rT - torch.tensor(target, dtype=torch.float64, device=DEV)) ** 2).sum()
    loss.backward()
    g_r = 2.0 * (rT.detach().cpu().numpy() - target)
    cr_, cv_, cbc = ol.vjp(pop.r(), pop.v(), 0.0, dt, la.ALL_COMPONENTS, steps, g_r, np.zeros_like(g_r),
                           aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr)
    got = np.concatenate([r0.grad.cpu().numpy(), v0.grad.cpu().numpy()], 1)
    assert _normwise(got, np.concatenate([cr_, cv_], 1)).max() < 1e-8
    assert np.max(np.abs(bc.grad.cpu().numpy() - cbc)) < 1e-6 * np.max(np.abs(cbc))
    plan.close()


def test_shape_errors_are_raised_before_launch():
    adj = _adj()
    plan = adj.AdjointPlan(8, 4, 10.0)
    r = torch.zeros((7, 3), dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        adj.propagate(r, r, None, plan)
    with pytest.raises(ValueError):
        adj.propagate(r.float(), r.float(), None, plan)
    plan.close()
