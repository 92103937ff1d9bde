"""Pins the oracle's reference gradient (oracle/gradient.c, complex-step differentiation of
the oracle's own Yoshida-4 propagation) before it is used to check the engine's adjoint.

* closed form: with every component off the map is r_T = r0 + T v0, v_T = v0;
* central finite differences of the real propagation (independent of any derivative code);
* the symplectic invariant M^T J M = J of the Jacobian of a conservative force model
  (Yoshida's composition of symplectic drifts and gradient kicks), which drag violates.
"""
import numpy as np
import pytest

import ladds_amd as la
import oracle_lib as ol

MU = 398600.4418
R_E = 6378.1363
SMOOTH = ol.KEP | ol.J2 | ol.C22 | ol.S22 | ol.SOL | ol.LUN  # no shadow, no density table rows


def _population(n, seed):
    pop = la.generate_population(1, seed, n)
    return pop, pop.r(), pop.v()


def _full_jacobian(r, v, mask, steps, dt, **kw):
    n = r.shape[0]
    M = np.zeros((n, 6, 6))
    for k in range(6):
        g = np.zeros((n, 6))
        g[:, k] = 1.0
        gr, gv, _ = ol.vjp(r, v, 0.0, dt, mask, steps, g[:, :3], g[:, 3:], **kw)
        M[:, k, :3] = gr
        M[:, k, 3:] = gv
    return M


def test_straight_line_closed_form():
    pop, r, v = _population(16, 21)
    rng = np.random.default_rng(1)
    g_r, g_v = rng.normal(size=(16, 3)), rng.normal(size=(16, 3))
    steps, dt = 100, 10.0
    gr, gv, gbc = ol.vjp(r, v, 0.0, dt, 0, steps, g_r, g_v, bc_inv=pop.bc_inv)
    T = steps * dt
    np.testing.assert_allclose(gr, g_r, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(gv, g_v + T * g_r, rtol=1e-12, atol=1e-12)
    assert np.all(gbc == 0.0)


def _fd_rv(r, v, mask, steps, dt, g_r, g_v, rel, **kw):
    """Central differences of <g, state_T> w.r.t. r0 and v0, step rel * |x| per object."""
    def f(rr, vv):
        R, V = ol.propagate(rr, vv, 0.0, dt, mask, steps, **kw)
        return (g_r * R).sum(1) + (g_v * V).sum(1)

    out = np.zeros((r.shape[0], 6))
    for k in range(6):
        base = r if k < 3 else v
        h = rel * np.linalg.norm(base, axis=1)
        xp, xm = base.copy(), base.copy()
        xp[:, k % 3] += h
        xm[:, k % 3] -= h
        if k < 3:
            out[:, k] = (f(xp, v) - f(xm, v)) / (2 * h)
        else:
            out[:, k] = (f(r, xp) - f(r, xm)) / (2 * h)
    return out


@pytest.mark.parametrize("mask", [ol.KEP, ol.KEP | ol.J2, SMOOTH])
def test_matches_central_differences(mask):
    """One day at dt = 60 s: normwise agreement per object within 1e-7 (FD truncation and
    rounding are ~1e-8 at h = 1e-6 |x|)."""
    pop, r, v = _population(32, 22)
    rng = np.random.default_rng(2)
    g_r, g_v = rng.normal(size=(32, 3)), rng.normal(size=(32, 3)) * 1e3
    gr, gv, _ = ol.vjp(r, v, 0.0, 60.0, mask, 1440, g_r, g_v, aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr)
    fd = _fd_rv(r, v, mask, 1440, 60.0, g_r, g_v, 1e-6, aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr)
    cs = np.concatenate([gr, gv], 1)
    err = np.max(np.abs(fd - cs), 1) / np.max(np.abs(cs), 1)
    assert err.max() < 1e-7, err.max()


def test_ballistic_coefficient_against_central_differences():
    """d<g, state_T>/d bc_inv on circular 250-400 km orbits under Kepler + drag, 4 h."""
    n = 32
    rng = np.random.default_rng(3)
    alt = np.linspace(250.0, 400.0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    w = np.cross(u, rng.normal(size=(n, 3)))
    w /= np.linalg.norm(w, axis=1)[:, None]
    rad = R_E + alt
    r, v = u * rad[:, None], w * np.sqrt(MU / rad)[:, None]
    bc = np.full(n, 0.05)
    g_r, g_v = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 1e3
    mask, steps = ol.KEP | ol.DRAG, 240
    _, _, gbc = ol.vjp(r, v, 0.0, 60.0, mask, steps, g_r, g_v, bc_inv=bc)

    def f(b):
        R, V = ol.propagate(r, v, 0.0, 60.0, mask, steps, bc_inv=b)
        return (g_r * R).sum(1) + (g_v * V).sum(1)

    h = 1e-3 * bc
    fd = (f(bc + h) - f(bc - h)) / (2 * h)
    assert np.max(np.abs(fd - gbc) / np.abs(gbc)) < 1e-5


@pytest.mark.parametrize("mask,symplectic", [(ol.KEP, True), (ol.KEP | ol.J2, True), (0x7F, True), (0xFF, False)])
def test_symplectic_jacobian(mask, symplectic):
    """M^T J M = J for the one-day map of every conservative model (incl. the SRP branch
    and time-dependent Sun/Moon/Earth rotation); drag breaks it."""
    pop, r, v = _population(8, 23)
    M = _full_jacobian(r, v, mask, 1440, 60.0, aom=pop.aom, bc_inv=pop.bc_inv, cr=pop.cr)
    J = np.block([[np.zeros((3, 3)), np.eye(3)], [-np.eye(3), np.zeros((3, 3))]])
    E = np.einsum("nji,jk,nkl->nil", M, J, M) - J
    if symplectic:
        assert np.max(np.abs(E)) < 1e-6
    else:
        assert np.max(np.abs(E)) > 1e-3
