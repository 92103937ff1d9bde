"""Differentiable FP64 Yoshida-4 propagation for PyTorch (ROCm).

A thin binding of the ``ladds_adjoint_*`` entry points of ``include/ladds_amd.h``,
used only for this form: the forward pass and the hand-written adjoint both run as
gfx950 kernels on torch's current HIP stream, on tensors torch owns.  There is no
autograd tape through the propagation: ``YoshidaFlow.backward`` calls the adjoint
kernel, which recomputes each checkpoint segment and sweeps it in reverse.

    plan = AdjointPlan(n=4096, nsteps=1440, dt=60.0, components=ALL_COMPONENTS)
    rT, vT = propagate(r0, v0, bc_inv, plan, aom=aom, cr=cr)   # (n, 3) float64 on cuda
    loss = ((rT - target) ** 2).sum()
    loss.backward()                                            # r0.grad, v0.grad, bc_inv.grad

No LADDS counterpart: the reference propagates forward only (YoshidaIntegrator, called at
src/ladds/Simulation.cpp:297); the forward pass here is that step in FP64.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _abi
from ._abi import ALL_COMPONENTS, lib

__all__ = ["AdjointPlan", "YoshidaFlow", "propagate"]


def _hip_runtimes() -> set:
    with open("/proc/self/maps") as f:
        return {ln.split()[-1] for ln in f if "libamdhip64" in ln}


# torch bundles its own HIP runtime; the engine library must be bound to the same one
# (import torch before ladds_amd), or the two runtimes contend for the device.
if len(_hip_runtimes()) > 1:
    raise ImportError("two HIP runtimes are loaded (torch's and ROCm's): import torch before ladds_amd "
                      "so that the engine library binds to torch's runtime")


class AdjointPlan:
    """Per-(n, nsteps, dt, t0, components) state: the sub-step environments on the device
    and the backward workspace.  One plan serves any number of forward/backward pairs;
    calls on one plan must not overlap (they share the workspace)."""

    def __init__(self, n: int, nsteps: int, dt: float, t0: float = 0.0, components: int = ALL_COMPONENTS,
                 checkpoint_every: int = 64, device: int = 0):
        cfg = _abi.AdjointConfig(n, nsteps, checkpoint_every, dt, t0, components, device)
        h = C.c_void_p()
        rc = lib.ladds_adjoint_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"ladds_adjoint_create failed with status {rc}")
        self._h = h
        self.n, self.nsteps, self.dt, self.t0 = n, nsteps, dt, t0
        self.components, self.device = components, device
        self.checkpoint_doubles = int(lib.ladds_adjoint_checkpoint_doubles(h))

    def close(self):
        if self._h:
            lib.ladds_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = lib.ladds_adjoint_last_error(self._h)
            raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _rows(x: torch.Tensor, plan: AdjointPlan, name: str) -> torch.Tensor:
    """(n, 3) float64 on the plan's device -> contiguous [3][n] block."""
    if x.dtype != torch.float64 or x.shape != (plan.n, 3):
        raise ValueError(f"{name} must be a ({plan.n}, 3) float64 tensor, got {tuple(x.shape)} {x.dtype}")
    if x.device.type != "cuda" or x.device.index != plan.device:
        raise ValueError(f"{name} must live on cuda:{plan.device}")
    return x.detach().t().contiguous()


def _column(x, plan: AdjointPlan, name: str):
    if x is None:
        return None
    if x.dtype != torch.float64 or x.shape != (plan.n,) or x.device.type != "cuda" or x.device.index != plan.device:
        raise ValueError(f"{name} must be a ({plan.n},) float64 tensor on cuda:{plan.device}")
    return x.detach().contiguous()


class YoshidaFlow(torch.autograd.Function):
    """(r0, v0, bc_inv) -> (r_T, v_T) after plan.nsteps FP64 Yoshida-4 steps."""

    @staticmethod
    def forward(ctx, r0, v0, bc_inv, aom, cr, plan: AdjointPlan):
        R0, V0 = _rows(r0, plan, "r0"), _rows(v0, plan, "v0")
        bc, am, c = _column(bc_inv, plan, "bc_inv"), _column(aom, plan, "aom"), _column(cr, plan, "cr")
        RT, VT = torch.empty_like(R0), torch.empty_like(V0)
        ck = torch.empty(plan.checkpoint_doubles, dtype=torch.float64, device=R0.device)
        stream = torch.cuda.current_stream(R0.device).cuda_stream
        plan.check(lib.ladds_adjoint_forward(plan._h, _ptr(R0), _ptr(V0), _ptr(am), _ptr(bc), _ptr(c), None,
                                             _ptr(RT), _ptr(VT), _ptr(ck), C.c_void_p(stream)),
                   "ladds_adjoint_forward")
        ctx.plan, ctx.ck, ctx.params = plan, ck, (am, bc, c)
        ctx.has_bc = bc_inv is not None
        return RT.t(), VT.t()

    @staticmethod
    def backward(ctx, g_rT, g_vT):
        plan, ck = ctx.plan, ctx.ck
        am, bc, c = ctx.params
        dev = ck.device
        GR = None if g_rT is None else g_rT.detach().t().contiguous()
        GV = None if g_vT is None else g_vT.detach().t().contiguous()
        g_r0 = torch.empty((3, plan.n), dtype=torch.float64, device=dev)
        g_v0 = torch.empty((3, plan.n), dtype=torch.float64, device=dev)
        g_bc = torch.empty(plan.n, dtype=torch.float64, device=dev) if ctx.needs_input_grad[2] else None
        stream = torch.cuda.current_stream(dev).cuda_stream
        plan.check(lib.ladds_adjoint_backward(plan._h, _ptr(am), _ptr(bc), _ptr(c), None, _ptr(ck), _ptr(GR),
                                              _ptr(GV), _ptr(g_r0), _ptr(g_v0), _ptr(g_bc), C.c_void_p(stream)),
                   "ladds_adjoint_backward")
        return g_r0.t(), g_v0.t(), g_bc, None, None, None


def propagate(r0, v0, bc_inv, plan: AdjointPlan, aom=None, cr=None):
    """Differentiable FP64 propagation; gradients flow to r0, v0 and bc_inv."""
    return YoshidaFlow.apply(r0, v0, bc_inv, aom, cr, plan)
