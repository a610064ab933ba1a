"""Tangent-linear and adjoint of scalar transport over a window of RK3 steps.

The operator is the species part of the dycore step for one scalar:
advect_scalar followed by rk_update_scalar in each RK stage
(src/wrf_jax/dyn_em/module_advect_em_fast.py:1437-1769,
src/wrf_jax/dyn_em/module_em.py:1437-1598), plain 5th/3rd-order fluxes, with
the mass fluxes and column masses that a dycore run produced:

    q1 = U1(q0, q0),  q2 = U2(q0, q1),  q3 = U3(q0, q2)   per step,
    U_rk(q_old, q_in) = (m_old q_old + dt_rk msfty A_rk q_in) / m_new.

``Trajectory`` runs the dycore and checkpoints, per RK stage, the dynamics
each stage of the operator needs (ru_m, rv_m, ww_m, mu_1, mu_2): the adjoint
runs the stages backwards over those checkpoints.  The operator is linear in
q for given dynamics, so its tangent-linear model is the operator itself
and no q needs storing.  ``Transport.forward`` / ``adjoint`` call the HIP
kernels through the C ABI (wrfdc_transport_stage / _adjoint);
``TransportFunction`` exposes the pair to torch.autograd.
"""

from __future__ import annotations

import ctypes

from . import _lib


def _torch():
    import torch

    return torch


class Trajectory:
    """Dynamics of ``nsteps`` dycore steps of a Domain, one checkpoint per RK stage."""

    def __init__(self, domain, nsteps: int, dtbc0: float = 0.0):
        torch = _torch()
        self.domain = domain
        self.nsteps = int(nsteps)
        nst = 3 * self.nsteps
        b = domain.buffers
        self.stages = []
        ptrs = (ctypes.c_void_p * (5 * nst))()
        for s in range(nst):
            st = [torch.empty_like(b["ru_m"]), torch.empty_like(b["rv_m"]), torch.empty_like(b["ww_m"]),
                  torch.empty_like(b["mu_1"]), torch.empty_like(b["mu_2"])]
            self.stages.append(st)
            for q, t in enumerate(st):
                ptrs[5 * s + q] = t.data_ptr()
        lib = domain.lib
        domain._check(lib.wrfdc_record_stages(domain.ctx, ptrs, nst), "wrfdc_record_stages")
        try:
            domain.step(self.nsteps, dtbc0)
            domain.synchronize()
            got = lib.wrfdc_recorded_stages(domain.ctx)
        finally:
            lib.wrfdc_record_stages(domain.ctx, None, 0)
        if got != nst:
            raise RuntimeError(f"recorded {got} of {nst} stages")
        self._ptrs = [(ctypes.c_void_p * 5)(*[t.data_ptr() for t in st]) for st in self.stages]

    def stage_ptrs(self, n: int, rk: int):
        return self._ptrs[3 * n + (rk - 1)]


class Transport:
    """Forward (= tangent-linear) and adjoint of the transport over a Trajectory."""

    def __init__(self, trajectory: Trajectory):
        self.traj = trajectory
        self.dom = trajectory.domain
        self.lib = self.dom.lib

    def _check(self, rc, what):
        self.dom._check(rc, what)

    def _p(self, t):
        return ctypes.c_void_p(t.data_ptr())

    def forward(self, q0):
        torch = _torch()
        q_old = q0.contiguous()
        a, b = torch.zeros_like(q_old), torch.zeros_like(q_old)
        for n in range(self.traj.nsteps):
            out = torch.zeros_like(q_old)
            self._check(self.lib.wrfdc_transport_stage(self.dom.ctx, 1, self.traj.stage_ptrs(n, 1), self._p(q_old), self._p(q_old),
                                                       self._p(a)), "wrfdc_transport_stage")
            self._check(self.lib.wrfdc_transport_stage(self.dom.ctx, 2, self.traj.stage_ptrs(n, 2), self._p(a), self._p(q_old),
                                                       self._p(b)), "wrfdc_transport_stage")
            self._check(self.lib.wrfdc_transport_stage(self.dom.ctx, 3, self.traj.stage_ptrs(n, 3), self._p(b), self._p(q_old),
                                                       self._p(out)), "wrfdc_transport_stage")
            q_old = out
        return q_old

    # the tangent-linear model of a linear operator is the operator
    tangent_linear = forward

    def adjoint(self, qbar_end):
        torch = _torch()
        qbar = qbar_end.contiguous()
        for n in reversed(range(self.traj.nsteps)):
            acc = torch.zeros_like(qbar)
            b2 = torch.zeros_like(qbar)
            b1 = torch.zeros_like(qbar)
            self._check(self.lib.wrfdc_transport_stage_adjoint(self.dom.ctx, 3, self.traj.stage_ptrs(n, 3), self._p(qbar),
                                                               self._p(acc), self._p(b2)), "wrfdc_transport_stage_adjoint")
            self._check(self.lib.wrfdc_transport_stage_adjoint(self.dom.ctx, 2, self.traj.stage_ptrs(n, 2), self._p(b2),
                                                               self._p(acc), self._p(b1)), "wrfdc_transport_stage_adjoint")
            self._check(self.lib.wrfdc_transport_stage_adjoint(self.dom.ctx, 1, self.traj.stage_ptrs(n, 1), self._p(b1),
                                                               self._p(acc), self._p(acc)), "wrfdc_transport_stage_adjoint")
            qbar = acc
        return qbar


def _autograd_function():
    torch = _torch()

    class TransportFunction(torch.autograd.Function):
        """q_end = M q0 with M the transport over a Trajectory; the backward
        pass is the hand-written adjoint M^T."""

        @staticmethod
        def forward(ctx, q0, transport):
            ctx.transport = transport
            return transport.forward(q0.detach())

        @staticmethod
        def backward(ctx, grad_out):
            return ctx.transport.adjoint(grad_out.contiguous()), None

    return TransportFunction


_FN = None


def transport_apply(transport: Transport, q0):
    """Differentiable application of the transport (torch.autograd)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(q0, transport)


__all__ = ["Trajectory", "Transport", "transport_apply"]
